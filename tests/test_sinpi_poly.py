"""The sinc^2 cell's FP32 arithmetic (CPU).

Reads the coefficients the kernel ships (``kSinPiOverPi`` in csrc/score.hip)
and evaluates the kernel's cell in FP32: x = r P(r^2) / u with r = u - rint(u),
then fminf(x, 1) and x^2.  The reciprocal is emulated correctly rounded (the
hardware's v_rcp_f32 adds at most one more ulp).  The sine polynomial must
match sin(pi r)/pi to the FP32 rounding level, including near r = 0 where the
line centres sit, and the whole cell must match FP64 [sin(pi u)/(pi u)]^2 to
below 6e-7 relative, with u == 0 giving the limit 1."""

import re
from pathlib import Path

import numpy as np

SRC = Path(__file__).resolve().parents[1] / "pyrite_amd" / "csrc" / "score.hip"


def _coefficients():
    text = SRC.read_text()
    m = re.search(r"kSinPiOverPi\[5\]\s*=\s*\{([^}]*)\}", text)
    assert m, "kSinPiOverPi not found"
    return [np.float32(float(v.strip().rstrip("f"))) for v in m.group(1).split(",")]


def _poly(r, c):
    r = r.astype(np.float32)
    r2 = (r * r).astype(np.float32)
    s = np.full_like(r, c[4])
    for k in (3, 2, 1, 0):  # fmaf: one rounding per step
        s = (s.astype(np.float64) * r2 + np.float64(c[k])).astype(np.float32)
    return s


def _cell(u64, c):
    r = (u64 - np.rint(u64)).astype(np.float32)
    uf = u64.astype(np.float32)
    s = _poly(r, c)
    with np.errstate(divide="ignore", invalid="ignore"):
        rc = (np.float32(1.0) / uf).astype(np.float32)
        x = ((r * rc).astype(np.float32) * s).astype(np.float32)
    x = np.fmin(x, np.float32(1.0))  # fminf: NaN -> 1
    return (x * x).astype(np.float32)


def test_polynomial_matches_sin_pi_r_over_pi_to_fp32_rounding():
    c = _coefficients()
    assert c[0] == np.float32(1.0)  # the leading term is r
    r = np.linspace(-0.5, 0.5, 1_000_001).astype(np.float32)
    ref = np.sin(np.pi * r.astype(np.float64)) / np.pi
    got = (_poly(r, c) * r).astype(np.float64)
    live = ref != 0
    rel = np.abs(got[live] - ref[live]) / np.abs(ref[live])
    assert rel.max() < 3e-7, rel.max()
    tiny = np.float32(10.0) ** -np.arange(1, 12, dtype=np.float32)
    for x in tiny:  # near the line centre the sine is relatively exact
        val = float(_poly(np.array([x]), c)[0]) * float(x)
        assert abs(val / (np.sin(np.pi * float(x)) / np.pi) - 1.0) < 3e-7
    # a plausible mistake (a dropped r^8 term) fails
    bad = list(c)
    bad[4] = np.float32(0.0)
    assert (np.abs((_poly(r, bad) * r).astype(np.float64)[live] - ref[live]) / np.abs(ref[live])).max() > 1e-5


def test_cell_matches_fp64_profile():
    c = _coefficients()
    rng = np.random.default_rng(7)
    u = np.concatenate([rng.uniform(-0.5, 0.5, 400_000), rng.uniform(-60, 60, 400_000),
                        rng.uniform(-5e4, 5e4, 200_000), np.array([0.5, -0.5, 1.0, -3.0, 1e-30, -1e-12])])
    exact = (np.sin(np.pi * (u - np.rint(u))) / (np.pi * u)) ** 2
    got = _cell(u, c).astype(np.float64)
    live = exact > 0
    rel = np.abs(got[live] - exact[live]) / exact[live]
    assert rel.max() < 6e-7, rel.max()
    assert np.all(got <= 1.0)
    # u == 0 is the profile's limit 1 (0 * inf = NaN, which fminf maps to 1)
    assert _cell(np.array([0.0, -0.0]), c).tolist() == [1.0, 1.0]
    # integer u > 0 is an exact zero of the profile
    assert _cell(np.array([1.0, 2.0, -7.0]), c).tolist() == [0.0, 0.0, 0.0]
