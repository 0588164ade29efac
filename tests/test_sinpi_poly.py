"""The sinc^2 cell's FP32 sine polynomial (CPU).

Reads the coefficients the kernel ships (``kSinPi`` in csrc/score.hip) and
evaluates the kernel's Horner scheme in FP32 over [-1/2, 1/2]: its relative
error against FP64 sin(pi r) must stay at the FP32 rounding level, including
near r = 0 where the line centres sit (the profile there is (sin(pi r)/(pi u))^2
with u ~ r, so a relative error of the sine is the profile's)."""

import re
from pathlib import Path

import numpy as np

SRC = Path(__file__).resolve().parents[1] / "pyrite_amd" / "csrc" / "score.hip"


def _coefficients():
    text = SRC.read_text()
    m = re.search(r"kSinPi\[5\]\s*=\s*\{([^}]*)\}", text)
    assert m, "kSinPi not found"
    return [np.float32(float(v.strip().rstrip("f"))) for v in m.group(1).split(",")]


def _horner(r, c):
    r = r.astype(np.float32)
    r2 = (r * r).astype(np.float32)
    s = c[4]
    for k in (3, 2, 1, 0):  # fmaf: one rounding per step
        s = (s.astype(np.float64) * r2 + np.float64(c[k])).astype(np.float32)
    return (s * r).astype(np.float32)


def test_polynomial_matches_sin_pi_r_to_fp32_rounding():
    c = _coefficients()
    assert abs(float(c[0]) - np.pi) < 1e-6  # the leading term is pi r
    r = np.linspace(-0.5, 0.5, 1_000_001).astype(np.float32)
    ref = np.sin(np.pi * r.astype(np.float64))
    got = _horner(r, c).astype(np.float64)
    live = ref != 0
    rel = np.abs(got[live] - ref[live]) / np.abs(ref[live])
    assert rel.max() < 3e-7, rel.max()
    tiny = np.float32(10.0) ** -np.arange(1, 12, dtype=np.float32)
    for x in tiny:  # near the line centre the sine is relatively exact
        assert abs(float(_horner(np.array([x]), c)[0]) / np.sin(np.pi * float(x)) - 1.0) < 3e-7
    # a plausible mistake (a dropped r^8 term) fails
    bad = list(c)
    bad[4] = np.float32(0.0)
    assert (np.abs(_horner(r, bad).astype(np.float64)[live] - ref[live]) / np.abs(ref[live])).max() > 1e-5
