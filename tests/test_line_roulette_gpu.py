"""Per-primary line roulette on the device (GPU).

Draws and selection are bit-for-bit the oracle's (which reproduces the
reference's pinned draws); floor 1 is the analog tally exactly; and the
rouletted tally is the inverse-probability sum of the analog primaries it
keeps (the reference's estimator identity, test_line_roulette.py:213-224).
"""

import numpy as np
import pytest

from oracle import line_roulette as olr

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("seed,first,n", [(7, 0, 3), (7, 0, 5000), (123456789, 10**9, 4097), (2**64 - 1, 5, 100)])
def test_device_draws_are_the_oracle_draws(device, seed, first, n):
    from pyrite_amd.line_roulette import roulette_uniforms

    got = roulette_uniforms(seed, first, n, device).cpu().numpy()
    np.testing.assert_array_equal(got, olr.roulette_uniforms(seed, np.arange(first, first + n)))


def test_device_selection_is_the_oracle_selection(device):
    import torch

    from pyrite_amd import _lib

    rng = np.random.default_rng(5)
    m = rng.lognormal(0.0, 1.5, 3000)
    m[rng.random(m.size) < 0.1] = 0.0
    for first, pilot, thr, floor, seed in ((0, 8, 1.3, 0.5, 7), (100, 128, 0.7, 0.25, 99), (0, 10**6, 1.0, 0.5, 1)):
        md = torch.as_tensor(m, device=device)
        p = torch.empty(m.size, dtype=torch.float64, device=device)
        k = torch.empty(m.size, dtype=torch.uint8, device=device)
        _lib.call("hr_line_screen", md.data_ptr(), m.size, first, pilot, thr, floor, seed, p.data_ptr(),
                  k.data_ptr(), None)
        wp, wk = olr.select(m, first, pilot, thr, floor, seed)
        np.testing.assert_array_equal(p.cpu().numpy(), wp)
        np.testing.assert_array_equal(k.cpu().numpy().astype(bool), wk)


def _pipe(route, roulette, Ne=4000):
    from pyrite_amd.pipeline import CaseConfig, CasePipeline

    cfg = CaseConfig(Ne=Ne, line_route=route, E0_keV=2000.0, thickness_ang=2.0e5, line_roulette=roulette)
    return CasePipeline(cfg)


@pytest.mark.parametrize("route", ["dense", "window"])
def test_floor_one_equals_analog_scoring(device, route):
    a = _pipe(route, None)
    a.step(0, 1)
    b = _pipe(route, {"floor": 1.0, "keep_fraction": 0.25, "pilot_histories": 8})
    b.step(0, 1)
    la, lb = a.spectra()[0], b.spectra()[0]
    assert la.sum() > 0
    if route == "dense":
        np.testing.assert_array_equal(lb, la)  # same lines, same weights, same fixed summation order
    else:
        np.testing.assert_allclose(lb, la, rtol=1e-12, atol=1e-15 * la.max())


def test_rouletted_tally_is_the_inverse_probability_sum_of_kept_primaries(device):
    import torch

    policy = {"floor": 0.5, "keep_fraction": 0.25, "pilot_histories": 8}
    p = _pipe("dense", policy)
    res = p.transport(0, compact=False)
    seg = p.records(res, first_id=0)
    scr = p.line_screen(p.cfg.Ne)
    w = scr.row_weights(p.lines, seg, seg["prim"], p.cfg.Ne)
    mass = scr.mass.cpu().numpy()
    # the screen's masses are the sum of w pi / a over each primary's lines
    E_r, a, lw = p.lines.lines(seg, 1.0 / p.cfg.Ne)
    keep = (lw != 0) & (a > 0)
    n_dir, n_g = p.lines.n_dir, len(p.lines.tables.wm)
    seg_of_line = (torch.arange(E_r.shape[0], device=device) // n_g) % seg["L"].shape[0]
    want = torch.zeros(p.cfg.Ne, dtype=torch.float64, device=device)
    want.index_add_(0, seg["prim"][seg_of_line[keep]], lw[keep].double() * np.pi / a[keep])
    np.testing.assert_allclose(mass, want.cpu().numpy(), rtol=1e-12, atol=0)
    # selection = the oracle's, with the pilot's quantile as threshold
    thr = olr.threshold(mass[:8], 0.25)
    assert scr.threshold == thr
    wp, wk = olr.select(mass, 0, 8, thr, 0.5, p.cfg.seed)
    np.testing.assert_array_equal(scr.probability.cpu().numpy(), wp)
    np.testing.assert_array_equal(scr.kept.cpu().numpy().astype(bool), wk)
    assert np.any(wp < 1.0) and not np.all(wk[wp < 1.0]), "fixture must cull a primary"
    # rouletted tally vs analog tallies of the sure and floor strata
    prim = seg["prim"].cpu().numpy()
    sure = torch.as_tensor((wk & (wp == 1.0))[prim].astype(np.float32), device=device)
    low = torch.as_tensor((wk & (wp < 1.0))[prim].astype(np.float32), device=device)
    s_r = p.lines.spectrum(seg, p.cfg.Ne, row_weight=w).cpu().numpy()
    s_sure = p.lines.spectrum(seg, p.cfg.Ne, row_weight=sure).cpu().numpy()
    s_low = p.lines.spectrum(seg, p.cfg.Ne, row_weight=low).cpu().numpy()
    assert s_low.sum() > 0
    np.testing.assert_allclose(s_r, s_sure + s_low / 0.5, rtol=1e-9, atol=1e-12 * s_r.max())
    rec = scr.record()
    assert rec["primaries_at_floor"] == int((wp < 1.0).sum())
    assert rec["line_mass_analog"] == pytest.approx(mass.sum(), rel=1e-12)
