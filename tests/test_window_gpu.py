"""Support-windowed line reduction on the device against the FP64 oracle (GPU).

Bar (the scoring-only workload's): <= 1e-6 relative per bin on identical
segments and lines, plus 1e-12 of the spectrum's peak for bins where the
exact window part and the tail cancel to nearly zero.
"""

import numpy as np
import pytest

from tests.test_scoring_gpu import _dev, _segments

pytestmark = pytest.mark.gpu


def _stream():
    import ctypes as C

    import torch

    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _close(got, want):
    err = np.abs(got - want)
    tol = 1e-6 * np.abs(want) + 1e-12 * np.abs(want).max()
    assert np.all(err <= tol), (err / np.maximum(np.abs(want), 1e-300)).max()


@pytest.mark.parametrize("x_c", [100.0, 10.0, 1.0e12])
def test_window_reduce_matches_oracle(oracle, device, x_c):
    import torch

    from pyrite_amd import _lib

    rng = np.random.default_rng(11)
    n_lines = 20000
    n, E0, dE = 4096, 1000.0, 99000.0 / 4096
    E_r = rng.uniform(-10000.0, 120000.0, n_lines)  # includes lines outside the grid
    a = np.exp(rng.uniform(np.log(0.01), np.log(20.0), n_lines))
    w = rng.uniform(0.0, 1.0, n_lines).astype(np.float32)
    want = oracle.window_reduce(E_r, a, w, E0, dE, n, x_c)
    t = [torch.as_tensor(v, device=device) for v in (E_r, a, w)]
    out = torch.zeros(n, dtype=torch.float64, device=device)
    ws = _lib.DeviceScratch(device).get(_lib.workspace_bytes("hr_window_reduce_workspace_bytes", n_lines, n))
    _lib.call("hr_window_reduce", t[0].data_ptr(), t[1].data_ptr(), t[2].data_ptr(), n_lines, n, E0, dE, x_c,
              out.data_ptr(), ws, _stream())
    _close(out.cpu().numpy(), want)


def test_pixel_spectra_match_oracle(oracle, device):
    """Synthetic scoring workload at small scale: 64 reflections, 16 pixels, 4096 bins."""
    from pyrite_amd.spectrum.crystal import polarization_pairs
    from pyrite_amd.spectrum.lines import LineScorer, synthetic_line_tables

    rng = np.random.default_rng(3)
    n_E = 4096
    E0, dE = 1000.0 + 0.5 * 99000.0 / n_E, 99000.0 / n_E
    E_grid = E0 + dE * np.arange(n_E)
    tables = synthetic_line_tables(64, E_grid, rng)
    side = 4
    u = (np.arange(side) - (side - 1) / 2) * 0.02
    th, ph = np.meshgrid(1.0 + u, u)
    n_hats = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], -1).reshape(-1, 3)
    weights = np.full(n_hats.shape[0], 1e-4)
    seg = _segments(rng, 1500, thickness=5.0e6, E_range=(1000.0, 20000.0))
    sc = LineScorer(tables, E_grid, 5.0e6, n_hats, weights, device)
    got = sc.window_spectra(_dev(seg, device), 1.0, x_c=100.0).cpu().numpy()
    for d in range(n_hats.shape[0]):
        es, ep = polarization_pairs(n_hats[d], tables.g)
        oE, oa, ow, og = oracle.cxr_lines(seg, 5.0e6, n_hats[d], tables.as_dict(), es, ep)
        w32 = np.where(og, ow * weights[d], 0.0).astype(np.float32)
        want = oracle.window_reduce(oE, oa, w32, E0, dE, n_E, 100.0)
        _close(got[d], want)


def test_window_reduce_repeats_to_summation_order(device):
    """The block's FP64 LDS atomics land in hardware order (as the reference's
    device route's global FP64 atomics do, spectrum/cxr/_node_window_cuda.py:42,91),
    so repeated launches may differ, but only in FP64 summation order: every
    bin within 1e-13 of the spectrum's peak, where one line's deposit at its
    own node (weight >= 0.01 here, profile ~1) is a million times larger."""
    import torch

    from pyrite_amd import _lib

    rng = np.random.default_rng(12)
    n_lines = 50000
    n, E0, dE = 4096, 1000.0, 99000.0 / 4096
    E_r = rng.uniform(1000.0, 100000.0, n_lines)
    a = np.exp(rng.uniform(np.log(0.05), np.log(5.0), n_lines))
    w = rng.uniform(0.01, 1.0, n_lines).astype(np.float32)
    t = [torch.as_tensor(v, device=device) for v in (E_r, a, w)]
    ws = _lib.DeviceScratch(device).get(_lib.workspace_bytes("hr_window_reduce_workspace_bytes", n_lines, n))
    outs = []
    for _ in range(4):
        out = torch.zeros(n, dtype=torch.float64, device=device)
        _lib.call("hr_window_reduce", t[0].data_ptr(), t[1].data_ptr(), t[2].data_ptr(), n_lines, n, E0, dE, 100.0,
                  out.data_ptr(), ws, _stream())
        outs.append(out.cpu().numpy())
    peak = np.abs(outs[0]).max()
    assert 1e-13 * peak < 1e-6 * 0.01
    for o in outs[1:]:
        assert np.abs(o - outs[0]).max() <= 1e-13 * peak
