"""Overlapped batches (transport of batch s+1 while batch s is scored) give
the tallies of back-to-back steps bit for bit (GPU)."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("route", ["dense", "window"])
def test_overlapped_steps_equal_sequential_steps(device, route):
    from pyrite_amd.pipeline import CaseConfig, CasePipeline

    cfg = CaseConfig(Ne=6000, line_route=route, E0_keV=3000.0, thickness_ang=3.0e5)
    # equal arenas: equal sub-batches, so equal line groups in the FP32 cells
    a = CasePipeline(cfg, arena_bytes=2 << 30)
    for b in range(4):
        a.step(b, 4)
    la, ba = a.spectra()
    p = CasePipeline(cfg, arena_bytes=2 << 30)
    p.run_steps(0, 4, overlap=True)
    lb, bb = p.spectra()
    assert p.histories == a.histories and p.segments == a.segments
    assert la.sum() > 0 and ba.sum() > 0
    if route == "dense":
        np.testing.assert_array_equal(lb, la)
        np.testing.assert_array_equal(bb, ba)
    else:  # LDS FP64 atomics in the windowed route: summation order varies
        np.testing.assert_allclose(lb, la, rtol=1e-12, atol=1e-15 * la.max())
        np.testing.assert_array_equal(bb, ba)


def test_row_threads_gather_equals_slot_threads(device):
    """hr_vr_gather with one thread per written row (rows found by binary
    search over the offsets, including particles without rows) writes exactly
    what one thread per slot writes."""
    import ctypes as C

    import torch

    from pyrite_amd import _lib

    rng = np.random.default_rng(9)
    n, cap = 777, 13
    cnt = rng.integers(0, cap + 4, n)  # some empty, some clamped
    cnt[:5] = 0
    cnt[-3:] = 0
    dev = device
    seg_count = torch.as_tensor(cnt, dtype=torch.int64, device=dev)
    c = seg_count.clamp(max=cap)
    off = (torch.cumsum(c, 0) - c + 40).contiguous()  # rows start at an offset, as for a second slot group
    n_rows = int(c.sum())
    rows = {k: torch.as_tensor(rng.normal(size=(n * cap, 3) if w == 3 else n * cap), device=dev)
            for k, w in (("seg_mid", 3), ("seg_dir", 3), ("seg_len", 1), ("seg_E", 1), ("seg_E_end", 1))}
    weight = torch.as_tensor(rng.uniform(size=n), device=dev)
    outs = []
    for n_out in (0, n_rows):
        o = {k: torch.zeros((40 + n_rows,) + ((3,) if k in ("r_mid", "v_hat") else ()), dtype=torch.float32,
                            device=dev) for k in ("r_mid", "v_hat", "L", "E", "w")}
        prim = torch.full((40 + n_rows,), -7, dtype=torch.int64, device=dev)
        g = _lib.VrGatherArgs()
        g.n, g.rows, g.n_out = n, cap, n_out
        g.seg_count, g.row_offset, g.weight = seg_count.data_ptr(), off.data_ptr(), weight.data_ptr()
        g.seg_mid, g.seg_dir = rows["seg_mid"].data_ptr(), rows["seg_dir"].data_ptr()
        g.seg_len, g.seg_E, g.seg_E_end = rows["seg_len"].data_ptr(), rows["seg_E"].data_ptr(), \
            rows["seg_E_end"].data_ptr()
        g.r_mid, g.v_hat, g.L, g.E = o["r_mid"].data_ptr(), o["v_hat"].data_ptr(), o["L"].data_ptr(), \
            o["E"].data_ptr()
        g.row_weight, g.row_prim = o["w"].data_ptr(), prim.data_ptr()
        _lib.call("hr_vr_gather", C.byref(g), None)
        torch.cuda.synchronize()
        outs.append({k: v.cpu().numpy() for k, v in o.items()} | {"prim": prim.cpu().numpy()})
    for k in outs[0]:
        np.testing.assert_array_equal(outs[1][k], outs[0][k], err_msg=k)
    assert (outs[0]["prim"][40:] >= 0).all() and (outs[0]["prim"][:40] == -7).all()
