"""Continuation rounds against the uninterrupted device transport (GPU).

The reference's bar for its continuation driver (continuation.py:9-19): a
history split over launches continues exactly as one uninterrupted run.  So
every row column, the per-history row count and exit code and the final state
must equal ``hr_transport_run``'s bit for bit, for lane populations that force
refills, for short rounds, for the packed tail and for both transport
precisions.  (The uninterrupted run is itself pinned to the oracle by
test_transport_gpu.py.)
"""

import numpy as np
import pytest

from tests.test_transport_gpu import _gpu_run
from tests.transport_cases import FRAMED, HOPG_5MEV, MULTILAYER, SI_30KEV_STRAGGLE

pytestmark = pytest.mark.gpu


def _continued(case, device, lanes, R, precision="fp64", max_steps=400_000, sort_lanes=False):
    import torch

    from pyrite_amd import _lib
    from pyrite_amd.transport.continuation import ContinuationDriver, canonical_rows
    from pyrite_amd.transport.core import DeviceTables, TransportControl

    lt, lut = case.tables()
    pos, dirs, E, E_cut = case.state()
    tabs = DeviceTables(lt, lut, device)
    ctl = TransportControl(max_steps=max_steps, energy_model=case.energy_model, max_dE_frac=case.max_dE_frac,
                           straggle_on=case.straggle, finite_footprint=case.finite is not None,
                           width_ang=case.finite[0] if case.finite else 0.0,
                           height_ang=case.finite[1] if case.finite else 0.0, precision=precision)
    drv = ContinuationDriver(tabs, ctl, lanes, R, device, float(E_cut[0]), sort_lanes=sort_lanes)
    pos_d, dirs_d, E_d = (torch.as_tensor(x, device=device) for x in (pos, dirs, E))

    def fill(start, n, p, d, e, clock, keys):
        p[:n] = pos_d[start:start + n]
        d[:n] = dirs_d[start:start + n]
        e[:n] = E_d[start:start + n]
        clock[:n] = 0.0
        _lib.call("hr_stream_keys", case.seed, start, n, keys.data_ptr(), None)

    rounds = []

    def consume(rnd):
        # copy what this launch produced: the next launch reuses the buffers
        cols = {k: v[: rnd.n_lanes * rnd.rows_per_round].clone() for k, v in rnd.rows.items()}
        done = (rnd.history >= 0) & (rnd.exit_code != 6)
        lanes_done = torch.nonzero(done).squeeze(1)
        fin = {"history": rnd.history[lanes_done].clone(), "code": rnd.exit_code[lanes_done].clone(),
               "pos": drv.pos[lanes_done].clone(), "dirs": drv.dirs[lanes_done].clone(),
               "E_keV": drv.E[lanes_done].clone(), "clock": drv.clock[lanes_done].clone(),
               "stragg_dE": drv.stragg[lanes_done].clone(), "nseg": drv.nseg[lanes_done].clone()}
        rounds.append((rnd.n_lanes, rnd.rows_per_round, cols, rnd.seg_count.clone(), rnd.history.clone(),
                       rnd.row_start.clone(), fin))

    res = drv.run(0, case.Ne, fill, consume)
    torch.cuda.synchronize()
    return res, rounds


@pytest.mark.parametrize("case,lanes,R,precision,max_steps,sort_lanes", [
    (SI_30KEV_STRAGGLE, 64, 3, "fp64", 400_000, False),
    (HOPG_5MEV, 50, 7, "fp64", 400_000, False),
    (HOPG_5MEV, 96, 5, "mixed", 400_000, False),
    (MULTILAYER, 40, 2, "fp64", 400_000, False),
    (FRAMED, 128, 16, "fp64", 400_000, False),
    (HOPG_5MEV, 37, 4, "fp64", 30, False),  # step limit reached inside a round
    (SI_30KEV_STRAGGLE, 64, 3, "mixed", 400_000, True),  # lanes re-ordered by energy every launch
], ids=["si30s", "hopg5", "hopg5-mixed", "stack", "prism", "hopg5-step-limit", "si30s-mixed-sorted"])
def test_continuation_is_bitwise_the_uninterrupted_run(device, case, lanes, R, precision, max_steps, sort_lanes):
    import torch

    from pyrite_amd.transport.continuation import ContinuationRound, canonical_rows

    cap = 4096
    ref_rows, ref_cnt, ref_code, ref_st = _gpu_run(case, device, cap=cap, precision=precision, max_steps=max_steps)
    assert ref_cnt.max() <= cap
    if max_steps < 1000:
        assert np.any(ref_code == 4), "fixture must hit the step limit"
    keep = (np.arange(cap)[None, :] < ref_cnt[:, None]).reshape(-1)
    ref = {k: v[keep] for k, v in ref_rows.items()}
    n_rows = int(ref_cnt.sum())
    offsets = torch.as_tensor(np.concatenate([[0], np.cumsum(ref_cnt)[:-1]]), device=device)

    res, rounds = _continued(case, device, lanes, R, precision, max_steps, sort_lanes)
    assert res.n_segments == n_rows
    assert res.rounds == len(rounds) and len(rounds) > 2
    seen = np.zeros(n_rows, dtype=np.int64)
    # the frozen schema has no end-state / event columns (batching.py:398-400)
    got = {k: np.zeros_like(v) for k, v in ref.items() if k in rounds[0][2]}
    cnt = np.full(case.Ne, -1, dtype=np.int64)
    code = np.full(case.Ne, -1, dtype=np.int64)
    st = {k: np.zeros_like(v) for k, v in ref_st.items()}
    for n_lanes, Rr, cols, seg_count, history, row_start, fin in rounds:
        rnd = ContinuationRound(0, n_lanes, Rr, cols, seg_count, history, row_start, None)
        slot, canon = canonical_rows(rnd, offsets)
        c = canon.cpu().numpy()
        np.add.at(seen, c, 1)
        s = slot.cpu().numpy()
        for k in got:
            got[k][c] = cols[k].cpu().numpy()[s]
        h = fin["history"].cpu().numpy()
        cnt[h] = fin["nseg"].cpu().numpy()
        code[h] = fin["code"].cpu().numpy()
        for k in st:
            st[k][h] = fin[k].cpu().numpy()
    assert np.all(seen == 1)
    np.testing.assert_array_equal(cnt, ref_cnt)
    np.testing.assert_array_equal(code, ref_code)
    assert len(got) == (12 if case.energy_model == 1 else 7)
    for k in got:
        np.testing.assert_array_equal(got[k], ref[k], err_msg=k)
    for k in ref_st:
        if k == "stragg_dE" and not case.straggle:
            continue
        np.testing.assert_array_equal(st[k], ref_st[k], err_msg=k)


@pytest.mark.parametrize("route", ["dense", "window"])
def test_streamed_case_matches_the_batched_route(device, route):
    """Streamed transport -> scoring gives the batched route's tallies (same rows;
    only the summation order of the tallies differs)."""
    from pyrite_amd.pipeline import CaseConfig, CasePipeline

    cfg = CaseConfig(Ne=3000, line_route=route, E0_keV=2000.0, thickness_ang=2.0e5)
    a = CasePipeline(cfg)
    res = a.step(0, 1)
    la, ba = a.spectra()
    b = CasePipeline(cfg)
    sres = b.stream(0, cfg.Ne, lanes=700, rows_per_round=6)
    lb, bb = b.spectra()
    assert sres.n_segments == res.n_segments and sres.counts == res.counts
    assert la.sum() > 0 and ba.sum() > 0
    # sinc^2 cells are summed in FP32 over groups of 4 lines and brem cells
    # over groups of 8 rows before each FP64 accumulate, so another row order
    # moves the last bits at about 1e-8
    np.testing.assert_allclose(lb, la, rtol=1e-7, atol=1e-12 * la.max())
    np.testing.assert_allclose(bb, ba, rtol=1e-7, atol=1e-10 * ba.max())
