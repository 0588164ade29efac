"""Detector-cone splitting on the device against the FP64 oracle (GPU).

* the single-event copy (``hr_cone_event``) against the oracle's restatement
  on the same inputs, to rtol 1e-12 (host/device libm ulps);
* split keys and the copy records (``hr_cone_collect``: order, ladder
  roulette, copy stream keys) bit for bit on the same slots;
* device invariances: the energy gate above the beam leaves transport analog
  bit for bit, primaries follow their analog rows until cone entry, and a
  launch rerun for copies beyond the slots gives the same copies;
* the in-cone path-length tally (primary plus weighted copies, the copies
  transported on the device) matches analog within 4 sigma and its variance
  drops;
* the HOPG 5 MeV case with the policy's defaults: line and brem spectra per
  bin against the oracle's split case (copies scored per weight stratum), and
  at a 60 deg detector against the analog device case, to the binomial /
  4 sigma bar;
* the learned copy roulette: keep decisions bit for bit against the rule
  restated on the host, expectation kept, and the case with it per bin
  against plain splitting.
"""

import math

import numpy as np
import pytest

from tests.test_case_gpu import _agree
from tests.test_cone_split_oracle import HBN_50KEV, HBN_50KEV_ELSEPA, N_HAT, THETA_C

pytestmark = pytest.mark.gpu


def _cone_args(n_hat=N_HAT, theta=THETA_C, target=0.05, angle_scale=0.0, min_energy_keV=0.0):
    from pyrite_amd import _lib

    c = _lib.ConeSplitArgs()
    n = np.asarray(n_hat, dtype=float)
    n = n / np.linalg.norm(n)
    c.on = 1
    c.nx, c.ny, c.nz = (float(v) for v in n)
    c.cos_c, c.sin_c = math.cos(theta), math.sin(theta)
    c.target_weight, c.angle_scale, c.min_energy_keV = target, angle_scale, min_energy_keV
    return c


def test_single_event_copy_matches_oracle(oracle, device):
    import ctypes as C

    import torch

    from pyrite_amd import _lib

    rng = np.random.default_rng(4)
    n = 4000
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    d[:50] = (0.0, 0.0, 1.0)  # along n and against it: the frame's fallback branch
    d[50:100] = (0.0, 0.0, -1.0)
    alpha = 10.0 ** rng.uniform(-7, 1, n)
    E = rng.uniform(50.0, 5000.0, n)
    keys = np.array([oracle.cone_event_key(int(k), f) for f, k in enumerate(rng.integers(0, 2**63, n))],
                    dtype=np.uint64)
    for target, u_ref in ((1e-3, 0.0), (0.3, 0.5)):
        c = _cone_args((0.0, 0.0, 1.0), 0.7, target, u_ref)
        out = torch.zeros((n, 5), dtype=torch.float64, device=device)
        t = [torch.as_tensor(x, device=device).contiguous() for x in (d, alpha, E)]
        kt = torch.as_tensor(keys.view(np.int64), device=device).contiguous()
        _lib.check(_lib.load().hr_cone_event(C.byref(c), n, t[0].data_ptr(), t[1].data_ptr(), t[2].data_ptr(),
                                             kt.data_ptr(), out.data_ptr(), None), "hr_cone_event")
        got = out.cpu().numpy()
        made, ref = oracle.cone_event(d, alpha, E, keys, dict(n_hat=(0, 0, 1), half_angle_rad=0.7,
                                                              target_weight=target, angle_scale=u_ref))
        assert made.sum() > 100
        np.testing.assert_array_equal(got[:, 0] == 1.0, made)
        np.testing.assert_allclose(got[made, 1:4], ref[made, :3], rtol=1e-12, atol=1e-14)
        np.testing.assert_allclose(got[made, 4], ref[made, 3], rtol=1e-12)


def _device_run(case, device, seed, ne, cone=None, slots=16, keys=None, state=None, cap=4096):
    """Transport on the device with or without splitting; returns host rows,
    counts, codes and (with cone) the core's collected copies."""
    import torch

    from pyrite_amd import _lib
    from pyrite_amd.cone_split import ConeSplitCore
    from pyrite_amd.transport.core import DeviceTables, TransportControl, alloc_rows, make_hip_transport_core

    lt, lut = case.tables()
    tabs = DeviceTables(lt, lut, device)
    ctl = TransportControl(energy_model=1, straggle_on=False, finite_footprint=case.finite is not None,
                           width_ang=case.finite[0] if case.finite else 0.0,
                           height_ang=case.finite[1] if case.finite else 0.0)
    if keys is None:
        keys = torch.empty(ne, dtype=torch.uint64, device=device)
        _lib.call("hr_stream_keys", seed, 0, ne, keys.data_ptr(), None)
    if state is None:
        bd = np.asarray(case.beam_dir, dtype=float)
        pos = torch.zeros((ne, 3), dtype=torch.float64, device=device)
        dirs = torch.as_tensor(np.tile(bd / np.linalg.norm(bd), (ne, 1)), device=device).contiguous()
        E = torch.full((ne,), case.E0, dtype=torch.float64, device=device)
        clock = torch.zeros(ne, dtype=torch.float64, device=device)
    else:
        pos, dirs, E, clock = (x.clone() for x in state)
    alive = torch.ones(ne, dtype=torch.uint8, device=device)
    Ecut = torch.full((ne,), case.E_cut, dtype=torch.float64, device=device)
    stragg = torch.zeros(ne, dtype=torch.float64, device=device)
    rows = alloc_rows(ne, cap, True, device)
    cnt = torch.zeros(ne, dtype=torch.int64, device=device)
    code = torch.zeros(ne, dtype=torch.int8, device=device)
    core = make_hip_transport_core(tabs, ctl) if cone is None else ConeSplitCore(tabs, ctl, cone, device, slots)
    core((0, ne, cap, keys), (alive, clock, pos, dirs, E, Ecut), rows, (cnt, code), (False, stragg))
    copies = core.copies() if cone is not None else None
    torch.cuda.synchronize()
    host = {k: v.cpu().numpy() for k, v in rows.items() if k != "work"}
    return host, cnt.cpu().numpy(), code.cpu().numpy(), copies, core, tabs, ctl


def _request(**kw):
    from pyrite_amd.cone_split import ConeSplit

    base = dict(max_copies_per_electron=1024, weight_ladder=True)
    base.update(kw)
    return ConeSplit(N_HAT, THETA_C, base.pop("target_weight", 0.05), **base)


def test_split_keys_and_collect_match_oracle(oracle, device):
    import ctypes as C

    import torch

    from pyrite_amd import _lib
    from pyrite_amd.cone_split import split_keys

    ne = 3000
    keys = torch.empty(ne, dtype=torch.uint64, device=device)
    _lib.call("hr_stream_keys", 3, 0, ne, keys.data_ptr(), None)
    sk = split_keys(keys).cpu().numpy()
    np.testing.assert_array_equal(sk, oracle.cone_split_keys(oracle.stream_keys(3, ne)))
    _, _, _, copies, core, _, _ = _device_run(HBN_50KEV, device, 3, ne, cone=_request())
    # the oracle's collect on the very slots the device wrote
    s = core._slots
    k = core.k
    rec = {"split_key": sk, "cap": k, "count": s["count"][:ne].cpu().numpy(),
           "killed": s["killed"][:ne].cpu().numpy()}
    for name in ("pos", "dir", "E_keV", "clock", "weight", "flight"):
        rec[name] = np.ascontiguousarray(s[name][:ne * k].cpu().numpy())
    c = oracle.ConeSplit()
    c.on, c.cap = 1, k
    c.split_key = oracle._p(rec["split_key"])
    for name in ("count", "killed", "pos", "dir", "E_keV", "clock", "weight", "flight"):
        setattr(c, name, oracle._p(rec[name]))
    rec["struct"] = c
    ref = oracle.cone_collect(rec, ladder=True)
    got = {kk: v.cpu().numpy() for kk, v in copies.items()}
    assert ref["parent"].size > 50
    for name in ("parent", "flight", "stream_key", "weight", "pos", "dir", "E_keV", "clock"):
        np.testing.assert_array_equal(got[name], ref[name], err_msg=name)
    assert ref["max_count"] <= k


def test_gate_above_beam_prefix_and_rerun_invariance(device):
    ne = 1500
    rows_a, cnt_a, code_a, _, _, _, _ = _device_run(HBN_50KEV, device, 7, ne)
    rows_g, cnt_g, code_g, cp_g, _, _, _ = _device_run(HBN_50KEV, device, 7, ne, cone=_request(min_energy_keV=60.0))
    assert cp_g is None or cp_g["weight"].numel() == 0
    np.testing.assert_array_equal(cnt_a, cnt_g)
    np.testing.assert_array_equal(code_a, code_g)
    for k in ("seg_mid", "seg_dir", "seg_len", "seg_E"):
        np.testing.assert_array_equal(rows_a[k], rows_g[k])
    rows_s, cnt_s, code_s, cp, core, _, _ = _device_run(HBN_50KEV, device, 7, ne, cone=_request())
    killed = cp["killed"].cpu().numpy().astype(bool)
    assert killed.any() and cp["weight"].numel() > 0
    cap = 4096
    for e in range(ne):
        a = rows_a["seg_mid"][e * cap:e * cap + cnt_a[e]]
        b = rows_s["seg_mid"][e * cap:e * cap + cnt_s[e]]
        if killed[e]:
            assert code_s[e] == 6 and b.shape[0] <= a.shape[0]
            np.testing.assert_array_equal(a[:b.shape[0]], b)
        else:
            np.testing.assert_array_equal(a, b)
    d = cp["dir"].cpu().numpy()
    assert np.all(d @ np.asarray(N_HAT) >= math.cos(THETA_C) - 1e-12)
    # one slot per electron forces reruns with room; the copies are the same
    _, cnt_r, _, cp_r, core_r, _, _ = _device_run(HBN_50KEV, device, 7, ne, cone=_request(), slots=1)
    assert core_r.reruns >= 1
    np.testing.assert_array_equal(cnt_r, cnt_s)
    for name in ("parent", "flight", "stream_key", "weight", "pos", "dir"):
        np.testing.assert_array_equal(cp_r[name].cpu().numpy(), cp[name].cpu().numpy(), err_msg=name)


def _in_cone_tally(rows, cnt, n, cap=4096):
    ids = np.repeat(np.arange(n), np.minimum(cnt, cap))
    keep = (np.arange(cap)[None, :] < cnt[:, None]).reshape(-1)
    v, L = rows["seg_dir"][keep], rows["seg_len"][keep]
    inside = v @ np.asarray(N_HAT) >= math.cos(THETA_C)
    return np.bincount(ids, weights=L * inside, minlength=n), np.bincount(ids, weights=L, minlength=n)


def test_tabulated_single_event_copy_matches_oracle(oracle, device):
    """hr_cone_event_tab (the reference's law: tabulated rows, band by
    bisection on the mixed quantile) against the oracle on the same inputs."""
    import ctypes as C

    import torch

    from pyrite_amd import _lib
    from pyrite_amd.transport.core import DeviceTables
    from pyrite_amd.transport.tables import build_transport_lut

    lt, _ = HBN_50KEV_ELSEPA.tables()
    lut = build_transport_lut(1.0, 50.0, lt)
    tabs = DeviceTables(lt, lut, device)
    rng = np.random.default_rng(8)
    n = 4000
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    d[:50] = (0.0, 0.0, 1.0)
    el = rng.integers(0, 2, n).astype(np.int64)  # B (screened-Rutherford shape) and N (vendor shape)
    E = np.exp(rng.uniform(np.log(1.0), np.log(50.0), n))
    keys = np.array([oracle.cone_event_key(int(k), f) for f, k in enumerate(rng.integers(0, 2**63, n))],
                    dtype=np.uint64)
    for target, u_ref in ((1e-3, 0.0), (0.3, 0.5)):
        c = _cone_args((0.0, 0.0, 1.0), 0.7, target, u_ref)
        out = torch.zeros((n, 5), dtype=torch.float64, device=device)
        t = [torch.as_tensor(x, device=device).contiguous() for x in (d, el, E)]
        kt = torch.as_tensor(keys.view(np.int64), device=device).contiguous()
        _lib.check(_lib.load().hr_cone_event_tab(C.byref(c), C.byref(tabs.elsepa_args), n, t[0].data_ptr(),
                                                 t[1].data_ptr(), t[2].data_ptr(), kt.data_ptr(), out.data_ptr(),
                                                 None), "hr_cone_event_tab")
        got = out.cpu().numpy()
        made, ref = oracle.cone_event_tab(d, lt.elsepa, el, E, keys, dict(n_hat=(0, 0, 1), half_angle_rad=0.7,
                                                                          target_weight=target, angle_scale=u_ref))
        assert made.sum() > 100
        np.testing.assert_array_equal(got[:, 0] == 1.0, made)
        np.testing.assert_allclose(got[made, 1:4], ref[made, :3], rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(got[made, 4], ref[made, 3], rtol=1e-12)


@pytest.mark.parametrize("case", [HBN_50KEV, HBN_50KEV_ELSEPA], ids=lambda c: c.name)
def test_split_in_cone_tally_matches_analog_on_device(device, case):
    from pyrite_amd.cone_split import history_sums

    ne = 20000
    va, vb, ma, mb = [], [], [], []
    for seed in range(3):
        rows, cnt, _, _, _, _, _ = _device_run(case, device, 100 + seed, ne)
        ca, la = _in_cone_tally(rows, cnt, ne)
        rows, cnt, _, cp, _, _, _ = _device_run(case, device, 100 + seed, ne, cone=_request())
        cb, lb = _in_cone_tally(rows, cnt, ne)
        nc = int(cp["weight"].numel())
        assert nc > 0
        rc, cc, _, _, _, _, _ = _device_run(case, device, 0, nc, keys=cp["stream_key"],
                                            state=(cp["pos"], cp["dir"], cp["E_keV"], cp["clock"]))
        c2, l2 = _in_cone_tally(rc, cc, nc)
        par, w = cp["parent"].cpu().numpy(), cp["weight"].cpu().numpy()
        cb = history_sums(cb, c2, par, w, ne)
        lb = history_sums(lb, l2, par, w, ne)
        ma.append([ca.mean(), la.mean()])
        mb.append([cb.mean(), lb.mean()])
        va.append([ca.var(ddof=1) / ne, la.var(ddof=1) / ne])
        vb.append([cb.var(ddof=1) / ne, lb.var(ddof=1) / ne])
    ma, mb = np.mean(ma, 0), np.mean(mb, 0)
    se = np.sqrt((np.sum(va, 0) + np.sum(vb, 0)) / 9.0)
    assert np.all(np.abs(ma - mb) < 4.0 * se), (ma, mb, se)
    assert np.sum(vb, 0)[0] < np.sum(va, 0)[0]


def test_hopg_case_with_cone_split_agrees_with_oracle_and_analog(device):
    """At the 90 deg detector: the device's split case per bin against the
    oracle's split case.  Against analog at a 60 deg detector, where the
    analog run samples the cone's population well: with the detector at 90 deg
    only about 1 primary in 1.2e4 enters the cone and its contributions are
    heavy-tailed (long grazing paths under the exit face), so an analog
    sample of this size misses most of them and is no reference there."""
    from oracle import case as ocase
    from pyrite_amd.pipeline import CaseConfig, CasePipeline
    from tests.test_case_gpu import BATCHES, PER_BATCH, _gpu_batches

    split = CasePipeline(CaseConfig(Ne=PER_BATCH, cone_split={}))
    g_lines, g_brem = _gpu_batches(split)
    assert split.cone_stats["copies"] > 0 and split.cone_stats["copy_rows"] > 0
    lines, brem = [], []
    n = PER_BATCH
    for b in range(BATCHES):
        ln, br, _ = ocase.run_case(split, 40_000_000 + b * n, n, cone=split.cone_request)
        lines.append(ln / n)
        brem.append(br / (4.0 * np.pi * n) * split.omega)
    c_lines, c_brem = np.array(lines), np.array(brem)
    _agree(g_brem, c_brem, "split brem vs oracle")
    _agree(g_lines, c_lines, "split lines vs oracle")
    split.release()
    cs = {"target_weight": 0.05}  # about one copy per primary at 60 deg
    split60 = CasePipeline(CaseConfig(Ne=4096, theta_obs_deg=60.0, cone_split=cs))
    s_lines, s_brem = _gpu_batches(split60)
    analog60 = CasePipeline(CaseConfig(Ne=4096, theta_obs_deg=60.0, seed=2))
    a_lines, a_brem = _gpu_batches(analog60)
    _agree(s_brem, a_brem, "split vs analog brem (60 deg)")
    _agree(s_lines, a_lines, "split vs analog lines (60 deg)")


def test_copy_roulette_decisions_and_expectation(oracle, device):
    """``hr_cone_roulette``: keep flags bit for bit against the rule restated on
    the host (p = clamp(w I / T, p_min, 1), draw counter 0 of
    splitmix64(key ^ 0x61C8864680B583EB)), survivors carry w/p, and the
    weight sum keeps its expectation."""
    import torch

    from pyrite_amd import _lib

    rng = np.random.default_rng(9)
    n = 200_000
    w = np.exp2(rng.integers(-12, 2, n)).astype(np.float64)
    log_imp = rng.normal(0.0, 2.0, n).astype(np.float32)
    keys = rng.integers(0, 2**63, n, dtype=np.int64).view(np.uint64)
    log_T, p_min = 0.5, 1.0 / 64
    wt = torch.as_tensor(w, device=device).clone()
    keep = torch.empty(n, dtype=torch.uint8, device=device)
    li = torch.as_tensor(log_imp, device=device)
    kt = torch.as_tensor(keys.view(np.int64), device=device)
    _lib.call("hr_cone_roulette", n, li.data_ptr(), kt.data_ptr(), log_T, p_min, wt.data_ptr(), keep.data_ptr(), None)
    got_keep = keep.cpu().numpy().astype(bool)
    got_w = wt.cpu().numpy()
    s = np.log(w) + log_imp.astype(np.float64) - log_T
    p = np.where(s >= 0.0, 1.0, np.maximum(p_min, np.exp(s)))
    u = np.array([oracle.stream_uniform(oracle.splitmix64(int(k) ^ 0x61C8864680B583EB), 0) for k in keys])
    want = (p >= 1.0) | (u < p)
    np.testing.assert_array_equal(got_keep, want)
    np.testing.assert_allclose(got_w[want], (w / p)[want], rtol=1e-15)
    assert np.all(got_w[~want] == 0.0)
    assert np.all(got_w[s >= 0.0] == w[s >= 0.0])
    # E[sum w'] = sum w
    var = np.sum(w * w * (1.0 - p) / p)
    assert abs(got_w.sum() - w.sum()) < 4.0 * math.sqrt(var)


def test_learned_copy_roulette_case_is_unbiased(device):
    """The HOPG case with cone splitting (no angle roulette) and the learned
    copy roulette, trained on a few batches, against the same splitting
    without the roulette, per bin (disjoint histories): the roulette depends
    only on the copy's state, so tallies keep their expectation whatever the
    network predicts."""
    from pyrite_amd.pipeline import CaseConfig, CasePipeline
    from pyrite_amd.vr.cone_roulette import LearnedConeRoulette
    from tests.test_case_gpu import _gpu_batches

    cs = {"angle_scale": None}
    learned = CasePipeline(CaseConfig(Ne=4096, cone_split=cs))
    lr = LearnedConeRoulette(learned, keep_quantile=0.9)
    x, y, w = lr.training_set(4, first_batch=70_000)
    lr.fit(x, y, w)
    learned.cone_roulette = lr
    g_lines, g_brem = _gpu_batches(learned)
    assert 0 < lr.stats["kept"] < lr.stats["seen"]
    learned.release()
    plain = CasePipeline(CaseConfig(Ne=4096, seed=3, cone_split=cs))
    p_lines, p_brem = _gpu_batches(plain)
    _agree(g_brem, p_brem, "learned-roulette brem vs plain split")
    _agree(g_lines, p_lines, "learned-roulette lines vs plain split")


def test_overlapped_cone_steps_equal_back_to_back(device):
    """``run_steps`` with cone splitting on two streams (primaries and copy
    generations on one, scoring on the other) gives the tallies of ``step()``
    back to back, bit for bit: the same copies, pooled and transported in the
    same order, scored in the same order."""
    import torch

    from pyrite_amd.pipeline import CaseConfig, CasePipeline

    cfg = CaseConfig(Ne=4096, cone_split={"angle_scale": None})
    a = CasePipeline(cfg, arena_bytes=1 << 30)
    for b in range(4):
        a.step(100 + b, 4)
    la, ba = a.spectra()
    ca = dict(a.cone_stats)
    a.release()
    o = CasePipeline(cfg, arena_bytes=1 << 30)
    o.run_steps(100, 4, overlap=True)
    torch.cuda.synchronize()
    lo, bo = o.spectra()
    assert ca["copies"] > 0 and ca == {k: o.cone_stats[k] for k in ca}
    np.testing.assert_array_equal(la, lo)
    np.testing.assert_array_equal(ba, bo)


def test_flat_entry_cone_split_record_blocks_and_tally(device):
    """simulate_trajectories(_cone_split=ConeSplit) as the reference's
    transport_cone_split: the primaries' result carries the copy records and
    the copies' own transport; parents index the call's primaries (block-local
    under electron_start, and a block equals that slice of the single call);
    primary plus weighted copies keep the in-cone path tally of analog."""
    from pyrite_amd.cone_split import ConeSplit, history_sums
    from pyrite_amd.transport.api import simulate_trajectories

    comp = [("B", 0.0553), ("N", 0.0553)]
    kw = dict(composition=comp, E_cut_keV=1.0, crystal_width_mm=50.0, crystal_height_mm=50.0,
              keep_segments_on_device=False)
    req = ConeSplit(N_HAT, THETA_C, 0.05, max_copies_per_electron=1024, weight_ladder=True)
    full = simulate_trajectories(50.0, 3000, 5.0e4, seed=4, _cone_split=req, **kw)
    rec = full["cone_split"]
    assert rec["parent"].size > 50 and rec["copies"]["Ne"] == rec["parent"].size
    assert np.all(full["exit_code"][rec["killed"]] == 6)
    np.testing.assert_array_equal(rec["copies"]["initial_r_ang"], rec["r_ang"])
    assert np.all(rec["v_hat"] @ np.asarray(N_HAT) >= math.cos(THETA_C) - 1e-12)
    block = simulate_trajectories(50.0, 1000, 5.0e4, seed=4, _cone_split=req, electron_start=2000, **kw)
    b = block["cone_split"]
    sel = rec["parent"] >= 2000
    np.testing.assert_array_equal(b["parent"], rec["parent"][sel] - 2000)
    for name in ("weight", "stream_keys", "r_ang", "v_hat", "E_keV"):
        np.testing.assert_array_equal(b[name], rec[name][sel], err_msg=name)

    ma, mb, va, vb = [], [], [], []
    n = 10_000
    for seed in range(3):
        a = simulate_trajectories(50.0, n, 5.0e4, seed=50 + seed, **kw)
        s = simulate_trajectories(50.0, n, 5.0e4, seed=50 + seed, _cone_split=req, **kw)
        ta = np.bincount(a["electron_id"], weights=a["L_ang"] * (a["v_hat"] @ np.asarray(N_HAT) >= math.cos(THETA_C)),
                         minlength=n)
        tp = np.bincount(s["electron_id"], weights=s["L_ang"] * (s["v_hat"] @ np.asarray(N_HAT) >= math.cos(THETA_C)),
                         minlength=n)
        c = s["cone_split"]
        co = c["copies"]
        tc = np.bincount(co["electron_id"], weights=co["L_ang"] * (co["v_hat"] @ np.asarray(N_HAT) >= math.cos(THETA_C)),
                         minlength=c["parent"].size)
        tb = history_sums(tp, tc, c["parent"], c["weight"], n)
        ma.append(ta.mean())
        mb.append(tb.mean())
        va.append(ta.var(ddof=1) / n)
        vb.append(tb.var(ddof=1) / n)
    se = math.sqrt((sum(va) + sum(vb)) / 9.0)
    assert abs(np.mean(ma) - np.mean(mb)) < 4.0 * se, (np.mean(ma), np.mean(mb), se)
    assert sum(vb) < sum(va)
