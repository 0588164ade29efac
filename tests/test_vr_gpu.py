"""Learned variance reduction on the device (GPU).

* the weight window, child streams and line masses match the oracle (counts
  and keys bit for bit, weights to one rounding);
* a round long enough for every history reproduces the analog kernel's rows
  bit for bit (same step, same streams);
* suspending and resuming (rounds without splitting) and the learned
  splitting/roulette keep the analog expectation of weighted observables
  (transmitted and backscattered weight, rows per electron, peak tally)
  within 4 sigma;
* trained on analog tallies, the learned importance raises the peak-window
  figure of merit over the plain analog run.
"""

import ctypes as C
import dataclasses

import numpy as np
import pytest

from tests.transport_cases import SI_30KEV_STRAGGLE

pytestmark = pytest.mark.gpu


def _stream():
    import torch

    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def test_window_and_spawn_match_oracle(oracle, device):
    import torch

    from pyrite_amd import _lib

    rng = np.random.default_rng(0)
    n = 50_000
    ec = np.where(rng.random(n) < 0.8, 6, rng.integers(0, 6, n)).astype(np.int8)
    y = rng.normal(0.0, 2.0, n).astype(np.float32)
    w = np.exp(rng.uniform(np.log(1 / 64), 0.0, n))
    keys = rng.integers(0, 2**63, n, dtype=np.uint64)
    args = (0.7, 1 / 16, 1.0, 32)
    want_c, want_w = oracle.vr_window(ec, y, w, keys, *args)
    d = {k: torch.as_tensor(v, device=device) for k, v in (("ec", ec), ("y", y), ("w", w), ("k", keys))}
    count = torch.empty(n, dtype=torch.int32, device=device)
    wc = torch.empty(n, dtype=torch.float64, device=device)
    a = _lib.VrWindowArgs()
    a.n = n
    a.exit_code, a.log_importance, a.weight, a.key = (d[k].data_ptr() for k in ("ec", "y", "w", "k"))
    a.log_I_ref, a.w_min, a.w_max, a.n_max = args
    a.ratio_lo, a.ratio_hi = 1.0, 1.0
    a.count, a.w_child = count.data_ptr(), wc.data_ptr()
    _lib.check(_lib.load().hr_vr_window(C.byref(a), _stream()), "hr_vr_window")
    got_c = count.cpu().numpy()
    np.testing.assert_array_equal(got_c, want_c)
    np.testing.assert_allclose(wc.cpu().numpy(), want_w, rtol=1e-15, atol=0)

    # spawn: copies of every continuing particle with child streams
    total = int(got_c.sum())
    off = torch.cumsum(count.to(torch.int64), 0) - count.to(torch.int64)
    pos = torch.as_tensor(rng.normal(size=(n, 3)), device=device)
    dirs = torch.as_tensor(rng.normal(size=(n, 3)), device=device)
    E = torch.as_tensor(rng.uniform(1, 2, n), device=device)
    clock = torch.as_tensor(rng.uniform(0, 1, n), device=device)
    prim = torch.arange(n, dtype=torch.int64, device=device)
    out = {k: torch.empty(s, dtype=dt, device=device) for k, s, dt in (
        ("pos", (total, 3), torch.float64), ("dirs", (total, 3), torch.float64), ("E", total, torch.float64),
        ("clock", total, torch.float64), ("w", total, torch.float64), ("key", total, torch.uint64),
        ("prim", total, torch.int64))}
    sp = _lib.VrSpawnArgs()
    sp.n = n
    sp.count, sp.offset, sp.w_child = count.data_ptr(), off.data_ptr(), wc.data_ptr()
    sp.pos, sp.dirs, sp.E, sp.clock = pos.data_ptr(), dirs.data_ptr(), E.data_ptr(), clock.data_ptr()
    sp.key, sp.prim = d["k"].data_ptr(), prim.data_ptr()
    sp.out_pos, sp.out_dirs, sp.out_E = out["pos"].data_ptr(), out["dirs"].data_ptr(), out["E"].data_ptr()
    sp.out_clock, sp.out_weight = out["clock"].data_ptr(), out["w"].data_ptr()
    sp.out_key, sp.out_prim = out["key"].data_ptr(), out["prim"].data_ptr()
    _lib.check(_lib.load().hr_vr_spawn(C.byref(sp), _stream()), "hr_vr_spawn")
    parent = np.repeat(np.arange(n), got_c)
    j = np.arange(total) - np.repeat(off.cpu().numpy(), got_c)
    np.testing.assert_array_equal(out["prim"].cpu().numpy(), parent)
    np.testing.assert_array_equal(out["pos"].cpu().numpy(), pos.cpu().numpy()[parent])
    np.testing.assert_array_equal(out["w"].cpu().numpy(), wc.cpu().numpy()[parent])
    ck = out["key"].cpu().numpy()
    sample = rng.choice(total, 500, replace=False)
    for s in sample:
        assert int(ck[s]) == oracle.vr_child_key(int(keys[parent[s]]), int(j[s]))


def test_second_weight_and_line_roulette_match_oracle(oracle, device):
    import torch

    from pyrite_amd import _lib

    rng = np.random.default_rng(4)
    n = 20_000
    ec = np.where(rng.random(n) < 0.8, 6, rng.integers(0, 6, n)).astype(np.int8)
    y = rng.normal(0.0, 2.0, n).astype(np.float32)
    w = np.exp(rng.uniform(np.log(1 / 64), 0.0, n))
    w2 = rng.uniform(0.5, 2.0, n)
    keys = rng.integers(0, 2**63, n, dtype=np.uint64)
    want_c, want_w, want_w2 = oracle.vr_window(ec, y, w, keys, 0.7, 1 / 16, 1.0, 32, weight2=w2,
                                               ratio_lo=0.5, ratio_hi=2.0)
    d = {k: torch.as_tensor(v, device=device) for k, v in (("ec", ec), ("y", y), ("w", w), ("k", keys), ("w2", w2))}
    count = torch.empty(n, dtype=torch.int32, device=device)
    wc = torch.empty(n, dtype=torch.float64, device=device)
    w2c = torch.empty(n, dtype=torch.float64, device=device)
    a = _lib.VrWindowArgs()
    a.n = n
    a.exit_code, a.log_importance, a.weight, a.key = (d[k].data_ptr() for k in ("ec", "y", "w", "k"))
    a.log_I_ref, a.w_min, a.w_max, a.n_max = 0.7, 1 / 16, 1.0, 32
    a.ratio_lo, a.ratio_hi = 0.5, 2.0
    a.count, a.w_child = count.data_ptr(), wc.data_ptr()
    a.weight2, a.w2_child = d["w2"].data_ptr(), w2c.data_ptr()
    _lib.check(_lib.load().hr_vr_window(C.byref(a), _stream()), "hr_vr_window")
    np.testing.assert_array_equal(count.cpu().numpy(), want_c)
    np.testing.assert_allclose(wc.cpu().numpy(), want_w, rtol=1e-15, atol=0)
    np.testing.assert_allclose(w2c.cpu().numpy(), want_w2, rtol=1e-15, atol=0)
    # line roulette: survivors bit for bit
    n_seg, n_dir, n_g = 5000, 3, 2
    lw = rng.uniform(0.1, 1.0, n_seg * n_dir * n_g).astype(np.float32)
    mass = rng.uniform(0.0, 3.0, n_seg)
    want = oracle.line_roulette(lw, n_seg, n_dir, n_g, mass, 2.0, key=99, row0=1000)
    lw_d = torch.as_tensor(lw, device=device)
    m_d = torch.as_tensor(mass, device=device)
    _lib.call("hr_line_roulette", lw_d.data_ptr(), n_seg, n_dir, n_g, m_d.data_ptr(), 2.0, 99, 1000, _stream())
    np.testing.assert_array_equal(lw_d.cpu().numpy(), want)


def test_line_mass_matches_oracle(oracle, device):
    import torch

    from pyrite_amd import _lib

    rng = np.random.default_rng(1)
    n_seg, n_dir, n_g = 5000, 3, 2
    m = n_seg * n_dir * n_g
    E_r = rng.uniform(1000.0, 9000.0, m)
    a = rng.uniform(0.01, 10.0, m)
    w = rng.uniform(0.0, 1.0, m).astype(np.float32)
    want = oracle.line_mass(E_r, a, w, n_seg, n_dir, n_g, 3000.0, 4500.0, 47.9)
    t = [torch.as_tensor(v, device=device) for v in (E_r, a, w)]
    out = torch.empty(n_seg, dtype=torch.float64, device=device)
    _lib.call("hr_line_mass", t[0].data_ptr(), t[1].data_ptr(), t[2].data_ptr(), n_seg, n_dir, n_g,
              3000.0, 4500.0, 47.9, out.data_ptr(), _stream())
    np.testing.assert_allclose(out.cpu().numpy(), want, rtol=1e-14, atol=0)


def _round_run(case, device, rows, cap):
    """One hr_vr_round over a bank of the case's primaries."""
    import torch

    from pyrite_amd import _lib
    from pyrite_amd.transport.core import DeviceTables, TransportControl, alloc_rows, transport_args

    lt, lut = case.tables()
    pos, dirs, E, E_cut = case.state()
    Ne = case.Ne
    tabs = DeviceTables(lt, lut, device)
    ctl = TransportControl(energy_model=case.energy_model, straggle_on=case.straggle)
    keys = torch.empty(Ne, dtype=torch.uint64, device=device)
    _lib.call("hr_stream_keys", case.seed, 0, Ne, keys.data_ptr(), None)
    st = [torch.as_tensor(x, device=device).contiguous() for x in (pos, dirs, E, E_cut)]
    clock = torch.zeros(Ne, dtype=torch.float64, device=device)
    stragg = torch.zeros(Ne, dtype=torch.float64, device=device)
    rws = alloc_rows(Ne, cap, True, device)
    cnt = torch.zeros(Ne, dtype=torch.int64, device=device)
    code = torch.zeros(Ne, dtype=torch.int8, device=device)
    a = transport_args(tabs, ctl, (0, Ne, cap, keys), (None, clock, st[0], st[1], st[2], st[3]), rws,
                       (cnt, code), (case.straggle, stragg))
    _lib.check(_lib.load().hr_vr_round(C.byref(a), rows, None, _stream()), "hr_vr_round")
    torch.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in rws.items()}, cnt.cpu().numpy(), code.cpu().numpy()


def test_long_round_reproduces_analog_rows(device):
    from tests.test_transport_gpu import _gpu_run

    case = dataclasses.replace(SI_30KEV_STRAGGLE, Ne=200)
    cap = 4096
    rows_v, cnt_v, code_v = _round_run(case, device, cap, cap)
    rows_a, cnt_a, code_a, _ = _gpu_run(case, device, cap=cap)
    assert cnt_v.max() < cap
    np.testing.assert_array_equal(cnt_v, cnt_a)
    np.testing.assert_array_equal(code_v, code_a)
    keep = (np.arange(cap)[None, :] < cnt_a[:, None]).reshape(-1)
    for k in ("seg_mid", "seg_dir", "seg_len", "seg_E", "seg_E_end", "seg_t_end", "seg_event"):
        np.testing.assert_array_equal(rows_v[k][keep], rows_a[k][keep], err_msg=k)


def test_short_round_suspends(device):
    case = dataclasses.replace(SI_30KEV_STRAGGLE, Ne=300)
    rows_v, cnt_v, code_v = _round_run(case, device, 3, 8)
    assert cnt_v.max() <= 3
    assert np.all(cnt_v[code_v == 6] == 3)
    assert (code_v == 6).mean() > 0.5


@pytest.fixture(scope="module")
def small_case(device):
    from pyrite_amd.pipeline import CaseConfig, CasePipeline

    # the 5 MeV HOPG (002) geometry on a 20 um slab: same physics, short histories
    cfg = CaseConfig(Ne=4096, thickness_ang=2.0e5, n_side=2)
    return CasePipeline(cfg)


def _zs(a_vals, b_vals):
    a, b = np.asarray(a_vals, float), np.asarray(b_vals, float)
    se = np.sqrt(a.var(ddof=1) / a.size + b.var(ddof=1) / b.size)
    return abs(a.mean() - b.mean()) / max(se, 1e-300)


def _weighted_observables(vr, n_batches, Ne, use_vr, first):
    import torch

    out = []
    p = vr.pipe
    for b in range(n_batches):
        st = {}
        lines = torch.zeros(p.E_grid.size, dtype=torch.float64, device=p.dev)
        brem = torch.zeros_like(lines)
        vr.run_batch(first + b, Ne, Ne, lines, brem, use_vr=use_vr, stats=st)
        peak = vr.peak_tally(lines.cpu().numpy())
        out.append((st["rows"] / Ne, st.get("exit_2", 0.0) / Ne, st.get("exit_1", 0.0) / Ne, peak))
    return np.array(out)


def _analog_observables(pipe, vr, n_batches, first):
    import torch

    out = []
    for b in range(n_batches):
        res = pipe.transport(first + b)
        Ne = pipe.cfg.Ne
        code = res.exit_code.cpu().numpy()
        lines = torch.zeros(pipe.E_grid.size, dtype=torch.float64, device=pipe.dev)
        brem = torch.zeros_like(lines)
        from pyrite_amd.spectrum.lines import segments_fp32

        seg = segments_fp32({
            "r_mid": res.rows["seg_mid"], "v_hat": res.rows["seg_dir"], "L_ang": res.rows["seg_len"],
            "E_keV": res.rows["seg_E"], "E_repr_keV": 0.5 * (res.rows["seg_E"] + res.rows["seg_E_end"]),
        }, pipe.dev)
        pipe.lines.spectrum(seg, Ne, lines)
        pipe.brem.reduce(seg, brem)
        peak = vr.peak_tally(lines.cpu().numpy())
        wmass = float(vr.row_masses(seg, Ne).sum())
        out.append((res.n_segments / Ne, (code == 2).mean(), (code == 1).mean(), peak, wmass))
    return np.array(out)


def test_rounds_without_splitting_keep_analog_expectation(small_case):
    from pyrite_amd.vr.importance import LearnedImportance, VRConfig

    vr = LearnedImportance(small_case, VRConfig(rows_per_round=3))
    rnd = _weighted_observables(vr, 12, small_case.cfg.Ne, False, first=100)
    ana = _analog_observables(small_case, vr, 12, first=200)
    for col, name in enumerate(("rows", "transmitted", "backscattered", "peak")):
        if ana[:, col].std() == 0 and rnd[:, col].std() == 0:
            assert ana[0, col] == rnd[0, col]
            continue
        assert _zs(rnd[:, col], ana[:, col]) < 4.0, name


def test_learned_importance_unbiased(small_case):
    from pyrite_amd.vr.importance import LearnedImportance, VRConfig

    vr = LearnedImportance(small_case, VRConfig(train_rows=400_000, epochs=4))
    x, y = vr.training_set(4, first_batch=10_000)
    hist = vr.fit(x, y)
    assert hist[-1] < hist[0]
    Ne = small_case.cfg.Ne
    for roulette in (False, True):
        lrn = []
        for b in range(12):
            import torch

            st = {}
            lines = torch.zeros(small_case.E_grid.size, dtype=torch.float64, device=small_case.dev)
            brem = torch.zeros_like(lines)
            vr.run_batch(300 + b, Ne, Ne, lines, brem, use_vr=True, stats=st, roulette=roulette)
            lrn.append((st["rows"] / Ne, st.get("exit_2", 0.0) / Ne, st.get("exit_1", 0.0) / Ne,
                        vr.peak_tally(lines.cpu().numpy()), st["wmass"], st["rows2"] / Ne))
        lrn = np.array(lrn)
        ana = _analog_observables(small_case, vr, 12, first=400)
        for col, name in enumerate(("rows", "transmitted", "backscattered", "peak", "window mass")):
            if ana[:, col].std() == 0 and lrn[:, col].std() == 0:
                continue
            assert _zs(lrn[:, col], ana[:, col]) < 4.0, (name, roulette)
        # the brem weight follows one copy per primary: its weighted rows match the analog rows
        assert _zs(lrn[:, 5], ana[:, 0]) < 4.0


@pytest.fixture(scope="module")
def hopg_case(device):
    from pyrite_amd.pipeline import CaseConfig, CasePipeline

    return CasePipeline(CaseConfig(Ne=16384))


def test_overlapped_rounds_give_the_one_stream_tallies(small_case):
    """Scoring each round on a second stream beside the next round's transport
    keeps the scoring order: line and brem tallies equal the one-stream
    schedule's bit for bit."""
    import torch

    from pyrite_amd.vr.importance import LearnedImportance, VRConfig

    vr = LearnedImportance(small_case, VRConfig(train_rows=200_000, epochs=2))
    x, y = vr.training_set(2, first_batch=30_000)
    vr.fit(x, y)
    Ne, p = small_case.cfg.Ne, small_case
    out = []
    for overlap in (False, True):
        vr.overlap = overlap
        lines = torch.zeros(p.E_grid.size, dtype=torch.float64, device=p.dev)
        brem = torch.zeros_like(lines)
        for b in range(3):
            vr.run_batch(40_000 + b, Ne, 3 * Ne, lines, brem, use_vr=True, roulette=True)
        out.append((lines.cpu().numpy(), brem.cpu().numpy()))
    vr.overlap = True
    assert out[0][0].sum() > 0 and out[0][1].sum() > 0
    np.testing.assert_array_equal(out[1][0], out[0][0])
    np.testing.assert_array_equal(out[1][1], out[0][1])


def _vr_spectra(vr, n_batches, Ne, mode, first):
    import torch

    p = vr.pipe
    lines_all, brem_all, wm = [], [], []
    for b in range(n_batches):
        lines = torch.zeros(p.E_grid.size, dtype=torch.float64, device=p.dev)
        brem = torch.zeros_like(lines)
        st = {}
        vr.run_batch(first + b, Ne, Ne, lines, brem, use_vr=(mode == "vr"), stats=st, roulette=(mode == "vr"))
        lines_all.append(lines.cpu().numpy())
        brem_all.append(brem.cpu().numpy() / Ne)
        wm.append(st["wmass"])
    return np.array(lines_all), np.array(brem_all), np.array(wm)


def _per_bin_agree(a, b, name, p_trip=0.0027):
    from scipy.stats import binom

    se = np.sqrt(a.var(0, ddof=1) / len(a) + b.var(0, ddof=1) / len(b))
    live = se > 0
    z = (a.mean(0) - b.mean(0))[live] / se[live]
    n = int(live.sum())
    trips = int((np.abs(z) > 3.0).sum())
    assert n >= 50 and trips <= int(binom.ppf(1.0 - 1e-4, n, p_trip)), (name, trips, n)
    ta, tb = a.sum(1), b.sum(1)
    zt = (ta.mean() - tb.mean()) / np.sqrt(ta.var(ddof=1) / len(ta) + tb.var(ddof=1) / len(tb))
    assert abs(zt) < 4.0, (name, zt)


def test_learned_importance_is_unbiased_per_bin_and_gains_on_the_smooth_observable(hopg_case):
    """Config 5 geometry (100 um HOPG, 5 MeV): learned splitting/roulette plus
    line roulette against the analog round run, per bin of the line and brem
    spectra (3-sigma trips within the binomial count, integrals within 4
    sigma), over 48 disjoint batches of 16,384 histories each.  The variance
    gain is checked on the pooled window line mass, the smooth observable of
    the peak (the peak tally itself is too heavy-tailed for 48 batches; its
    figure of merit is the bench's)."""
    from pyrite_amd.vr.importance import LearnedImportance, VRConfig

    vr = LearnedImportance(hopg_case, VRConfig(train_rows=600_000, epochs=5))
    x, y = vr.training_set(6, first_batch=10_000)
    vr.fit(x, y)
    Ne, nb = hopg_case.cfg.Ne, 48
    lv, bv, wv = _vr_spectra(vr, nb, Ne, "vr", first=50_000)
    la, ba, wa = _vr_spectra(vr, nb, Ne, "rounds", first=60_000)
    assert la.sum() > 0 and lv.sum() > 0
    _per_bin_agree(bv, ba, "brem")
    _per_bin_agree(lv, la, "lines")
    var_ratio = wv.var(ddof=1) / wa.var(ddof=1)
    print(f"window-mass variance per history, learned/analog: {var_ratio:.3f}")
    assert abs(wv.mean() - wa.mean()) < 4.0 * np.sqrt(wv.var(ddof=1) / nb + wa.var(ddof=1) / nb)
    assert var_ratio < 0.8
