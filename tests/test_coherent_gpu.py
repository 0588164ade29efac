"""The coherent line route on the device against the oracle (GPU).

* records: the device prologue on real transport rows equals the oracle's
  FP64 restatement (cxr/_batched.py:561-690) line by line;
* field: the device sum over the same records equals the oracle's FP64 field
  (cxr/_batched_finalize.py:201-235) within 2e-6 of the sum's L1 size per
  bin (FP32 profile and phase rotation, FP64 phase reduction and sums);
* the worked two-energy block (jit-spectrum-kernel-walkthrough.md:173-211)
  through the device field and finalize gives 1.756606;
* a straight segment and its collinear pieces give one spectrum;
* passes and streamed rounds change only the rounding (the FP32 groups of 8
  lines fall differently);
* the HOPG case's per-batch coherent spectra agree with the oracle's per bin
  (binomial count of 3-sigma trips and a 4-sigma integral), as the incoherent
  route does.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

F = ("E_c", "a", "phase_slope", "g_phase", "cs_re", "cs_im", "cp_re", "cp_im", "L_esc", "half_dL", "apb",
     "bma", "q")


def _host_rows(res):
    """Compacted FP64 host rows of a compact=False transport result."""
    out = {k: [] for k in ("seg_mid", "seg_dir", "seg_len", "seg_E", "seg_E_end", "seg_t0")}
    for _e0, m, cap, rows, cnt in res.slots:
        c = np.minimum(cnt[:m].cpu().numpy(), cap)
        keep = (np.arange(cap)[None, :] < c[:, None]).reshape(-1)
        for k in out:
            v = rows[k][: m * cap].cpu().numpy()
            v = v.reshape(m * cap, -1)[keep]
            out[k].append(v.reshape(-1, 3) if k in ("seg_mid", "seg_dir") else v.reshape(-1))
    return {k: np.concatenate(v) for k, v in out.items()}


@pytest.fixture(scope="module")
def hopg_batch(device):
    from pyrite_amd.pipeline import CaseConfig, CasePipeline

    pipe = CasePipeline(CaseConfig(Ne=400, coherent=True))
    res = pipe.transport(3, compact=False)
    return pipe, res


def test_device_records_equal_the_oracle(hopg_batch):
    from oracle import oracle as O
    from pyrite_amd.spectrum.crystal import polarization_pairs

    pipe, res = hopg_batch
    sc = pipe.coherent
    (_e0, m, cap, rows, cnt), = res.slots
    rec = sc.records(rows, cnt, m, cap).cpu().numpy()  # [13, n_dir*n_g, n_lines]
    h = _host_rows(res)
    seg = {"r_mid": h["seg_mid"], "v_hat": h["seg_dir"], "L": h["seg_len"],
           "E": 0.5 * (h["seg_E"] + h["seg_E_end"]), "t0": h["seg_t0"]}
    tab = pipe.line_tables.as_dict()
    n_g = len(tab["wm"])
    assert rec.shape[2] == len(seg["L"]) > 1000
    n_good = 0
    for d, n_hat in enumerate(pipe.n_hats):
        es, ep = polarization_pairs(n_hat, tab["g"])
        ref, good = O.coherent_lines(seg, pipe.cfg.thickness_ang, n_hat, tab, es, ep)
        for g in range(n_g):
            dev = rec[:, d * n_g + g, :].T  # [n_lines, 13]
            r = ref[:, g, :]
            dgood = np.any(dev[:, 4:8] != 0.0, axis=1)
            assert np.mean(dgood != good[:, g]) < 1e-3, (d, g)
            both = dgood & good[:, g]
            n_good += int(both.sum())
            if not both.any():  # grazing rays: the in-medium root does not settle (as on the incoherent route)
                continue
            for f in range(13):
                x, y = dev[both, f], r[both, f]
                scale = np.maximum(np.abs(y), 1e-300)
                if F[f] in ("g_phase", "phase_slope"):  # large phases: absolute in radians per eV / radians
                    tol = 1e-9 * np.max(np.abs(y)) + 1e-12
                    assert np.max(np.abs(x - y)) <= tol, (F[f], np.max(np.abs(x - y)))
                elif F[f] in ("half_dL", "bma", "q"):  # may cross zero
                    tol = 1e-9 * np.max(np.abs(y)) + 1e-300
                    assert np.max(np.abs(x - y)) <= tol, (F[f], np.max(np.abs(x - y)))
                else:
                    assert np.max(np.abs(x - y) / scale) < 1e-9, (F[f], np.max(np.abs(x - y) / scale))
    assert n_good > 0.3 * rec.shape[1] * rec.shape[2]


def test_device_field_equals_the_oracle_field(hopg_batch):
    import torch

    from oracle import oracle as O

    pipe, res = hopg_batch
    sc = pipe.coherent
    (_e0, m, cap, rows, cnt), = res.slots
    rec_d = sc.records(rows, cnt, m, cap).contiguous()
    n_lines = rec_d.shape[2]
    sc.reset()
    sc._launch_field(rec_d, n_lines)
    field = sc.field.cpu().numpy()
    rec = rec_d.cpu().numpy()
    dom = sc.dom.cpu().numpy()
    live = 0
    for row in range(sc.n_rows):
        bound = np.zeros(sc.n_E)
        ref = O.coherent_field(rec[:, row, :].T, pipe.E_grid, dom, bound=bound)
        live += bound.max() > 0
        err = np.abs(field[row] - ref).max(0)
        assert np.all(err <= 2e-6 * bound + 1e-300), (row, np.max(err / np.maximum(bound, 1e-300)))
    assert live >= sc.n_rows // 2
    # the finalize squares the same planes
    out = sc.spectrum(Ne=m).cpu().numpy()
    w = pipe.weights
    wm = pipe.line_tables.wm
    expect = sum(w[r // len(wm)] * wm[r % len(wm)] * (field[r] ** 2).sum(0) for r in range(sc.n_rows)) / m
    np.testing.assert_allclose(out, expect, rtol=1e-12, atol=1e-300)
    del torch


def test_worked_two_energy_block_on_the_device(device):
    import ctypes as C

    import torch

    from pyrite_amd import _lib

    dev = device
    rec = np.zeros((13, 1, 2))
    rec[1] = np.pi / 2
    rec[4], rec[6], rec[10] = 1.0, 2.0, 2.0
    rec[0, 0, 1], rec[3, 0, 1] = 1.0, -np.pi / 2
    rec_d = torch.as_tensor(rec, device=dev).contiguous()
    E = torch.tensor([0.0, 1.0], dtype=torch.float64, device=dev)
    dom = torch.zeros(2, dtype=torch.float64, device=dev)
    field = torch.zeros((1, 4, 2), dtype=torch.float64, device=dev)
    f = _lib.CoherentFieldArgs()
    f.rec, f.ld, f.n_lines, f.n_rows, f.n_E = rec_d.data_ptr(), 2, 2, 1, 2
    f.E_grid, f.dom, f.field = E.data_ptr(), dom.data_ptr(), field.data_ptr()
    f.ws = _lib.DeviceScratch(dev).get(_lib.workspace_bytes("hr_coherent_field_workspace_bytes", C.byref(f)))
    _lib.call("hr_coherent_field", C.byref(f), None)
    out = torch.zeros(2, dtype=torch.float64, device=dev)
    one = torch.ones(1, dtype=torch.float64, device=dev)
    quarter = torch.full((1,), 0.25, dtype=torch.float64, device=dev)
    _lib.call("hr_coherent_finalize", field.data_ptr(), 1, 1, one.data_ptr(), quarter.data_ptr(), 2, 1.0,
              out.data_ptr(), None)
    torch.cuda.synchronize()
    np.testing.assert_allclose(out.cpu().numpy(), 0.25 * (5.0 + 20.0 / np.pi**2), rtol=2e-7)


def test_straight_segment_and_its_pieces_give_one_spectrum(device):
    from pyrite_amd.pipeline import CaseTables, CaseConfig
    from pyrite_amd.spectrum.coherent import mc_cxr_coherent_spectrum

    t = CaseTables(CaseConfig(Ne=10))
    rng = np.random.default_rng(4)
    specs = []
    v = t.beam_dir + 0.05 * rng.normal(size=3)
    v /= np.linalg.norm(v)
    Ek, L, t0 = 3000.0, 2400.0, 5.0e4
    beta = np.sqrt(1.0 - 1.0 / (1.0 + Ek / 510.99895) ** 2)
    start = np.array([10.0, -4.0, 3.0e5])
    for ns in (1, 6):
        Ls = L / ns
        seg = {"r_mid": np.array([start + (i + 0.5) * Ls * v for i in range(ns)]), "v_hat": np.tile(v, (ns, 1)),
               "L_ang": np.full(ns, Ls), "E_keV": np.full(ns, Ek), "t_ang": t0 + np.arange(ns) * Ls / beta,
               "thickness_ang": 1.0e6, "Ne": 1}
        specs.append(mc_cxr_coherent_spectrum(seg, t.E_grid, "hopg", [(0, 0, 2)], n_hats=t.n_hats,
                                              weights=t.weights))
    assert specs[0].max() > 0
    np.testing.assert_allclose(specs[1], specs[0], rtol=0, atol=2e-5 * specs[0].max())


def test_passes_change_only_the_summation_order(hopg_batch, monkeypatch):
    from pyrite_amd.spectrum import coherent

    pipe, res = hopg_batch
    sc = pipe.coherent
    sc.reset()
    sc.add_transport(res, pipe.cfg.Ne)
    one = sc.field.cpu().numpy()
    monkeypatch.setattr(coherent, "RECORD_BYTES", 13 * sc.n_rows * 8 * 3001)
    sc.reset()
    sc.add_transport(res, pipe.cfg.Ne)
    many = sc.field.cpu().numpy()
    scale = np.abs(one).max()
    assert scale > 0
    # pass boundaries move the FP32 groups of 8 lines: rounding within the
    # parity bound (2e-6 of a bin's L1 size, which exceeds the field itself)
    np.testing.assert_allclose(many, one, rtol=0, atol=3e-5 * scale)


def test_streamed_rounds_give_the_batched_fields(device):
    """Continuation rounds (96 lanes, 8 rows per launch, refills and the packed
    tail) feed the same rows, so the fields differ from one batch's only in
    rounding."""
    from pyrite_amd.pipeline import CaseConfig, CasePipeline

    cfg = CaseConfig(Ne=500, coherent=True, E0_keV=2000.0, thickness_ang=2.0e5)
    a = CasePipeline(cfg)
    a.step(0, 1)
    fa = a.coherent.field.cpu().numpy()
    la, ba = a.spectra()
    b = CasePipeline(cfg)
    res = b.stream(0, 500, lanes=96, rows_per_round=8)
    assert res.rounds > 3
    fb = b.coherent.field.cpu().numpy()
    lb, bb = b.spectra()
    scale = np.abs(fa).max()
    assert scale > 0 and b.coherent.n_electrons == a.coherent.n_electrons == 500
    np.testing.assert_allclose(fb, fa, rtol=0, atol=3e-5 * scale)  # FP32 line groups fall differently
    np.testing.assert_allclose(lb, la, rtol=0, atol=1e-4 * la.max())
    np.testing.assert_allclose(bb, ba, rtol=1e-7, atol=1e-12 * ba.max())  # brem sums 8 rows in FP32


def test_overlapped_steps_give_the_sequential_fields(device):
    """Transport of batch s+1 beside the coherent scoring of batch s adds the
    same fields in the same order: bit for bit."""
    from pyrite_amd.pipeline import CaseConfig, CasePipeline

    cfg = CaseConfig(Ne=3000, coherent=True, E0_keV=3000.0, thickness_ang=3.0e5)
    # equal arenas: equal sub-batches, so equal line groups in the FP32 field cells
    a = CasePipeline(cfg, arena_bytes=2 << 30)
    for b in range(3):
        a.step(b, 3)
    b_ = CasePipeline(cfg, arena_bytes=2 << 30)
    b_.run_steps(0, 3, overlap=True)
    np.testing.assert_array_equal(b_.coherent.field.cpu().numpy(), a.coherent.field.cpu().numpy())
    la, ba = a.spectra()
    lb, bb = b_.spectra()
    assert la.sum() > 0
    np.testing.assert_array_equal(lb, la)
    np.testing.assert_array_equal(bb, ba)


BATCHES, PER_BATCH = 32, 300


def test_hopg_coherent_case_agrees_with_the_oracle_per_bin(device):
    from oracle import case as ocase
    from pyrite_amd.pipeline import CaseConfig, CasePipeline
    from test_case_gpu import _agree

    pipe = CasePipeline(CaseConfig(Ne=PER_BATCH, coherent=True))
    g_lines, c_lines = [], []
    for b in range(BATCHES):
        pipe.reset_coherent()
        pipe.spec_brem.zero_()
        pipe.step(b, 1)
        g_lines.append(pipe.spectra()[0])
        ln, _br, _ = ocase.run_case(pipe, 30_000_000 + b * PER_BATCH, PER_BATCH, coherent=True)
        c_lines.append(ln / PER_BATCH)
    g_lines, c_lines = np.array(g_lines), np.array(c_lines)
    assert g_lines.sum() > 0 and c_lines.sum() > 0
    _agree(g_lines, c_lines, "coherent lines")


def _grouped_scorer(pipe):
    from pyrite_amd.spectrum.coherent import CoherentScorer

    return CoherentScorer(pipe.line_tables, pipe.cfg.crystal, pipe.E_grid, pipe.cfg.thickness_ang, pipe.n_hats,
                          pipe.weights, pipe.coherent.device, grouped=True)


def test_device_grouped_term_equals_the_oracle(hopg_batch):
    """Each electron's field squared (hr_coherent_grouped) against the FP64
    oracle, within the field's parity bound carried through the square."""
    import torch

    from oracle import oracle as O

    pipe, res = hopg_batch
    sc = _grouped_scorer(pipe)
    (_e0, m, cap, rows, cnt), = res.slots
    rec_d = sc.records(rows, cnt, m, cap).contiguous()
    n_lines = rec_d.shape[2]
    c = cnt[:m].clamp(max=cap)
    gs = torch.cat((torch.zeros(1, dtype=torch.int64, device=c.device), torch.cumsum(c, 0)))
    print(f"grouped: {m} electrons, {n_lines} lines, {sc.n_rows} rows", flush=True)
    sc._launch_grouped(rec_d, n_lines, gs)
    G = sc.grouped.cpu().numpy()
    rec = rec_d.cpu().numpy()
    dom = sc.dom.cpu().numpy()
    gs_h = gs.cpu().numpy()
    live = 0
    for row in range(sc.n_rows):
        ref, size = O.coherent_grouped(rec[:, row, :].T, gs_h, pipe.E_grid, dom)
        live += ref.max() > 0
        err = np.abs(G[row] - ref)
        assert np.all(err <= 4 * 2e-6 * size + 1e-300), (row, np.max(err / np.maximum(size, 1e-300)))
    assert live >= sc.n_rows // 2


def test_blend_limits_and_the_physical_estimator_on_the_device(hopg_batch):
    """factor 1 (sampled) is the flat spectrum, factor 0 the grouped one, and a
    physical population follows G + (N-1)/(M-1) F (P - G) of the same planes."""
    from pyrite_amd.spectrum.coherent import gaussian_longitudinal_factor, pair_scale

    pipe, res = hopg_batch
    sc = _grouped_scorer(pipe)
    sc.add_transport(res, pipe.cfg.Ne)
    n_E = sc.n_E
    flat = sc.spectrum().cpu().numpy()
    np.testing.assert_allclose(sc.spectrum(factor=np.ones(n_E)).cpu().numpy(), flat, rtol=1e-12, atol=1e-300)
    field = sc.field.cpu().numpy()
    G = sc.grouped.cpu().numpy()
    w, wm = pipe.weights, pipe.line_tables.wm
    P = (field ** 2).sum(1)
    grouped_only = sum(w[r // len(wm)] * wm[r % len(wm)] * G[r] for r in range(sc.n_rows)) / pipe.cfg.Ne
    np.testing.assert_allclose(sc.spectrum(factor=np.zeros(n_E)).cpu().numpy(), grouped_only, rtol=1e-12,
                               atol=1e-300)
    assert np.all(G >= 0.0) and grouped_only.max() > 0
    F = gaussian_longitudinal_factor(pipe.E_grid, 2e-4)
    assert 0.0 < F.min() < F.max() < 1.0
    pair = pair_scale(6.2e6, pipe.cfg.Ne)
    got = sc.spectrum(factor=F, pair_scale=pair).cpu().numpy()
    want = sum(w[r // len(wm)] * wm[r % len(wm)] * (G[r] + pair * F * (P[r] - G[r]))
               for r in range(sc.n_rows)) / pipe.cfg.Ne
    np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-12 * np.abs(want).max())
    sampled = sc.spectrum(factor=F).cpu().numpy()
    want = sum(w[r // len(wm)] * wm[r % len(wm)] * ((1 - F) * G[r] + F * P[r])
               for r in range(sc.n_rows)) / pipe.cfg.Ne
    np.testing.assert_allclose(sampled, want, rtol=1e-12, atol=1e-300)


def test_grouped_passes_split_only_between_electrons(hopg_batch, monkeypatch):
    """Record passes end between electrons, so the grouped term of many passes
    is the one-pass term up to summation order; grouping per-segment rows by
    electron id gives the slot rows' grouping."""
    from pyrite_amd.spectrum import coherent

    pipe, res = hopg_batch
    sc = _grouped_scorer(pipe)
    sc.add_transport(res, pipe.cfg.Ne)
    one = sc.grouped.cpu().numpy()
    scale = one.max()
    assert scale > 0
    monkeypatch.setattr(coherent, "RECORD_BYTES", 13 * sc.n_rows * 8 * 3001)
    sc.reset()
    sc.add_transport(res, pipe.cfg.Ne)
    # each electron's FP32 line groups are the same in every pass: only the
    # chunk partials' summation order moves
    np.testing.assert_allclose(sc.grouped.cpu().numpy(), one, rtol=0, atol=1e-10 * scale)
    # the same lines as one row per segment, electrons given by id
    (_e0, m, cap, rows, cnt), = res.slots
    h = _host_rows(res)
    c = np.minimum(cnt[:m].cpu().numpy(), cap)
    elec = np.repeat(np.arange(m), c)
    seg = {"r_mid": h["seg_mid"], "v_hat": h["seg_dir"], "L_ang": h["seg_len"], "E_keV": h["seg_E"],
           "t_ang": h["seg_t0"]}
    srows, scnt, n = coherent.segment_rows(seg, sc.device)
    srows["seg_E_end"] = srows["seg_E"].new_tensor(h["seg_E_end"])
    sc.reset()
    sc.add_rows(srows, scnt, n, 1, electron=elec)
    np.testing.assert_allclose(sc.grouped.cpu().numpy(), one, rtol=0, atol=1e-10 * scale)
    sc.reset()
    sc.add_rows(srows, scnt, n, 1)  # each segment its own electron: the incoherent limit
    assert np.max(np.abs(sc.grouped.cpu().numpy() - one)) > 1e-3 * scale


def test_stream_route_refuses_the_grouped_term(device):
    from pyrite_amd.pipeline import CaseConfig, CasePipeline

    pipe = CasePipeline(CaseConfig(Ne=64, coherent=True, coherent_grouped=True))
    with pytest.raises(ValueError, match="continuation rounds"):
        pipe.stream(0, 64)
