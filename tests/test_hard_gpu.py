"""Hard-event modes on the device against the FP64 oracle (GPU).

* radiative LUT columns (hr_radiative_lut) and the photon sampler
  (hr_radiative_sample) against the oracle's restatement, rtol 1e-12 (the
  cell integrals take device and host log/atan/log1p, an ulp apart);
* transport with the shell and radiative modes: the first steps of every
  history row for row (rtol 1e-12; the mode events, their losses, channels and
  photons included), full histories within 4 sigma of the oracle in their
  observables and mode-event counts, and every device history's energy
  balance exact.
"""

import ctypes as C
import math

import numpy as np
import pytest

from pyrite_amd.transport import hard_events as he
from pyrite_amd.transport.tables import build_layer_tables, build_transport_lut

pytestmark = pytest.mark.gpu

GE = [("Ge", 0.0442)]
W_SI = [("W", 0.02), ("Si", 0.03)]


def _tables(comp, E0, thick, rad_on, shell_on, rad_cut=5000.0, elsepa_atomic=False):
    if elsepa_atomic:  # ELSEPA tables carrying the shell mode's energy-dependent atomic-electron factor
        from tests.transport_cases import elsepa_test_tables

        layers = [(0.0, thick, list(comp))]
        lt = build_layer_tables(layers, "elsepa", atomic_electron_deflection="kawrakow", hard_cutoff_eV=1000.0,
                                elastic_tables=elsepa_test_tables(layers, np.geomspace(4.0, 1.2 * E0, 40)))
    else:
        lt = build_layer_tables([(0.0, thick, list(comp))])
    lut0 = build_transport_lut(5.0, E0, lt)
    shell = he.build_shell_tables(lt, lut0.E_grid, 1000.0) if shell_on else None
    rad = he.build_radiative_tables(lt, rad_cut) if rad_on else None
    return lt, lut0, he.lut_for(lut0, shell), rad, shell


def test_radiative_lut_and_sampler_match_oracle(oracle, device):
    import torch

    from pyrite_amd import _lib

    lt, lut0, lut, rad, _ = _tables(W_SI, 30000.0, 1.0e6, True, False, rad_cut=2000.0)
    modes = he.DeviceHardModes(lt, lut0, device, rad=rad)
    soft_d, rate_d = (x.cpu().numpy() for x in modes.radiative_columns())
    ro = oracle.RadOracle(rad)
    soft_o, rate_o = ro.layer_lut(lt, lut.E_grid)
    assert rate_o.min() > 0 and soft_o.min() > 0
    np.testing.assert_allclose(soft_d, soft_o, rtol=1e-12)
    np.testing.assert_allclose(rate_d, rate_o, rtol=1e-12)
    rng = np.random.default_rng(4)
    n = 2000
    E = np.exp(rng.uniform(np.log(6e3), np.log(3e7), n))
    u = rng.random(n)
    for slot in (0, 1):
        k = torch.empty(n, dtype=torch.float64, device=device)
        E_d = torch.as_tensor(E, device=device)
        u_d = torch.as_tensor(u, device=device)
        _lib.check(_lib.load().hr_radiative_sample(C.byref(modes.record), slot, float(rad.plasma[0]), n,
                                                   E_d.data_ptr(), u_d.data_ptr(), k.data_ptr(), None),
                   "hr_radiative_sample")
        want = np.array([ro.sample(slot, E[i], u[i], plasma_eV=rad.plasma[0]) for i in range(n)])
        np.testing.assert_allclose(k.cpu().numpy(), want, rtol=1e-12)


def _gpu(lt, lut0, lut, rad, shell, device, E0, n, seed, straggle, cap, max_steps=400_000):
    import torch

    from pyrite_amd import _lib
    from pyrite_amd.transport.core import DeviceTables, TransportControl, alloc_rows, make_hip_transport_core

    modes = he.DeviceHardModes(lt, lut0, device, rad=rad, shell=shell)
    tabs = DeviceTables(lt, lut, device)
    ctl = TransportControl(max_steps=max_steps, energy_model=1, straggle_on=straggle)
    core = make_hip_transport_core(tabs, ctl, modes)
    modes.set_keys(seed, 0, n)
    keys = torch.empty(n, dtype=torch.uint64, device=device)
    _lib.call("hr_stream_keys", seed, 0, n, keys.data_ptr(), None)
    pos = torch.zeros((n, 3), dtype=torch.float64, device=device)
    dirs = torch.zeros((n, 3), dtype=torch.float64, device=device)
    dirs[:, 2] = 1.0
    E = torch.full((n,), E0, dtype=torch.float64, device=device)
    clock = torch.zeros(n, dtype=torch.float64, device=device)
    alive = torch.ones(n, dtype=torch.uint8, device=device)
    Ecut = torch.full((n,), 5.0, dtype=torch.float64, device=device)
    stragg = torch.zeros(n, dtype=torch.float64, device=device)
    rows = alloc_rows(n, cap, True, device, modes.row_columns)
    cnt = torch.zeros(n, dtype=torch.int64, device=device)
    code = torch.zeros(n, dtype=torch.int8, device=device)
    core((0, n, cap, keys), (alive, clock, pos, dirs, E, Ecut), rows, (cnt, code), (straggle, stragg))
    torch.cuda.synchronize()
    host = {k: v.cpu().numpy() for k, v in rows.items()}
    st = {"pos": pos.cpu().numpy(), "dirs": dirs.cpu().numpy(), "E_keV": E.cpu().numpy(),
          "stragg_dE": stragg.cpu().numpy()}
    return host, cnt.cpu().numpy(), code.cpu().numpy(), st, modes


def _oracle(oracle, lt, lut, rad, shell, modes, E0, n, seed, straggle, cap, max_steps=400_000):
    radarg = None
    if rad is not None:
        # the oracle runs on its own moment columns (pinned to the device's above)
        soft, rate = oracle.RadOracle(rad).layer_lut(lt, lut.E_grid)
        radarg = (rad, oracle.stream_keys(seed ^ he.RADIATIVE_STREAM_SALT, n), soft, rate)
    return oracle.run_transport(lt, lut, oracle.stream_keys(seed, n), np.zeros((n, 3)),
                                np.tile([0.0, 0.0, 1.0], (n, 1)), np.full(n, E0), np.full(n, 5.0), cap=cap,
                                max_steps=max_steps, straggle_on=straggle, rad=radarg, shell=shell)


CASES = [
    ("ge_shell_rad", GE, 5000.0, 2.0e6, True, True, True),
    ("ge_shell", GE, 3000.0, 1.0e6, False, True, False),
    ("wsi_rad", W_SI, 20000.0, 1.0e6, True, False, True),
]


def test_shell_mode_on_elsepa_tables_with_atomic_deflection_matches_oracle(oracle, device):
    """Shell mode, tabulated angles and the energy-dependent 1 + xi(E)/Z on the
    table rates together: first steps row for row against the oracle."""
    lt, lut0, lut, rad, shell = _tables(GE, 3000.0, 1.0e6, False, True, elsepa_atomic=True)
    assert lt.elsepa is not None
    n, seed, steps = 20000, 12, 4
    g_rows, g_cnt, g_code, g_st, modes = _gpu(lt, lut0, lut, rad, shell, device, 3000.0, n, seed, False, steps,
                                              max_steps=steps)
    o_rows, o_cnt, o_code, o_st = _oracle(oracle, lt, lut, rad, shell, modes, 3000.0, n, seed, False, steps,
                                          max_steps=steps)
    np.testing.assert_array_equal(g_cnt, o_cnt)
    np.testing.assert_array_equal(g_code, o_code)
    used = (np.arange(steps)[None, :] < g_cnt[:, None]).reshape(-1)
    np.testing.assert_array_equal(g_rows["seg_event"][used], o_rows["seg_event"][used])
    assert (o_rows["seg_event"][used] == 8).sum() > 100
    for k in ("seg_len", "seg_E", "seg_E_end", "seg_t_end"):
        np.testing.assert_allclose(g_rows[k][used], o_rows[k][used], rtol=1e-12, atol=1e-300)
    for k in ("E_keV", "dirs"):
        np.testing.assert_allclose(g_st[k], o_st[k], rtol=1e-11, atol=1e-12)


@pytest.mark.parametrize("name,comp,E0,thick,rad_on,shell_on,straggle", CASES, ids=[c[0] for c in CASES])
def test_first_steps_match_oracle(oracle, device, name, comp, E0, thick, rad_on, shell_on, straggle):
    lt, lut0, lut, rad, shell = _tables(comp, E0, thick, rad_on, shell_on)
    n, seed, steps = 20000, 9, 4
    g_rows, g_cnt, g_code, g_st, modes = _gpu(lt, lut0, lut, rad, shell, device, E0, n, seed, straggle, steps,
                                              max_steps=steps)
    o_rows, o_cnt, o_code, o_st = _oracle(oracle, lt, lut, rad, shell, modes, E0, n, seed, straggle, steps,
                                          max_steps=steps)
    np.testing.assert_array_equal(g_cnt, o_cnt)
    np.testing.assert_array_equal(g_code, o_code)
    used = (np.arange(steps)[None, :] < g_cnt[:, None]).reshape(-1)
    ev = o_rows["seg_event"][used]
    np.testing.assert_array_equal(g_rows["seg_event"][used], ev)
    if shell_on:
        assert (ev == 8).sum() > 100
        np.testing.assert_array_equal(g_rows["seg_hard_ch"][used], o_rows["seg_hard_ch"][used])
        np.testing.assert_allclose(g_rows["seg_hard_W"][used], o_rows["seg_hard_W"][used], rtol=1e-12)
    if rad_on:
        assert (ev == 9).sum() > 3
        np.testing.assert_array_equal(g_rows["seg_rad_Z"][used], o_rows["seg_rad_Z"][used])
        np.testing.assert_allclose(g_rows["seg_rad_k"][used], o_rows["seg_rad_k"][used], rtol=1e-12)
    for k in ("seg_len", "seg_E", "seg_E_end", "seg_t_end"):
        np.testing.assert_allclose(g_rows[k][used], o_rows[k][used], rtol=1e-12, atol=1e-300)
    for k in ("E_keV", "dirs"):
        np.testing.assert_allclose(g_st[k], o_st[k], rtol=1e-11, atol=1e-12)


@pytest.mark.parametrize("name,comp,E0,thick,rad_on,shell_on,straggle", CASES, ids=[c[0] for c in CASES])
def test_full_histories_agree_and_balance_energy(oracle, device, name, comp, E0, thick, rad_on, shell_on,
                                                 straggle):
    lt, lut0, lut, rad, shell = _tables(comp, E0, thick, rad_on, shell_on)
    n, seed, cap = 3000, 21, 8192
    g_rows, g_cnt, g_code, g_st, modes = _gpu(lt, lut0, lut, rad, shell, device, E0, n, seed, straggle, cap)
    o_rows, o_cnt, o_code, o_st = _oracle(oracle, lt, lut, rad, shell, modes, E0, n, seed, straggle, cap)
    assert g_cnt.max() <= cap and o_cnt.max() <= cap

    def obs(rows, cnt, code, st):
        used = (np.arange(cap)[None, :] < cnt[:, None]).reshape(-1)
        owner = np.repeat(np.arange(n), cnt)
        ev = rows["seg_event"][used]
        out = {"rows": cnt.astype(float), "transmitted": (code == 2).astype(float),
               "backscattered": (code == 1).astype(float), "E_final": np.where(code != 0, st["E_keV"], 5.0)}
        if shell_on:
            out["hard_inelastic"] = np.bincount(owner, ev == 8, minlength=n)
        if rad_on:
            out["hard_radiative"] = np.bincount(owner, ev == 9, minlength=n)
        return out

    go, oo = obs(g_rows, g_cnt, g_code, g_st), obs(o_rows, o_cnt, o_code, o_st)
    for k in go:
        a, b = go[k], oo[k]
        se = math.sqrt(a.var() / n + b.var() / n)
        assert abs(a.mean() - b.mean()) <= 4.0 * se + 1e-12 * abs(b.mean()), (k, a.mean(), b.mean(), se)
    # every device history: E0 = final energy + soft row losses + debited hard losses
    used = (np.arange(cap)[None, :] < g_cnt[:, None]).reshape(-1)
    ev = g_rows["seg_event"][used]
    soft = g_rows["seg_E"][used] - g_rows["seg_E_end"][used]
    debited = np.zeros_like(soft)
    if shell_on:
        debited += np.where(ev == 8, g_rows["seg_hard_W"][used], 0.0)
    if rad_on:
        debited += np.where(ev == 9, g_rows["seg_rad_k"][used] * 1e-3, 0.0)
    owner = np.repeat(np.arange(n), g_cnt)
    total = np.bincount(owner, soft + debited, minlength=n) + g_st["E_keV"]
    np.testing.assert_allclose(total, E0, rtol=1e-12)


def test_simulate_trajectories_modes(device):
    from pyrite_amd.transport.api import simulate_trajectories

    out = simulate_trajectories(E0_keV=5000.0, Ne=400, thickness_ang=2.0e6, composition=GE, seed=3,
                                straggling=True, keep_segments_on_device=False,
                                inelastic_model="shell-soft-hard", radiative_model="analytic-soft-hard")
    ev = out["event_kind"]
    assert (ev == 8).sum() > 0 and (ev == 9).sum() > 0
    assert np.all((out["hard_inelastic_W_keV"] > 0) == (out["hard_inelastic_channel"] >= 0))
    assert out["radiative"]["cutoff_eV"] == 5000.0 and out["inelastic"]["tables"] == "analytic-substitute"
    with pytest.raises(NotImplementedError, match="BremsLib"):
        simulate_trajectories(E0_keV=50.0, Ne=4, thickness_ang=1e4, composition=GE, radiative_model="bremslib-soft-hard")
    with pytest.raises(ValueError, match="electron cutoff"):
        simulate_trajectories(E0_keV=50.0, Ne=4, thickness_ang=1e4, composition=GE,
                              radiative_model="analytic-soft-hard", radiative_cutoff_eV=6000.0)
