"""Physics pins of the oracle's line prologue and transport rate column (CPU).

Each check compares the oracle with something its own formulas do not
contain, so a dropped factor, a wrong sign or a swapped operand fails:

* recoil line positions: the harmonic defect ``2 E2 - E4`` of Huang et al.
  2023 (SI Table S4, graphite rows), through the oracle's exact-recoil lines,
  within 2 sigma; the classical (hbar -> 0) lines give no defect
  (the reference's test_line_recoil_exact.py:163-209);
* absolute line weights and the PXR/CBS interference: Shi et al. 2024 Fig. 2
  (30 nm Si [001], theta = 100 deg): the 5 MeV classical peak heights within
  25 % (measured 13-16 % low: the couplings here have no anomalous f', which
  the reference takes from xraydb, absent offline) and the QM /
  classical-PXR area ratios (measured 0.2-0.4 % from Shi at 5 MeV, 15 % at
  30 keV where the lines sit below the Si K edge and f' matters most; the
  reference's bar is 15 %, test_shi_2024_fig2_intensity.py:121-162).  A
  dropped factor 2 or pi in the weight misses these by 50-200 %;
* the in-medium resonance root against the exact Snell refraction of the
  photon at the exit surface, to first order in the decrement;
* the screened-Rutherford rate column of the transport LUT against a
  numerical angular integral of the screened Rutherford cross section with
  CODATA e^2.
"""

import json
from pathlib import Path

import numpy as np
import pytest

from pyrite_amd.spectrum.crystal import CRYSTALS, Crystal, polarization_pairs
from pyrite_amd.spectrum.lines import keep_window, line_tables

GOLDEN = Path(__file__).parent / "golden"
BEAM = np.array([0.0, 0.0, 1.0])
HBARC = 1973.269804
MC2_EV = 510998.95


def _lines(oracle, crystal, rows, E_grid, r_mid, v_hat, L, T_keV, thickness, n_hat, cbs=True):
    t = line_tables(crystal, rows, E_grid)
    tabs = t.as_dict()
    if not cbs:
        tabs = dict(tabs, u_re=np.zeros_like(tabs["u_re"]), u_im=np.zeros_like(tabs["u_im"]))
    es, ep = polarization_pairs(n_hat, t.g)
    seg = {"r_mid": np.asarray(r_mid, dtype=np.float32)[None], "v_hat": np.asarray(v_hat, dtype=np.float32)[None],
           "L": np.array([L], dtype=np.float32), "E": np.array([T_keV], dtype=np.float32)}
    E_r, a, w, good = oracle.cxr_lines(seg, thickness, n_hat, tabs, es, ep)
    return E_r, a, w, good


# ---- Huang 2023: recoil harmonic defect ------------------------------------------
def _to_sample(axis):
    k = np.cross(axis, BEAM)
    s, c = np.linalg.norm(k), float(axis @ BEAM)
    if s < 1e-15:
        return np.eye(3)
    k /= s
    K = np.array([[0.0, -k[2], k[1]], [k[2], 0.0, -k[0]], [-k[1], k[0], 0.0]])
    return np.eye(3) + s * K + (1.0 - c) * K @ K


def _huang(oracle, classical=False):
    data = json.loads((GOLDEN / "huang-2023-table-s4.json").read_text())
    meas, sim = [], []
    for name, T_keV, e2, e4, th_t, ph_t in data["rows"]:
        info = data["sets"][name]
        if info["material"] != "graphite":
            continue  # the hBN rows need an hBN crystal this build does not carry
        t, p, o = np.radians([th_t, ph_t, info["theta_obs_deg"]])
        R = _to_sample(np.array([np.sin(t) * np.cos(p), np.sin(t) * np.sin(p), np.cos(t)]))
        v_hat = R @ BEAM
        n_hat = R @ np.array([np.sin(o), 0.0, np.cos(o)])
        thick = 10.0 * info["thickness_nm"]
        L = thick / v_hat[2]
        pos = []
        for order in (2, 4):
            grid = np.linspace(300.0, 3000.0, 64)
            if classical:
                with oracle.classical_recoil():
                    E_r, *_ = _lines(oracle, "hopg", [(0, 0, order)], grid, 0.5 * L * v_hat, v_hat, L, T_keV,
                                     thick, n_hat)
            else:
                E_r, *_ = _lines(oracle, "hopg", [(0, 0, order)], grid, 0.5 * L * v_hat, v_hat, L, T_keV, thick,
                                 n_hat)
            pos.append(float(E_r[0]))
        meas.append(2.0 * e2 - e4)
        sim.append(2.0 * pos[0] - pos[1])
    return np.array(meas), np.array(sim), data["detector"]["calibration_uncertainty_eV"]


def test_huang_2023_harmonic_defect(oracle):
    meas, sim, sig_cal = _huang(oracle)
    assert meas.size == 10 and np.all(sim > 15.0)
    res = meas - sim
    sigma = np.hypot(res.std(ddof=1) / np.sqrt(res.size), sig_cal)
    assert abs(res.mean()) <= 2.0 * sigma, (res.mean(), sigma)
    _, classical, _ = _huang(oracle, classical=True)
    assert np.all(np.abs(classical) < 1.0)  # refraction only
    assert (meas - classical).mean() > 10.0 * sigma


# ---- Shi 2024: absolute heights and QM/CL areas ------------------------------------
SHI = json.loads((GOLDEN / "shi-2024-fig2.json").read_text())["curves"]
SI_001 = Crystal("si001", 5.4309, 5.4309, 5.4309, 90.0, 90.0, 90.0, CRYSTALS["silicon"].basis, 0.4632, (0, 0, 1))
N_SHI = np.array([np.sin(np.radians(100.0)), 0.0, np.cos(np.radians(100.0))])


def _shi_rows(nz, hmax=8):
    rows = []
    for h in range(-hmax, hmax + 1):
        for k in range(-hmax, hmax + 1):
            if len({h % 2, k % 2, nz % 2}) == 1 and (h % 2 == 1 or (h + k + nz) % 4 == 0):
                rows.append((h, k, nz))
    return rows


def _shi_spectrum(oracle, T_keV, nz, classical, cbs=True):
    cl, qm = SHI[f"{T_keV:g} keV CL nz=-{nz}"], SHI[f"{T_keV:g} keV QM nz=-{nz}"]
    grid = np.arange(min(cl["E_eV"][0], qm["E_eV"][0]), max(cl["E_eV"][-1], qm["E_eV"][-1]), 0.5)
    args = (oracle, SI_001, _shi_rows(nz), grid, [0.0, 0.0, 150.0], BEAM, 300.0, T_keV, 300.0, N_SHI)
    if classical:
        with oracle.classical_recoil():
            E_r, a, w, _ = _lines(*args, cbs=cbs)
    else:
        E_r, a, w, _ = _lines(*args, cbs=cbs)
    return grid, oracle.sincsq_reduce(E_r, a, w, grid)


@pytest.mark.parametrize("nz", [2, 3])
def test_shi_2024_mev_classical_heights(oracle, nz):
    """At 5 MeV the CBS share is small, so the classical peak is Shi's CL peak
    (photons / (eV sr electron), one electron through the film)."""
    _, total = _shi_spectrum(oracle, 5000.0, nz, classical=True)
    _, pxr = _shi_spectrum(oracle, 5000.0, nz, classical=True, cbs=False)
    cl = SHI[f"5000 keV CL nz=-{nz}"]
    assert total.max() == pytest.approx(cl["peak"], rel=0.25)
    assert (total.max() - pxr.max()) / total.max() < 0.12


@pytest.mark.parametrize(("T_keV", "nz", "rel"), [(30.0, 2, 0.2), (5000.0, 2, 0.02), (5000.0, 3, 0.02)])
def test_shi_2024_qm_over_classical_pxr_area(oracle, T_keV, nz, rel):
    """Shi's QM/CL area ratio is the exact-recoil PXR+CBS total over the classical PXR alone."""
    _, qm = _shi_spectrum(oracle, T_keV, nz, classical=False)
    _, cl_pxr = _shi_spectrum(oracle, T_keV, nz, classical=True, cbs=False)
    shi = SHI[f"{T_keV:g} keV QM nz=-{nz}"]["area_per_sr"] / SHI[f"{T_keV:g} keV CL nz=-{nz}"]["area_per_sr"]
    assert qm.sum() / cl_pxr.sum() == pytest.approx(shi, rel=rel)


# ---- in-medium root: first-order Snell --------------------------------------------
def test_in_medium_root_is_the_snell_refracted_resonance(oracle):
    """With a constant decrement delta, the root is the classical resonance of a
    photon refracted at the exit face: tangential k conserved, |k_in| = (1 -
    delta) w/c; the oracle's root agrees to O(delta^2)."""
    E_grid = np.linspace(2000.0, 20000.0, 64)
    t = line_tables(SI_001, [(1, 1, 3)], E_grid)
    for delta in (1e-5, 3e-5):
        tabs = dict(t.as_dict(), delta_tab=np.full_like(t.delta_tab, delta))
        for theta in (100.0, 120.0, 150.0):
            n_hat = np.array([np.sin(np.radians(theta)), 0.0, np.cos(np.radians(theta))])
            es, ep = polarization_pairs(n_hat, t.g)
            T = 300.0
            seg = {"r_mid": np.array([[0.0, 0.0, 5000.0]], np.float32), "v_hat": BEAM[None].astype(np.float32),
                   "L": np.array([1000.0], np.float32), "E": np.array([T], np.float32)}
            with oracle.classical_recoil():
                E_r, *_ = oracle.cxr_lines(seg, 1.0e4, n_hat, tabs, es, ep)
            g = t.g[0]
            beta = np.sqrt(1.0 - 1.0 / (1.0 + float(np.float32(T)) / (MC2_EV / 1e3)) ** 2)
            v = beta * BEAM
            # inside the slab the photon that leaves along n_hat travels along k_in
            kt = n_hat[:2]
            kz = np.sign(n_hat[2]) * np.sqrt((1.0 - delta) ** 2 - kt @ kt)
            k_in = np.r_[kt, kz]
            E_snell = HBARC * (v @ g) / (1.0 - v @ k_in)
            # second order in delta, amplified by 1/|n_z|^3 at grazing exit
            assert E_r[0] == pytest.approx(E_snell, rel=5 * delta**2 / abs(n_hat[2]) ** 3 + 1e-12), (delta, theta)
            with oracle.classical_recoil():
                E_vac, *_ = oracle.cxr_lines(seg, 1.0e4, n_hat, dict(tabs, delta_tab=np.zeros_like(tabs["delta_tab"])),
                                             es, ep)
            first_order = E_vac[0] * beta * delta / abs(n_hat[2]) / (1.0 - beta * n_hat[2])
            assert abs(E_r[0] - E_vac[0]) > 0.5 * first_order  # the shift itself is far above the bar


# ---- screened-Rutherford rate column ---------------------------------------------
def test_sr_rate_column_is_the_screened_rutherford_cross_section():
    """total_rate = sum n_i sigma_i with sigma the angular integral of
    (Z e^2 / (2 p v))^2 / (sin^2(theta/2) + alpha)^2 (alpha = the LUT's Joy
    screening column), e^2 = 1.43996e-10 keV cm; the reference's rounded
    coefficient 5.21e-21 * 4 pi sits 0.5 % from pi e^4."""
    from pyrite_amd.transport.tables import build_layer_tables, build_transport_lut

    e2 = 1.43996448e-10  # keV cm
    lt = build_layer_tables([(0.0, 1.0e6, [("Si", 0.03), ("C", 0.05)])])
    lut = build_transport_lut(5.0, 20000.0, lt)
    mc2 = MC2_EV * 1e-3
    x = np.concatenate([np.geomspace(1e-12, 1e-2, 20001), np.linspace(1e-2, 1.0, 20001)[1:]])
    for j in (0, 700, 1800, lut.n_energy - 1):
        E = lut.E_grid[j]
        pv = E * (E + 2 * mc2) / (E + mc2)
        rate = 0.0
        for i in range(2):
            Z, n = lt.L_Z[0, i], lt.L_ncm3[0, i]
            alpha = lut.alpha[0, i, j]
            dcs = (Z * e2 / (2 * pv)) ** 2 / (x + alpha) ** 2
            rate += n * 4 * np.pi * np.trapezoid(dcs, x)  # dOmega = 4 pi d(sin^2(theta/2))
        assert lut.total_rate[0, j] == pytest.approx(rate, rel=1e-2), E


def test_keep_window_brackets_the_grid():
    lo, hi = keep_window(np.linspace(1000.0, 50000.0, 1024))
    assert lo < 1000.0 and hi > 50000.0
