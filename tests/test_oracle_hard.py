"""Pins of the oracle's hard-event modes (CPU).

Radiative partition (_jit_radiative.py:26-288, dielectric_radiative.py):
soft first and hard zeroth moments against a dense numerical quadrature of
the same piecewise-linear chi (an independent NumPy interpolant), with and
without the Ter-Mikaelian factor; the photon sampler against the CDF it
inverts.  Shell mode (hard_inelastic.py:120-554): the Newton inversion
against the reference's own bisection oracle, the loss-law antiderivatives
against quadrature of their densities, the recoil kinematics against
momentum conservation, Q_- against its definition, and the soft-loss sampler
against its two moments.  Transport with the modes: per-history energy
balance and the row contract of the mode columns.
"""

import numpy as np
import pytest

from pyrite_amd.transport import hard_events as he
from pyrite_amd.transport.tables import build_layer_tables, build_transport_lut

MC2 = 510998.95


def _rad(comp=(("C", 0.1136),), cutoff=5000.0, dielectric=True):
    lt = build_layer_tables([(0.0, 1.0e6, list(comp))])
    return lt, he.build_radiative_tables(lt, cutoff, dielectric=dielectric)


def _chi_numpy(rt, slot, E_eV, kappa):
    """Independent interpolant: np.interp along kappa per row, linear in ln T between rows."""
    inc = rt.incident[slot]
    lT = np.log(E_eV * 1e-3)
    j = int(np.clip(np.searchsorted(np.log(inc), lT, side="right") - 1, 0, inc.size - 2))
    f = (lT - np.log(inc[j])) / (np.log(inc[j + 1]) - np.log(inc[j]))
    f = min(max(f, 0.0), 1.0)
    nom = rt.nominal[slot].copy()
    v0 = np.interp(kappa, np.r_[nom[:-1], rt.top[slot, j]], rt.chi[slot, j])
    v1 = np.interp(kappa, np.r_[nom[:-1], rt.top[slot, j + 1]], rt.chi[slot, j + 1])
    return v0 + f * (v1 - v0)


def _quad_moments(rt, slot, E_eV, cutoff, plasma):
    """Soft first [eV cm^2] and hard zeroth [cm^2] per atom by dense quadrature in ln k."""
    Z = rt.z[slot]
    scale = 1e-27 * Z * Z
    s = (1.0 + E_eV / MC2) * plasma
    F = (lambda k: k * k / (k * k + s * s)) if s > 0 else (lambda k: np.ones_like(k))
    # hard: int_kc^E chi/k F dk, nodes are kinks: integrate cell by cell
    nodes = np.unique(np.r_[rt.nominal[slot] * E_eV, cutoff, E_eV])
    hard = 0.0
    soft = 0.0
    for a, b in zip(nodes[:-1], nodes[1:]):
        if b <= a:
            continue
        if a >= cutoff:
            x = np.linspace(np.log(a), np.log(b), 4001)
            k = np.exp(x)
            hard += np.trapezoid(_chi_numpy(rt, slot, E_eV, k / E_eV) * F(k), x)
        elif b <= cutoff:
            k = np.linspace(a, b, 4001)
            soft += np.trapezoid(_chi_numpy(rt, slot, E_eV, k / E_eV) * F(k), k)
    return scale * soft, scale * hard


@pytest.mark.parametrize("E_keV,plasma", [(5000.0, 0.0), (5000.0, 30.0), (40.0, 30.0), (700.0, 0.0)])
def test_radiative_moments_match_quadrature(oracle, E_keV, plasma):
    lt, rt = _rad()
    ro = oracle.RadOracle(rt)
    E = E_keV * 1e3
    soft, rate, hard_first = ro.moments(0, E, plasma_eV=plasma)
    qs, qh = _quad_moments(rt, 0, E, rt.cutoff_eV, plasma)
    assert soft > 0 and rate > 0 and hard_first > 0
    np.testing.assert_allclose(soft, qs, rtol=2e-6)
    np.testing.assert_allclose(rate, qh, rtol=2e-6)


def test_radiative_sampler_inverts_its_cdf(oracle):
    lt, rt = _rad()
    ro = oracle.RadOracle(rt)
    E = 5.0e6
    _, rate, _ = ro.moments(0, E, plasma_eV=30.0)
    us = np.array([1e-9, 0.01, 0.2, 0.5, 0.77, 0.999, 1 - 1e-12])
    ks = np.array([ro.sample(0, E, u, plasma_eV=30.0) for u in us])
    assert np.all(np.diff(ks) > 0) and ks[0] >= rt.cutoff_eV and ks[-1] <= E
    for u, k in zip(us, ks):
        # CDF at k = zeroth moment on [kc, k] (the moment routine with the cutoff moved up to k)
        _, above, _ = ro.moments(0, E, cutoff_eV=k, plasma_eV=30.0)
        np.testing.assert_allclose((rate - above) / rate, u, rtol=1e-7, atol=1e-12)


def test_radiative_layer_lut_sums_element_moments(oracle):
    lt, rt = _rad(comp=(("Si", 0.03), ("C", 0.05)))
    ro = oracle.RadOracle(rt)
    E = np.array([10.0, 300.0, 5000.0])
    soft, rate = ro.layer_lut(lt, E)
    for j, Ek in enumerate(E):
        want_s = want_r = 0.0
        for i in range(2):
            s, r, _ = ro.moments(i, Ek * 1e3, plasma_eV=rt.plasma[0])
            want_s += rt.ncm3[i] * 1e-8 * s
            want_r += rt.ncm3[i] * 1e-8 * r
        np.testing.assert_allclose(soft[0, j], want_s * 1e-3, rtol=1e-14)
        np.testing.assert_allclose(rate[0, j], want_r, rtol=1e-14)


def test_plasma_energy_of_graphite():
    # hbar w_p = hbar c sqrt(4 pi r_e n_e): about 30.7 eV at 0.1136 atoms/A^3 (6 electrons each)
    assert he.plasma_energy_eV([("C", 0.1136)]) == pytest.approx(30.66, abs=0.01)


# ---- shell ---------------------------------------------------------------------
CHANNELS = [(284.2, 800.0, 0), (284.2, 800.0, 2), (1839.0, 2500.0, 0), (0.0, 25.0, 2), (11103.1, 13000.0, 0),
            (69525.0, 80000.0, 2), (99.8, 300.0, 2)]


@pytest.mark.parametrize("E", [6.0e3, 5.0e4, 1.0e6, 3.0e7])
def test_newton_transfer_equals_the_bisection_oracle(oracle, E):
    for U, W, branch in CHANNELS:
        for u in (0.0, 1e-6, 0.1, 0.5, 0.9, 0.999999):
            a = oracle.shell_sample_transfer(E, U, W, branch, 1000.0, u)
            b = oracle.shell_sample_transfer(E, U, W, branch, 1000.0, u, bisection=True)
            lo, hi, _ = oracle.shell_bounds(E, U, W, branch, 1000.0)
            if hi > lo:
                assert lo <= a <= hi
            assert a == pytest.approx(b, rel=1e-11, abs=1e-9), (E, U, W, branch, u)


def test_moller_antiderivative_matches_quadrature(oracle):
    for E, U in ((5.0e4, 0.0), (1.0e6, 284.2), (2.0e7, 1839.0)):
        prime = E + U
        a2 = (E / (E + MC2)) ** 2
        lo, hi = 1000.0, 0.5 * prime
        w = np.geomspace(lo, hi * (1 - 1e-9), 200_001)
        r = w / (prime - w)
        F = 1.0 + r * r - (1.0 - a2) * r + a2 * (w / prime) ** 2
        x = np.log(w)
        quad = np.trapezoid(F / w, x)  # int F/W^2 dW = int F/W d(ln W)
        assert oracle.moller_j0(E, prime, w[-1]) - oracle.moller_j0(E, prime, lo) == pytest.approx(quad, rel=1e-6)


def test_close_collision_conserves_momentum(oracle):
    p = lambda T: np.sqrt(T * (T + 2 * MC2))  # noqa: E731
    for E in (1e4, 3e5, 5e6):
        for W in (1200.0, 0.1 * E, 0.45 * E):
            c1, c2 = oracle.shell_cosines(E, 0.0, 20.0, 2, W, 0.3)
            s1, s2 = np.sqrt(1 - c1 * c1), np.sqrt(1 - c2 * c2)
            p0, p1, p2 = p(E), p(E - W), p(W)
            assert p1 * c1 + p2 * c2 == pytest.approx(p0, rel=1e-12)
            assert p1 * s1 == pytest.approx(p2 * s2, rel=1e-9)


def test_distant_recoil_closes_the_momentum_triangle(oracle):
    p = lambda T: np.sqrt(T * (T + 2 * MC2))  # noqa: E731
    E, U, W = 2.0e6, 1839.0, 2500.0
    for u in (0.0, 0.3, 0.9):
        c1, cq = oracle.shell_cosines(E, U, W, 0, W, u)
        # p0 = p1 + q as vectors, in-plane: both cosines are measured from p0
        p0, p1 = p(E), p(E - W)
        # q from the primary's law of cosines, then the secondary's angle must close the triangle
        q2 = p0 * p0 + p1 * p1 - 2 * p0 * p1 * c1
        q = np.sqrt(q2)
        assert p1 * c1 + q * cq == pytest.approx(p0, rel=1e-9)
    # u = 0 is the minimum momentum transfer: qmin(E, W) solves Q(Q+2mc^2) = (p0 - p1)^2
    Q = oracle.qmin(E, W)
    assert Q * (Q + 2 * MC2) == pytest.approx((p(E) - p(E - W)) ** 2, rel=1e-9)


@pytest.mark.parametrize("mean,var", [(50.0, 1.0), (3.0, 2.0), (1.0, 5.0)])
def test_soft_loss_sampler_has_its_two_moments(oracle, mean, var):
    n = 100_000
    x = np.array([oracle.shell_soft_loss(mean, var, oracle.stream_key(3, i))[0] for i in range(n)])
    assert np.all(x >= 0.0)
    se_m = np.sqrt(var / n)
    assert abs(x.mean() - mean) < 5 * se_m
    assert x.var() == pytest.approx(var, rel=0.03)


def test_soft_loss_sampler_limits(oracle):
    assert oracle.shell_soft_loss(0.0, 1.0, 1)[0] == 0.0
    assert oracle.shell_soft_loss(2.5, 0.0, 1) == (2.5, 0)


# ---- transport with the modes ---------------------------------------------------
def _mode_run(oracle, comp, E0, thick, n, straggle, rad_on, shell_on, seed=5):
    layers = [(0.0, thick, list(comp))]
    lt = build_layer_tables(layers)
    lut0 = build_transport_lut(5.0, E0, lt)
    shell = he.build_shell_tables(lt, lut0.E_grid, 1000.0) if shell_on else None
    lut = he.lut_for(lut0, shell)
    rad = None
    if rad_on:
        rt = he.build_radiative_tables(lt, 5000.0)
        soft, rate = oracle.RadOracle(rt).layer_lut(lt, lut.E_grid)
        rad = (rt, oracle.stream_keys(seed ^ he.RADIATIVE_STREAM_SALT, n), soft, rate)
    keys = oracle.stream_keys(seed, n)
    rows, cnt, code, st = oracle.run_transport(lt, lut, keys, np.zeros((n, 3)), np.tile([0.0, 0.0, 1.0], (n, 1)),
                                               np.full(n, E0), np.full(n, 5.0), cap=4000, straggle_on=straggle,
                                               rad=rad, shell=shell)
    return rows, cnt, code, st


@pytest.mark.parametrize("straggle", [False, True])
def test_mode_rows_balance_energy_and_keep_their_contract(oracle, straggle):
    n, E0 = 300, 5000.0
    rows, cnt, code, st = _mode_run(oracle, [("Ge", 0.0442)], E0, 2.0e6, n, straggle, True, True)
    assert cnt.max() < 4000
    ev = rows["seg_event"]
    used = np.concatenate([np.arange(i * 4000, i * 4000 + c) for i, c in enumerate(cnt)])
    ev, W, ch, k, Z = (rows[c][used] for c in ("seg_event", "seg_hard_W", "seg_hard_ch", "seg_rad_k", "seg_rad_Z"))
    assert (ev == 8).sum() > 50 and (ev == 9).sum() > 5
    assert np.all((W > 0) == (ch >= 0))
    assert np.all(W[(ev != 8) & (ev != 7)] == 0) and np.all(W[ev == 8] >= 1.0 - 1e-12)  # keV, above W_c
    assert np.all((k > 0) == (Z > 0)) and np.all(Z[k > 0] == 32) and np.all(k[ev == 9] >= 5000.0 - 1e-6)
    # per history: E0 = final energy + soft losses along rows + the hard losses that were debited
    owner = np.repeat(np.arange(n), cnt)
    soft = rows["seg_E"][used] - rows["seg_E_end"][used]
    debited = np.where(ev == 8, W, 0.0) + np.where(ev == 9, k * 1e-3, 0.0)
    total = np.bincount(owner, soft + debited, minlength=n) + st["E_keV"]
    np.testing.assert_allclose(total, E0, rtol=1e-12)
