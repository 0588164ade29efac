"""Oracle stopping, Urban straggling, LUT and rotation, pinned by known answers.

Pins (reference file:line):
* crossovers H 1.36, B 2.66, Bi 10.46 keV and ratio==1 at the crossover
  (tests/montecarlo/collisions/test_stopping_berger_seltzer.py:190-223);
* non-relativistic limit and F^-(0) = 1 - ln 2 (same file, :60-110);
* Urban closure <dE> = C s to 1e-14 and the moments of the sampled law
  (tests/montecarlo/collisions/test_energy_loss_straggling.py:125-261);
* LUT worked example E = 20 sqrt(2) on [10, 20, 40] keV -> (i, f) = (1, 0.5)
  (docs/.../transport-jit-kernel-walkthrough.md:295-318);
* worked collision: beta(30 keV) = 0.3283761764, cos theta = 11/13 and the
  rotated direction (-4 sqrt 3/13, 0, 11/13) (walkthrough.md:175-253).
"""

import math

import numpy as np
import pytest

from pyrite_amd.transport.elements import TRANSPORT_ELEMENTS
from pyrite_amd.transport.tables import (
    MC2_KEV,
    _dEds_berger_seltzer,
    _dEds_joy_luo,
    build_layer_tables,
    build_transport_lut,
    crossover_keV,
    spliced_dEds,
)


def _elem(el):
    p = TRANSPORT_ELEMENTS[el]
    Z, A, J = float(p["Z"]), float(p["A"]), float(p["J_keV"])
    return Z, A, J, 0.731 + 0.0688 * math.log10(Z)


def test_crossovers_pinned():
    assert crossover_keV("H") == pytest.approx(1.36, abs=0.05)
    assert crossover_keV("B") == pytest.approx(2.66, abs=0.05)
    assert crossover_keV("Bi") == pytest.approx(10.46, abs=0.05)
    for el in TRANSPORT_ELEMENTS:
        Z, A, J, _ = _elem(el)
        E = crossover_keV(el)
        assert 1.0 < E < 300.0
        assert _dEds_joy_luo(Z, A, J, 1.0, E) == pytest.approx(_dEds_berger_seltzer(Z, A, J, 1.0, E), rel=1e-10)


def test_bs_prefactor_reduces_to_joy_luo_constant():
    assert 1.535e-6 * MC2_KEV == pytest.approx(7.85e-4, rel=1e-3)


def test_oracle_element_stopping_nonrelativistic_limit(oracle):
    # -(P mc^2/2E) rho Z/A [2 ln(E/I) + 1 - ln 2] as tau -> 0; force the BS branch (E_cross = 0).
    Z, A, J, k = _elem("Si")
    n = 0.04994
    coeff = n / 0.602214076 * Z  # = rho Z / A in mol/cm^3
    for E in (0.05, 0.02):
        got = oracle.dEds_element(J, k, coeff, 0.0, E)
        limit = -(1.535e-6 * MC2_KEV / (2.0 * E)) * coeff * (2.0 * math.log(E / J) + 1.0 - math.log(2.0))
        assert got == pytest.approx(limit, rel=5e-3 * (E / 0.02))


def test_oracle_stopping_splices_and_bragg_adds(oracle):
    layers = build_layer_tables([(0.0, 1e4, [("Si", 0.03), ("W", 0.01), ("C", 0.02)])])
    J, k, c, x = layers.L_J[0], layers.L_k[0], layers.L_coeff[0], layers.L_E_cross[0]
    E = np.geomspace(1.0, 3e4, 400)
    packed = np.array([oracle.dEds_packed(J, k, c, x, e) for e in E])
    summed = np.array([sum(oracle.dEds_element(J[i], k[i], c[i], x[i], e) for i in range(3)) for e in E])
    np.testing.assert_allclose(packed, summed, rtol=1e-13)
    # vectorised product builder agrees with the scalar oracle
    np.testing.assert_allclose(spliced_dEds(J, k, c, x, E), packed, rtol=1e-13)
    # branch choice: below each crossover the Joy-Luo law, above it Berger-Seltzer
    Z, A, Jsi, ksi = _elem("Si")
    Ec = crossover_keV("Si")
    lo, hi = 0.5 * Ec, 3.0 * Ec
    cs = 0.03 / 0.602214076 * Z
    rho_over = cs / Z * A  # rho such that rho Z / A = coeff
    assert oracle.dEds_element(Jsi, ksi, cs, Ec, lo) == pytest.approx(_dEds_joy_luo(Z, A, Jsi, rho_over, lo), rel=1e-12)
    assert oracle.dEds_element(Jsi, ksi, cs, Ec, hi) == pytest.approx(_dEds_berger_seltzer(Z, A, Jsi, rho_over, hi), rel=1e-12)


@pytest.mark.parametrize("el", ["C", "Si", "W"])
@pytest.mark.parametrize("E", [8.0, 300.0, 5000.0])
def test_urban_closure_mean_equals_C_s(oracle, el, E):
    Z, A, J, k = _elem(el)
    coeff = 0.05 / 0.602214076 * Z
    C = -oracle.dEds_element(J, k, coeff, crossover_keV(el), E)
    s = 137.0
    mean, var = oracle.urban_moments_element(Z, J, C, E, s)
    assert mean == pytest.approx(C * s, rel=1e-13)
    assert var > 0.0


def test_urban_e2_channel_off_below_20_Z2_eV(oracle):
    # W: E_2 = 10 Z^2 eV = 54.76 keV sits above T_up = E/2 below 109.5 keV, so f_2 = 0 and the
    # variance is s (Sigma_1 J^2 + Sigma_3 E_0 T_up) with the single level at I.
    Z, A, J, k = _elem("W")
    E, s = 60.0, 50.0
    C = 0.01
    mean, var = oracle.urban_moments_element(Z, J, C, E, s)
    tau = E / MC2_KEV
    beta_sq = 1.0 - 1.0 / (1.0 + tau) ** 2
    two = 2.0 * MC2_KEV * tau * (tau + 2.0)
    L_I = math.log(two / J) - beta_sq
    s1 = C * 0.45 / L_I * (1.0 / J) * L_I
    T_up = 0.5 * E
    s3 = C * 0.55 * (T_up - 0.01) / (0.01 * T_up * math.log(T_up / 0.01))
    assert var == pytest.approx(s * (s1 * J * J + s3 * 0.01 * T_up), rel=1e-12)


def test_urban_sampled_moments_match_analytic(oracle):
    Z, A, J, k = _elem("C")
    coeff = 0.1136 / 0.602214076 * Z
    n = 60_000
    # (20 keV, 4000 A): ~8.5 continuum quanta per draw, so the sample variance of
    # the 1/E^2 tail (fourth moment ~ T_up^3) settles to a few percent.
    for E, s in ((5000.0, 2.0e4), (20.0, 4000.0)):
        C = -oracle.dEds_element(J, k, coeff, crossover_keV("C"), E)
        mean, var = oracle.urban_moments_element(Z, J, C, E, s)
        draws = np.array([
            oracle.urban_sample_compound([Z], [J], [k], [coeff], [crossover_keV("C")], E, s,
                                         oracle.stream_key(5, i))[0]
            for i in range(n)
        ])
        assert abs(draws.mean() - mean) < 5.0 * math.sqrt(var / n)
        if E < 100.0:
            # the 1/E^2 continuum's fourth moment ~ T_up^3: check the variance where T_up/E_0 is small
            assert draws.var() == pytest.approx(var, rel=0.15)


@pytest.mark.parametrize("lam", [0.3, 2.5, 150.0])
def test_poisson_moments_and_draw_count(oracle, lam):
    n = 20_000
    vals = []
    for i in range(n):
        v, c = oracle.urban_poisson(lam, oracle.stream_key(9, i))
        vals.append(v)
        assert c == math.ceil(lam / 64.0)  # one uniform per chunk
    vals = np.array(vals, dtype=float)
    assert abs(vals.mean() - lam) < 5.0 * math.sqrt(lam / n)
    assert vals.var() == pytest.approx(lam, rel=0.06)


def test_lut_worked_example_and_clamps(oracle):
    lo, inv, n = math.log(10.0), 1.0 / math.log(2.0), 3
    i, f = oracle.lut_index_frac(20.0 * math.sqrt(2.0), lo, inv, n)
    assert i == 1 and f == pytest.approx(0.5, abs=1e-12)
    assert oracle.lut_index_frac(80.0, lo, inv, n) == (1, 1.0)  # upper clamp first
    assert oracle.lut_index_frac(5.0, lo, inv, n) == (0, 0.0)
    assert oracle.lut_index_frac(0.0, lo, inv, n) == (0, 0.0)  # -inf lands low
    # interpolated values of the worked example: 2e7,4e7 -> 3e7 ; lambda = 10/3 A
    rate = 2.0e7 + f * (4.0e7 - 2.0e7)
    assert 1e8 / rate == pytest.approx(10.0 / 3.0, rel=1e-12)


def test_worked_collision_rotation(oracle):
    alpha, R = 1.0 / 3.0, 0.25
    cos_t = 1.0 - 2.0 * alpha * R / (1.0 + alpha - R)
    assert cos_t == pytest.approx(11.0 / 13.0, rel=1e-15)
    d = oracle.rotate_direction([0.0, 0.0, 1.0], cos_t, 2.0 * math.pi * 0.25)
    np.testing.assert_allclose(d, [-4.0 * math.sqrt(3.0) / 13.0, 0.0, 11.0 / 13.0], atol=1e-15)
    assert np.linalg.norm(d) == pytest.approx(1.0, abs=1e-15)


def test_lut_tables_closed_forms():
    comp = [("C", 0.1136)]
    layers = build_layer_tables([(0.0, 1e6, comp)])
    lut = build_transport_lut(5.0, 5000.0, layers, audit=True)
    assert lut.n_energy == int(math.ceil(3.0 * 1024)) + 1  # 3 decades at 1024/decade
    assert lut.E_grid[0] == 5.0 and lut.E_grid[-1] == 5000.0
    # beta(30 keV) = 0.3283761764 (walkthrough.md:223-231) through the inv_beta column
    E = lut.E_grid
    j = np.argmin(abs(E - 30.0))
    g = 1.0 + E[j] / 510.99895
    assert 1.0 / lut.inv_beta[j] == pytest.approx(math.sqrt(1 - 1 / g**2), rel=1e-14)
    g30 = 1.0 + 30.0 / 510.99895
    assert math.sqrt(1.0 - 1.0 / g30**2) == pytest.approx(0.3283761764, abs=5e-11)
    # single element: cdf is 1, alpha = 3.4e-3 Z^0.67 / E
    assert np.all(lut.cdf == 1.0)
    np.testing.assert_allclose(lut.alpha[0, 0], 3.4e-3 * 6.0**0.67 / E, rtol=1e-14)
    assert lut.max_rel_interp_error < 1e-4
