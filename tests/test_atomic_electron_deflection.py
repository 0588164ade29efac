"""Atomic-electron deflection folded into the elastic rate (CPU).

After the reference's tests/montecarlo/collisions/test_atomic_electron_deflection.py:
continuous stopping scales both analytic rate numerators by (Z + 1)/Z and
leaves the denominators and the screening bit for bit; the opt-out is the
default bit for bit; shell mode (explicit hard collisions) and unknown models
are refused before any device work; at an unchanged angular law and the same
optical-depth draws every first flight shrinks by exactly Z/(Z + 1) (oracle
transport, every history, not only the mean).
"""

import numpy as np
import pytest

from pyrite_amd.transport.tables import build_layer_tables, build_transport_lut
from tests.transport_cases import MOTT_FIXTURE_DIR

SI = [(0.0, 2.0e5, [("Si", 0.04996)])]
STACK = [(0.0, 2.0e3, [("W", 0.0632)]), (2.0e3, 3.0e4, [("Si", 0.025), ("C", 0.05)])]


@pytest.mark.parametrize("model", ["sr", "mott"])
def test_continuous_mode_scales_the_analytic_rates(model):
    off = build_layer_tables(STACK, model, MOTT_FIXTURE_DIR)
    on = build_layer_tables(STACK, model, MOTT_FIXTURE_DIR, atomic_electron_deflection="kawrakow")
    Z = off.L_Z
    live = Z > 0
    scale = np.where(live, (Z + 1.0) / np.where(live, Z, 1.0), 1.0)
    np.testing.assert_allclose(on.L_sr_rate_numer, off.L_sr_rate_numer * scale, rtol=1e-15)
    np.testing.assert_allclose(on.L_mott_numer, off.L_mott_numer * scale, rtol=1e-15)
    assert on.L_sr_rate_numer[0, 0] == pytest.approx(off.L_sr_rate_numer[0, 0] * 75.0 / 74.0, rel=1e-15)
    for name in ("L_sr_joy_numer", "L_mott_denom1", "L_mott_denom2", "L_Z", "L_J", "L_coeff", "L_E_cross"):
        np.testing.assert_array_equal(getattr(on, name), getattr(off, name))
    lut_off, lut_on = build_transport_lut(5.0, 40.0, off), build_transport_lut(5.0, 40.0, on)
    np.testing.assert_array_equal(lut_on.alpha, lut_off.alpha)
    np.testing.assert_array_equal(lut_on.dEds, lut_off.dEds)
    # a one-element layer's rate scales by (Z+1)/Z exactly; a mixture's element CDF moves
    np.testing.assert_allclose(lut_on.total_rate[0], lut_off.total_rate[0] * 75.0 / 74.0, rtol=1e-14)
    assert not np.allclose(lut_on.cdf[1, 0], lut_off.cdf[1, 0], rtol=1e-6)


def test_opt_out_is_the_default_bit_for_bit():
    base = build_layer_tables(STACK)
    off = build_layer_tables(STACK, "sr", None, atomic_electron_deflection=None)
    for name in ("L_sr_rate_numer", "L_sr_joy_numer", "L_mott_numer", "L_mott_denom1", "L_mott_denom2"):
        np.testing.assert_array_equal(getattr(base, name), getattr(off, name))


def test_shell_mode_and_unknown_models_are_refused():
    with pytest.raises(ValueError, match="needs elastic_model='elsepa'"):
        build_layer_tables(SI, "sr", None, atomic_electron_deflection="kawrakow", hard_cutoff_eV=50.0)
    with pytest.raises(ValueError, match="atomic_electron_deflection must be None"):
        build_layer_tables(SI, "sr", None, atomic_electron_deflection="z1")
    from pyrite_amd.transport.api import simulate_trajectories

    with pytest.raises(ValueError, match="atomic_electron_deflection must be None"):
        simulate_trajectories(30.0, 4, 1e4, composition=[("Si", 0.05)], atomic_electron_deflection="z1")
    with pytest.raises(ValueError, match="needs elastic_model='elsepa'"):
        simulate_trajectories(30.0, 4, 1e4, composition=[("Si", 0.05)], atomic_electron_deflection="kawrakow",
                              inelastic_model="shell-soft-hard")


def test_first_flights_shrink_by_z_over_z_plus_one(oracle):
    n = 3000
    first = {}
    for model in (None, "kawrakow"):
        lt = build_layer_tables(SI, atomic_electron_deflection=model)
        lut = build_transport_lut(5.0, 30.0, lt)
        rows, cnt, code, _ = oracle.run_transport(lt, lut, oracle.stream_keys(11, n), np.zeros((n, 3)),
                                                  np.tile([0.0, 0.0, 1.0], (n, 1)), np.full(n, 30.0),
                                                  np.full(n, 5.0), cap=8, max_steps=4)
        assert np.all(cnt >= 1)
        first[model] = rows["seg_len"].reshape(n, 8)[:, 0]
    np.testing.assert_allclose(first["kawrakow"] / first[None], 14.0 / 15.0, rtol=1e-13)


# ---- energy-dependent xi (shell mode, ELSEPA tables) ------------------------------------------
MC2_EV = 510998.95


def _moller_sin2_moment(tau, tau_c):
    """g_M by adaptive quadrature of the Moller DCS times the primary's sin^2
    theta (units mc^2 = 2 pi r0^2 = 1; the nucleus's sin^2 moment per Z^2 is
    2 g_R / (beta^2 tau (tau + 2)))."""
    from scipy.integrate import quad

    b2 = tau * (tau + 2.0) / (tau + 1.0) ** 2

    def f(eps):
        w = eps * tau
        dcs = (1.0 / eps**2 + 1.0 / (1.0 - eps) ** 2 + (tau / (tau + 1.0)) ** 2
               - (2.0 * tau + 1.0) / (tau + 1.0) ** 2 / (eps * (1.0 - eps))) / (b2 * tau)
        return dcs * 2.0 * w / (tau * (tau - w + 2.0))

    return quad(f, tau_c / tau, 0.5, limit=200, epsabs=0.0, epsrel=1e-11)[0] * b2 * tau * (tau + 2.0) / 2.0


def _moller_closed_form(t, tc):
    """EGSnrc PIRS-701 Eq. 4.7.21 (the published closed form of g_M)."""
    d0 = (t + 2.0) / (t + 1.0)
    d1 = (t + 1.0) ** 2
    return (np.log(0.5 * t / tc) + (1.0 + d0 * d0) * np.log(2.0 * (t - tc + 2.0) / (t + 4.0))
            - (0.25 * (t + 2.0) ** 2 + (t + 2.0) * (t + 0.5) / d1) * np.log((t + 4.0) * (t - tc) / (t * (t - tc + 2.0)))
            + 0.5 * (t - 2.0 * tc) * (t + 2.0) * (1.0 / (t - tc) - 1.0 / d1))


def _sr_sin2_moment(eta):
    from scipy.integrate import quad

    return 0.5 * quad(lambda mu: (1.0 - mu * mu) / (1.0 - mu + 2.0 * eta) ** 2, -1.0, 1.0,
                      points=[1.0 - 1e-6, 1.0 - 1e-4, 1.0 - 1e-2], limit=400, epsabs=0.0, epsrel=1e-11)[0]


@pytest.mark.parametrize(("tau", "tau_c"), [(0.1, 0.01), (0.6, 1e-4), (0.6, 0.05), (1.5, 0.1), (5.0, 0.2)])
def test_moller_g_against_quadrature_and_the_closed_form(tau, tau_c):
    from pyrite_amd.transport.atomic_electrons import moller_g

    g = moller_g(np.array([tau]), tau_c)[0]
    assert g == pytest.approx(_moller_sin2_moment(tau, tau_c), rel=1e-9)
    assert g == pytest.approx(_moller_closed_form(tau, tau_c), rel=1e-12)


@pytest.mark.parametrize("eta", [1e-6, 1e-4, 1e-2, 0.3])
def test_screened_rutherford_g_against_the_direct_moment(eta):
    from pyrite_amd.transport.atomic_electrons import screened_rutherford_g

    assert screened_rutherford_g(eta) == pytest.approx(_sr_sin2_moment(eta), rel=1e-8)


def test_moller_g_vanishes_without_hard_collisions():
    from pyrite_amd.transport.atomic_electrons import moller_g

    g = moller_g(np.array([0.05, 0.1, 0.1000001]), 0.05)
    assert g[0] == 0.0 and g[1] == 0.0 and 0.0 < g[2] < 1e-5


@pytest.mark.parametrize(("element", "energy_keV", "cutoff_eV"),
                         [("Si", 10.0, 50.0), ("Si", 300.0, 50.0), ("Si", 300.0, 10_000.0), ("W", 300.0, 1_000.0),
                          ("W", 800.0, 10_000.0)])
def test_xi_against_an_independent_evaluation(element, energy_keV, cutoff_eV):
    """EGSnrc screening for one element written out, direct moments, xi = 1 - g_M/g_R."""
    from pyrite_amd.transport.atomic_electrons import atomic_electron_xi, elastic_rate_scale, moliere_screening_eta

    Z = {"Si": 14.0, "W": 74.0}[element]
    tau, tau_c = energy_keV * 1e3 / MC2_EV, cutoff_eV / MC2_EV
    comp = [(element, 0.05)]
    a2 = (1.0 / 137.035999084) ** 2
    ratio = (0.1569 / 7821.6) * np.exp(np.log(1.0 + 3.34 * a2 * Z * Z) + (2.0 / 3.0) * np.log(Z))
    eta = ratio / (4.0 * (MC2_EV * 1e-6) ** 2 * tau * (tau + 2.0))
    assert moliere_screening_eta(comp, tau) == pytest.approx(eta, rel=1e-8)
    xi = 1.0 - min(_moller_sin2_moment(tau, tau_c) / _sr_sin2_moment(eta), 1.0)
    got = atomic_electron_xi(energy_keV * 1e3, comp, cutoff_eV)
    assert got == pytest.approx(xi, rel=1e-7, abs=1e-10)
    assert elastic_rate_scale(Z, got) == pytest.approx(1.0 + xi / Z, rel=1e-9)


def test_xi_falls_with_energy_rises_with_cutoff_and_saturates():
    from pyrite_amd.transport.atomic_electrons import atomic_electron_xi

    si = [("Si", 0.04994)]
    E = np.geomspace(2.5e3, 1e6, 40)
    low, high = atomic_electron_xi(E, si, 1_000.0), atomic_electron_xi(E, si, 10_000.0)
    assert np.all(np.diff(low) < 0.0)
    assert np.all((0.0 < low) & (low <= high) & (high <= 1.0))
    assert atomic_electron_xi(1_999.0, si, 1_000.0) == 1.0  # no hard Moller collision below 2 W_c
    np.testing.assert_array_equal(atomic_electron_xi(E, si, None), 1.0)
    # Si, W_c = 50 eV: Moller above W_c already exceeds Z g_R, so xi clips to 0
    np.testing.assert_array_equal(atomic_electron_xi(np.geomspace(1e4, 1e6, 9), si, 50.0), 0.0)


def test_elsepa_nodes_carry_the_energy_dependent_factor():
    from pyrite_amd.transport.atomic_electrons import atomic_electron_xi

    E = np.array([4.0, 40.0, 400.0])
    table = {"energy_eV": E * 1e3, "total_elastic_cm2": 1e-17 / E, "mu": np.linspace(0.0, 1.0, 5),
             "dcs_cm2_sr": np.ones((3, 5))}
    layers = [(0.0, 1e5, [("Si", 0.04994)])]
    off = build_layer_tables(layers, "elsepa", elastic_tables=[[table]]).elsepa
    on = build_layer_tables(layers, "elsepa", atomic_electron_deflection="kawrakow", hard_cutoff_eV=50.0,
                            elastic_tables=[[table]]).elsepa
    xi = atomic_electron_xi(table["energy_eV"], layers[0][2], 50.0)
    np.testing.assert_allclose(on.log_rate - off.log_rate, np.log1p(xi / 14.0), rtol=1e-12, atol=1e-15)
    np.testing.assert_array_equal(on.cdf, off.cdf)
    cont = build_layer_tables(layers, "elsepa", atomic_electron_deflection="kawrakow", elastic_tables=[[table]])
    np.testing.assert_allclose(cont.elsepa.log_rate - off.log_rate, np.log(15.0 / 14.0), rtol=1e-12)
