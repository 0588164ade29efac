"""Decoherence blend: the oracle's grouped term and the host helpers, on the CPU.

The grouped term ``G = sum_e |sum_{j in e} field_j|^2``
(coherent_grouped_jit_kernel.py:62-230) is pinned by its limits: one group is
the squared all-line field, single-line groups are the incoherent sum, a
single line with ``q = 0, b - a = 0`` is the closed form
``|c|^2 (apb/2)^2 sinc^2(v)``, and the reference's estimator
(coherent_population.py:53-96) turns M identical sampled electrons of a
bunch of N into ``N |S|^2`` per sample.
"""

import numpy as np
import pytest

from oracle import oracle as O
from pyrite_amd.spectrum import coherent as coh


def _records(rng, n, live=True):
    rec = np.zeros((n, 13))
    rec[:, 0] = rng.uniform(3600.0, 3800.0, n)  # E_c [eV]
    rec[:, 1] = rng.uniform(0.005, 0.05, n)  # a [1/eV]
    rec[:, 2] = rng.uniform(-50.0, 50.0, n)  # phase slope [1/eV]
    rec[:, 3] = rng.uniform(-1e5, 1e5, n)  # lattice phase
    rec[:, 4:8] = rng.normal(size=(n, 4)) * (1.0 if live else 0.0)
    rec[:, 8] = rng.uniform(0.0, 1e4, n)  # escape length [A]
    rec[:, 9] = rng.uniform(-50.0, 50.0, n)  # half path change [A]
    rec[:, 10] = rng.uniform(1.0, 2.0, n)  # a + b
    rec[:, 11] = rng.uniform(-0.2, 0.2, n)  # b - a
    rec[:, 12] = rng.uniform(0.0, 0.3, n)  # q
    return rec


E = np.linspace(3000.0, 4500.0, 301)
DOM = np.full(E.size, 1e-4)


def test_one_group_is_the_squared_all_line_field():
    rec = _records(np.random.default_rng(1), 40)
    G, _ = O.coherent_grouped(rec, [0, 40], E, DOM)
    F = O.coherent_field(rec, E, DOM)
    np.testing.assert_allclose(G, (F * F).sum(0), rtol=1e-13)


def test_single_line_groups_are_the_incoherent_sum():
    rec = _records(np.random.default_rng(2), 25)
    G, _ = O.coherent_grouped(rec, np.arange(26), E, DOM)
    want = sum((O.coherent_field(rec[j:j + 1], E, DOM) ** 2).sum(0) for j in range(25))
    np.testing.assert_allclose(G, want, rtol=1e-13)
    # and differs from the coherent square of the whole set
    F = O.coherent_field(rec, E, DOM)
    assert np.max(np.abs(G - (F * F).sum(0))) > 1e-3 * np.max(G)


def test_single_line_is_the_sinc_squared_closed_form():
    rec = _records(np.random.default_rng(3), 1)
    rec[0, 11] = rec[0, 12] = 0.0
    rec[0, 5:8] = 0.0
    G, _ = O.coherent_grouped(rec, [0, 1], E, DOM)
    v = rec[0, 1] * (E - rec[0, 0]) - rec[0, 9] * DOM
    want = rec[0, 4] ** 2 * (0.5 * rec[0, 10]) ** 2 * np.sinc(v / np.pi) ** 2
    np.testing.assert_allclose(G, want, rtol=1e-10, atol=1e-14 * want.max())


def test_empty_groups_add_nothing():
    rec = _records(np.random.default_rng(4), 12)
    G1, _ = O.coherent_grouped(rec, [0, 5, 12], E, DOM)
    G2, _ = O.coherent_grouped(rec, [0, 0, 5, 5, 12, 12], E, DOM)
    np.testing.assert_array_equal(G1, G2)


def test_identical_electrons_give_N_per_sample():
    M, N = 6, 1.0e4
    one = _records(np.random.default_rng(5), 9)
    rec = np.tile(one, (M, 1))
    S2 = (O.coherent_field(one, E, DOM) ** 2).sum(0)
    G, _ = O.coherent_grouped(rec, np.arange(0, 9 * M + 1, 9), E, DOM)
    P = (O.coherent_field(rec, E, DOM) ** 2).sum(0)
    np.testing.assert_allclose(G, M * S2, rtol=1e-12)
    np.testing.assert_allclose(P, M * M * S2, rtol=1e-10)
    mixed = O.mixed_row_power(G, P, np.ones(E.size), coh.pair_scale(N, M))
    np.testing.assert_allclose(mixed / M, N * S2, rtol=1e-10)
    # one physical electron keeps only the self term
    np.testing.assert_array_equal(O.mixed_row_power(G, P, np.ones(E.size), coh.pair_scale(1.0, M)), G)


def test_sampled_blend_limits():
    G, P, F = np.array([1.0, 2.0]), np.array([5.0, 7.0]), np.array([0.0, 1.0])
    np.testing.assert_array_equal(O.mixed_row_power(G, P, F), [1.0, 7.0])


def test_gaussian_longitudinal_factor():
    E_eV = np.array([0.0, 100.0, 1000.0])
    np.testing.assert_array_equal(coh.gaussian_longitudinal_factor(E_eV, 0.0), 1.0)
    rms = 1e-3  # fs: sigma_z = 2.998 A
    want = np.exp(-((E_eV / 1973.269804 * rms * 2997.92458) ** 2))
    np.testing.assert_allclose(coh.gaussian_longitudinal_factor(E_eV, rms), want, rtol=1e-15)
    # F = 1/e where omega sigma_z = 1
    E1 = 1973.269804 / (rms * 2997.92458)
    np.testing.assert_allclose(coh.gaussian_longitudinal_factor([E1], rms), [np.exp(-1.0)], rtol=1e-14)
    with pytest.raises(ValueError):
        coh.gaussian_longitudinal_factor(E_eV, -1.0)


def test_population_and_pair_scale():
    assert coh.physical_bunch_electrons(1.0) == pytest.approx(6.241509074e6, rel=1e-9)
    assert coh.physical_bunch_electrons(0.0) == 0.0
    with pytest.raises(ValueError):
        coh.physical_bunch_electrons(1e-9)  # below one electron
    assert coh.pair_scale(0.0, 10) == 0.0
    assert coh.pair_scale(1.0, 10) == 0.0
    assert coh.pair_scale(101.0, 11) == 10.0
    with pytest.raises(ValueError):
        coh.pair_scale(10.0, 1)
    with pytest.raises(ValueError):
        coh.pair_scale(0.5, 10)


def test_pipeline_refuses_grouped_without_coherent_route():
    from pyrite_amd.pipeline import CaseConfig, CasePipeline

    with pytest.raises(ValueError, match="coherent route"):
        CasePipeline(CaseConfig(Ne=16, coherent_grouped=True))
