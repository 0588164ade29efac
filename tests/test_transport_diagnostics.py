"""Per-flight transport diagnostics (CPU), after the reference's
batching.py:98-217 and test_elsepa_elastic.py:358-378.

Pinned by limits the summaries must meet: a zero-length flight has no loss,
hazard change, clock error or overshoot; the hazard at both ends of a flight
is the LUT's own total rate at those energies (each model: analytic, Mott,
ELSEPA), so a planted change of model shows; a flight that loses energy below
its cutoff overshoots by exactly the difference; percentiles of a known set.
"""

import numpy as np
import pytest

from pyrite_amd.transport.diagnostics import flight_diagnostic_summary, layer_hazard, percentile_summary
from pyrite_amd.transport.tables import build_layer_tables, build_transport_lut, spliced_dEds
from tests.transport_cases import MOTT_FIXTURE_DIR, elsepa_test_tables

LAYERS = [(0.0, 2.0e3, [("W", 0.0632)]), (2.0e3, 3.0e4, [("Si", 0.025), ("C", 0.05)])]


def test_percentiles():
    s = percentile_summary(np.arange(101.0))
    assert s == {"p50": 50.0, "p90": 90.0, "p99": 99.0, "max": 100.0}
    assert percentile_summary([]) == {"p50": None, "p90": None, "p99": None, "max": None}


@pytest.mark.parametrize("model", ["sr", "mott", "elsepa"])
def test_hazard_is_the_luts_total_rate(model):
    kw = {"elastic_tables": elsepa_test_tables(LAYERS)} if model == "elsepa" else {}
    lt = build_layer_tables(LAYERS, model, MOTT_FIXTURE_DIR, **kw)
    lut = build_transport_lut(5.0, 40.0, lt)
    for L in range(2):
        np.testing.assert_allclose(layer_hazard(lt, L, lut.E_grid), lut.total_rate[L], rtol=1e-13)


def test_flight_limits():
    lt = build_layer_tables(LAYERS)
    E = np.array([30.0, 20.0, 6.0])
    L = np.array([0.0, 50.0, 4000.0])
    out = flight_diagnostic_summary(E, L, np.array([0, 1, 2]), np.array([0, 1, 1]), np.full(3, 5.0), lt)
    assert out["n_flights"] == 3
    # the zero-length flight contributes zeros; the long low-energy flight falls below its cutoff
    dEds = spliced_dEds(lt.L_J[1, :2], lt.L_k[1, :2], lt.L_coeff[1, :2], lt.L_E_cross[1, :2], np.array([6.0]))[0]
    E_end = 6.0 + dEds * 4000.0
    assert E_end < 5.0
    assert out["cutoff_overshoot_keV"]["max"] == pytest.approx(5.0 - E_end, rel=1e-12)
    assert out["fractional_energy_loss"]["max"] == pytest.approx((6.0 - E_end) / 6.0, rel=1e-12)
    loss1 = -spliced_dEds(lt.L_J[1, :2], lt.L_k[1, :2], lt.L_coeff[1, :2], lt.L_E_cross[1, :2],
                          np.array([20.0]))[0] * 50.0 / 20.0
    assert out["fractional_energy_loss"]["p50"] == pytest.approx(loss1, rel=1e-12)
    h0, h1 = layer_hazard(lt, 1, np.array([20.0])), layer_hazard(lt, 1, np.array([20.0 - loss1 * 20.0]))
    assert out["relative_hazard_change"]["p50"] == pytest.approx(abs(h1 - h0)[0] / h0[0], rel=1e-12)
