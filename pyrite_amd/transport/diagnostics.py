"""Per-flight transport error diagnostics (``collect_diagnostics=True``).

The reference's ``_flight_diagnostic_summary``
(``src/pyrite/montecarlo/transport/batching.py:98-217``): every row is one
physical flight; at its stored start energy the spliced stopping predicts an
end energy ``E_end = E_start + (dE/ds) L``, and the layer's elastic hazard
(the sum of its elements' rates under the run's elastic model) is evaluated
at both ends.  Only fixed-size percentile summaries (p50, p90, p99, max) are
kept: the fractional energy loss, the relative hazard change, the relative
clock error of a left-endpoint against a midpoint velocity, and the cutoff
overshoot.  Host-side and read-only: nothing feeds back into transport.
"""

from __future__ import annotations

import numpy as np

from .tables import MC2_KEV, LayerTables, _browning_rates, _sr_rates, spliced_dEds


def percentile_summary(values) -> dict:
    values = np.asarray(values, dtype=np.float64)
    if values.size == 0:
        return {"p50": None, "p90": None, "p99": None, "max": None}
    p50, p90, p99 = np.percentile(values, (50.0, 90.0, 99.0))
    return {"p50": float(p50), "p90": float(p90), "p99": float(p99), "max": float(values.max())}


def _beta(E):
    g = 1.0 + E / MC2_KEV
    return np.sqrt(1.0 - 1.0 / (g * g))


def layer_hazard(tables: LayerTables, L: int, E) -> np.ndarray:
    """Total elastic rate [1/cm] of layer ``L`` at energies ``E`` [keV] under the run's model."""
    E = np.asarray(E, dtype=np.float64)
    total = np.zeros(E.shape)
    for i in range(int(tables.L_nel[L])):
        if tables.elastic_model == "elsepa":
            from ..tables.elsepa import element_rate

            total += element_rate(tables.elsepa, L, i, E)
        elif tables.elastic_model == "mott":
            total += _browning_rates(tables.L_mott_numer[L, i], tables.L_mott_denom1[L, i],
                                     tables.L_mott_denom2[L, i], E)
        else:
            total += _sr_rates(tables.L_sr_rate_numer[L, i], tables.L_sr_joy_numer[L, i], E)
    return total


def flight_diagnostic_summary(E_start_keV, L_ang, elec_id, layer, E_cut_by_electrons, tables: LayerTables) -> dict:
    """Percentile summaries of the frozen-state errors of every flight (row)."""
    E_start = np.asarray(E_start_keV, dtype=np.float64)
    length = np.asarray(L_ang, dtype=np.float64)
    electron = np.asarray(elec_id, dtype=np.int64)
    lay = np.asarray(layer, dtype=np.int64)
    n = E_start.size
    stopping = np.zeros(n)
    hazard_start = np.zeros(n)
    hazard_end = np.zeros(n)
    for L in range(tables.n_layers):
        mask = lay == L
        if not mask.any():
            continue
        m = int(tables.L_nel[L])
        stopping[mask] = spliced_dEds(tables.L_J[L, :m], tables.L_k[L, :m], tables.L_coeff[L, :m],
                                      tables.L_E_cross[L, :m], E_start[mask])
    E_end = E_start + stopping * length
    for L in range(tables.n_layers):
        mask = lay == L
        if not mask.any():
            continue
        hazard_start[mask] = layer_hazard(tables, L, E_start[mask])
        with np.errstate(invalid="ignore", divide="ignore"):
            hazard_end[mask] = layer_hazard(tables, L, E_end[mask])
    midpoint = 0.5 * (E_start + E_end)
    with np.errstate(invalid="ignore", divide="ignore"):
        left_clock = length / _beta(E_start)
        midpoint_clock = length / _beta(midpoint)
        fractional_loss = np.maximum(0.0, (E_start - E_end) / E_start)
        relative_hazard_change = np.abs(hazard_end - hazard_start) / hazard_start
        relative_clock_error = np.abs(midpoint_clock - left_clock) / midpoint_clock
    cutoff = np.asarray(E_cut_by_electrons, dtype=np.float64)[electron]
    cutoff_overshoot = np.maximum(0.0, cutoff - E_end)
    return {
        "n_flights": int(n),
        "fractional_energy_loss": percentile_summary(fractional_loss),
        "relative_hazard_change": percentile_summary(relative_hazard_change),
        "relative_clock_error_estimate": percentile_summary(relative_clock_error),
        "cutoff_overshoot_keV": percentile_summary(cutoff_overshoot),
    }
