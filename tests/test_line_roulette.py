"""Per-primary line roulette: policy validation and the oracle's draws and
selection (CPU).  Mirrors tests/montecarlo/spectrum/test_line_roulette.py of
the reference: its pinned draws, draw-stream properties, policy refusals,
strata and the expectation over seeds."""

import numpy as np
import pytest

from oracle import line_roulette as olr
from pyrite_amd.line_roulette import LINE_ROULETTE_MODEL, LineRoulette

# tests/montecarlo/spectrum/test_line_roulette.py:28 (seed 7, primaries 0, 1, 2)
PINNED_DRAWS = [0.47716136674061527, 0.6692201585934083, 0.42749804493215904]
POLICY = LineRoulette(floor=0.5, keep_fraction=0.25, pilot_histories=8)


def test_draws_match_the_reference_pins_and_are_keyed_by_primary():
    assert olr.roulette_uniforms(7, [0, 1, 2]).tolist() == PINNED_DRAWS
    ids = np.arange(2000)
    draws = olr.roulette_uniforms(7, ids)
    np.testing.assert_array_equal(olr.roulette_uniforms(7, ids[::-1]), draws[::-1])
    np.testing.assert_array_equal(olr.roulette_uniforms(7, ids[500:]), draws[500:])
    assert not np.array_equal(olr.roulette_uniforms(8, ids), draws)
    assert np.all((draws >= 0.0) & (draws < 1.0))


def test_draw_rate_matches_the_floor_within_binomial_error():
    draws = olr.roulette_uniforms(11, np.arange(200_000))
    for floor in (0.1, 0.5, 0.75):
        assert abs(np.mean(draws < floor) - floor) < 5 * np.sqrt(floor * (1 - floor) / 200_000)


def test_policy_round_trips_and_refuses_bad_fields():
    assert set(POLICY.to_dict()) == {"floor", "keep_fraction", "pilot_histories", "model"}
    assert LineRoulette.from_dict(POLICY.to_dict()) == POLICY
    assert POLICY.model == LINE_ROULETTE_MODEL
    for bad in ({"floor": 0.0}, {"floor": 1.5}, {"floor": True}, {"floor": 0.5, "keep_fraction": 0.0},
                {"floor": 0.5, "pilot_histories": 0}, {"floor": 0.5, "model": "other"},
                {"floor": 0.5, "unknown": 1}):
        with pytest.raises((ValueError, TypeError)):
            LineRoulette.from_dict(bad)
    with pytest.raises(TypeError):
        LineRoulette.from_dict([("floor", 0.5)])


def test_policy_refuses_routes_that_are_not_per_history_sums():
    from pyrite_amd.pipeline import CaseConfig

    plain = CaseConfig(line_roulette=POLICY.to_dict())
    POLICY.validate_case(plain, 0.0)
    with pytest.raises(ValueError, match="coherent"):
        POLICY.validate_case(CaseConfig(coherent=True), 0.0)
    with pytest.raises(ValueError, match="streamed"):
        POLICY.validate_case(plain, 0.0, route="streamed")
    with pytest.raises(ValueError, match="max_dE_frac"):
        POLICY.validate_case(plain, 0.05)


def _masses(n=400, seed=3):
    rng = np.random.default_rng(seed)
    m = rng.lognormal(0.0, 1.5, n)
    m[rng.random(n) < 0.1] = 0.0
    return m


def test_pilot_and_heavy_primaries_are_always_scored():
    m = _masses()
    thr = olr.threshold(m[:8], POLICY.keep_fraction)
    assert thr == pytest.approx(np.sort(m[:8])[5] + 0.25 * (np.sort(m[:8])[6] - np.sort(m[:8])[5]))
    p, kept = olr.select(m, 0, 8, thr, POLICY.floor, 7)
    positive = m > 0.0
    assert np.all(p[:8] == 1.0) and np.all(kept[:8] == positive[:8])
    heavy = m >= thr
    heavy[:8] = False
    assert heavy.any() and np.all(kept[heavy]) and np.all(p[heavy] == 1.0)
    low = (m[8:] > 0.0) & (m[8:] < thr)
    assert np.all(p[8:][low] == 0.5)
    assert not np.all(kept[8:][low]) and np.any(kept[8:][low])
    assert not np.any(kept[~positive])


def test_expectation_over_seeds_recovers_the_analog_weights():
    m = _masses()
    thr = olr.threshold(m[:8], POLICY.keep_fraction)
    p, _ = olr.select(m, 0, 8, thr, POLICY.floor, 0)
    mean = np.zeros(m.size)
    n = 20_000
    for seed in range(n):
        _, kept = olr.select(m, 0, 8, thr, POLICY.floor, seed)
        mean += kept / p
    mean /= n
    sigma = np.sqrt((1.0 / p - 1.0) / n)
    pos = m > 0
    assert np.all(np.abs(mean[pos] - 1.0) <= 5 * sigma[pos] + 1e-12)
