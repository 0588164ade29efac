"""Numpy restatement of the reference's per-primary line roulette (TEST INFRASTRUCTURE ONLY).

Follows src/pyrite/montecarlo/runner/line_roulette.py:33-45 (roulette_uniforms:
the SplitMix64 finalizer of (seed ^ "LINEROUL") + golden * (id + 1), top 53
bits), :110-126 (LineScreen._resolve_threshold and _select: the pilot's
(1 - keep_fraction) quantile, probability ``floor`` for positive masses below
it) and :156-171 (the two strata, weights 1 and 1/floor).
Pins: the reference's own pinned draws for seed 7, primaries 0-2
(tests/montecarlo/spectrum/test_line_roulette.py:28,186), order and prefix
independence of the draws, and the binomial draw rate.
"""

from __future__ import annotations

import numpy as np

DRAW_SALT = np.uint64(0x4C494E45524F554C)
GOLDEN = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)


def roulette_uniforms(seed, primaries):
    ids = np.asarray(primaries, dtype=np.int64).astype(np.uint64)
    with np.errstate(over="ignore"):
        x = (np.uint64(int(seed) & 0xFFFFFFFFFFFFFFFF) ^ DRAW_SALT) + GOLDEN * (ids + np.uint64(1))
        x = (x ^ (x >> np.uint64(30))) * M1
        x = (x ^ (x >> np.uint64(27))) * M2
        x ^= x >> np.uint64(31)
    return (x >> np.uint64(11)).astype(np.float64) * 2.0**-53


def threshold(pilot_masses, keep_fraction):
    return float(np.quantile(np.asarray(pilot_masses, dtype=float), 1.0 - keep_fraction))


def select(mass, first_id, pilot, thr, floor, seed):
    """``(probability, kept)`` of primaries ``first_id + i`` with masses ``mass[i]``."""
    mass = np.asarray(mass, dtype=float)
    ids = int(first_id) + np.arange(mass.size)
    p = np.ones(mass.size)
    kept = mass > 0.0
    low = (ids >= pilot) & (mass > 0.0) & (mass < thr)
    p[low] = floor
    kept[low] = roulette_uniforms(seed, ids[low]) < floor
    return p, kept
