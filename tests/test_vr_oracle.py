"""Oracle of the learned variance reduction's decisions (CPU).

The weight window must conserve weight: a split conserves it exactly, a
roulette in expectation; counts are floor/ceil of the weight ratio; child
streams are distinct.  The line mass (expected window content of a narrow
sinc^2 line) is pinned against a brute-force average of the oracle's own
dense reduction over the line's offset to the bins, so a dropped pi, a
missing 1/dE or 1/a fails.
"""

import numpy as np
import pytest


def _inputs(n, seed=0):
    rng = np.random.default_rng(seed)
    exit_code = np.where(rng.random(n) < 0.8, 6, rng.integers(0, 6, n)).astype(np.int8)
    y = rng.normal(0.0, 2.0, n).astype(np.float32)
    w = np.exp(rng.uniform(np.log(1 / 64), 0.0, n))
    keys = rng.integers(0, 2**63, n, dtype=np.uint64) * np.uint64(2) + np.uint64(1)
    return exit_code, y, w, keys


def test_child_keys_distinct_and_deterministic(oracle):
    k = [oracle.vr_child_key(0x1234_5678_9ABC_DEF0, j) for j in range(64)]
    assert len(set(k)) == 64
    assert k[3] == oracle.vr_child_key(0x1234_5678_9ABC_DEF0, 3)
    # a child never reuses its parent's key or a sibling family's keys
    other = {oracle.vr_child_key(0x1234_5678_9ABC_DEF1, j) for j in range(64)}
    assert not other & set(k)


def test_window_counts_and_weight_conservation(oracle):
    ec, y, w, keys = _inputs(20000)
    log_I_ref, w_min, w_max, n_max = 0.5, 1 / 16, 1.0, 32
    count, wc = oracle.vr_window(ec, y, w, keys, log_I_ref, w_min, w_max, n_max)
    ended = ec != 6
    assert np.all(count[ended] == 0) and np.all(wc[ended] == 0)
    wt = np.clip(np.exp(log_I_ref - y.astype(float)), w_min, w_max)
    r = w / wt
    split = (~ended) & (r >= 1)
    assert np.all((count[split] == np.floor(r[split])) | (count[split] == np.minimum(np.ceil(r[split]), n_max)))
    np.testing.assert_allclose(count[split] * wc[split], w[split], rtol=1e-15)
    rou = (~ended) & (r < 1)
    assert set(np.unique(count[rou])) <= {0, 1}
    np.testing.assert_allclose(wc[rou & (count == 1)], wt[rou & (count == 1)], rtol=1e-15)
    # roulette conserves weight in expectation: sum of survivors' weight vs sum of w
    got = (count[rou] * wc[rou]).sum()
    want = w[rou].sum()
    sd = np.sqrt((wt[rou] ** 2 * r[rou] * (1 - r[rou])).sum())
    assert abs(got - want) < 4 * sd


def test_split_count_rounding_is_unbiased(oracle):
    # one state, many streams: E[count] = r for r in (1, n_max)
    n = 40000
    ec = np.full(n, 6, np.int8)
    y = np.zeros(n, np.float32)
    w = np.ones(n)
    keys = np.arange(1, n + 1, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)
    count, wc = oracle.vr_window(ec, y, w, keys, np.log(1 / 5.3), 1 / 16, 1.0, 32)
    assert set(np.unique(count)) == {5, 6}
    assert abs(count.mean() - 5.3) < 4 * np.sqrt(0.3 * 0.7 / n)


def test_line_mass_matches_offset_averaged_reduction(oracle):
    dE = 47.9
    grid = 1000.0 + dE * np.arange(200)
    n_dir, n_g, n_seg = 2, 1, 3
    rng = np.random.default_rng(4)
    a = rng.uniform(0.02, 0.5, n_dir * n_seg)  # widths from a few eV to 150 eV
    w = rng.uniform(0.5, 2.0, n_dir * n_seg).astype(np.float32)
    E_r = rng.uniform(3000.0, 9000.0, n_dir * n_seg)
    E_r[1] = 20000.0  # outside the window: no mass
    lo, hi = 2500.0, 10000.0
    mass = oracle.line_mass(E_r, a, w, n_seg, n_dir, n_g, lo, hi, dE)
    # brute force: average the dense reduction over the line's offset to the bins
    offs = (np.arange(64) + 0.5) / 64 * dE
    for j in range(n_seg):
        tot = 0.0
        for d in range(n_dir):
            k = d * n_seg + j
            if not lo <= E_r[k] < hi:
                continue
            acc = 0.0
            for o in offs:
                out = oracle.sincsq_reduce(np.array([E_r[k] + o]), np.array([a[k]]), np.array([w[k]]), grid)
                acc += out.sum()
            tot += acc / offs.size
        assert mass[j] == pytest.approx(tot, rel=2e-3), j
