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


def test_second_weight_stays_with_copy_zero(oracle):
    ec, y, w, keys = _inputs(20000, seed=5)
    w2 = np.random.default_rng(6).uniform(0.5, 2.0, w.size)
    count, wc, w2c = oracle.vr_window(ec, y, w, keys, 0.5, 1 / 16, 1.0, 32, weight2=w2)
    wt = np.clip(np.exp(0.5 - y.astype(float)), 1 / 16, 1.0)
    r = w / wt
    split = (ec == 6) & (r >= 1)
    np.testing.assert_array_equal(w2c[split], w2[split])
    rou = (ec == 6) & (r < 1)
    surv = rou & (count == 1)
    np.testing.assert_allclose(w2c[surv], w2[surv] / r[surv], rtol=1e-15)
    # expectation over the roulette: sum of surviving w2 vs sum of w2
    sd = np.sqrt((w2[rou] ** 2 * (1 - r[rou]) / r[rou]).sum())
    assert abs(w2c[rou].sum() - w2[rou].sum()) < 4 * sd


def test_line_roulette_keeps_the_expectation(oracle):
    rng = np.random.default_rng(8)
    n_seg, n_dir, n_g = 400, 3, 2
    w = rng.uniform(0.1, 1.0, n_seg * n_dir * n_g).astype(np.float32)
    mass = np.where(rng.random(n_seg) < 0.3, rng.uniform(2.0, 5.0, n_seg), rng.uniform(0.0, 1.0, n_seg))
    m_thr = 2.0
    acc = np.zeros_like(w, dtype=float)
    trials = 400
    for t in range(trials):
        out = oracle.line_roulette(w, n_seg, n_dir, n_g, mass, m_thr, key=1234 + t)
        acc += out
        sure = np.repeat(np.repeat(mass >= m_thr, n_g)[None, :], n_dir, 0).reshape(-1)
        np.testing.assert_array_equal(out[sure], w[sure])
    mean = acc / trials
    # a row of mass m survives with p = m/m_thr; its weight's estimate has sd w sqrt((1-p)/p / trials)
    p = np.clip(mass / m_thr, 1e-12, 1.0)
    pl = np.repeat(np.repeat(p, n_g)[None, :], n_dir, 0).reshape(-1)
    sd = w * np.sqrt((1 - pl) / pl / trials)
    ok = pl > 0.05
    assert np.all(np.abs(mean[ok] - w[ok]) < 5 * sd[ok] + 1e-6)


def test_window_band_leaves_moderate_ratios_alone(oracle):
    ec, y, w, keys = _inputs(20000, seed=9)
    count, wc = oracle.vr_window(ec, y, w, keys, 0.5, 1 / 16, 1.0, 32, ratio_lo=0.5, ratio_hi=2.0)
    wt = np.clip(np.exp(0.5 - y.astype(float)), 1 / 16, 1.0)
    r = w / wt
    mid = (ec == 6) & (r >= 0.5) & (r <= 2.0)
    assert mid.sum() > 1000
    np.testing.assert_array_equal(count[mid], 1)
    np.testing.assert_array_equal(wc[mid], w[mid])
    hi = (ec == 6) & (r > 2.0)
    np.testing.assert_allclose(count[hi] * wc[hi], w[hi], rtol=1e-15)
    lo = (ec == 6) & (r < 0.5)
    assert set(np.unique(count[lo])) <= {0, 1}
