"""Support-windowed sinc^2 reduction with the far tail: oracle pins (CPU).

* x_c huge: every window is broad, the result is the dense reduction;
* inside a line's window the result is its exact sinc^2 (the window
  subtracts the tail the convolution adds back);
* far from a line the result is its cycle-averaged tail w / (2 a^2 y^2),
  y = E - E_r, to O((eps/y)^3) -- a wrong sign or factor in the moments
  (m3 = 2 eps m2, m4 = 3 eps^2 m2) fails.
"""

import numpy as np
import pytest


def _lines(n, rng, lo=3000.0, hi=40000.0):
    E_r = rng.uniform(lo, hi, n)
    a = np.exp(rng.uniform(np.log(0.05), np.log(5.0), n))
    w = rng.uniform(0.1, 1.0, n).astype(np.float32)
    return E_r, a, w


def test_broad_windows_equal_dense(oracle):
    rng = np.random.default_rng(0)
    E_r, a, w = _lines(300, rng)
    n, E0, dE = 512, 1000.0, 97.0
    grid = E0 + dE * np.arange(n)
    dense = oracle.sincsq_reduce(E_r, a, w.astype(float), grid)
    win = oracle.window_reduce(E_r, a, w, E0, dE, n, 1.0e12)
    np.testing.assert_allclose(win, dense, rtol=1e-12, atol=0)


def test_window_interior_exact_and_far_tail_cycle_average(oracle):
    n, E0, dE = 2048, 1000.0, 24.0
    a, w = 2.0, np.float32(0.75)
    E_r = E0 + 1000 * dE + 7.3  # eps = 7.3 eV off the node
    x_c = 100.0
    out = oracle.window_reduce([E_r], [a], [w], E0, dE, n, x_c)
    E = E0 + dE * np.arange(n)
    y = E - E_r
    node = 1000
    R = int(np.ceil((7.3 + x_c / a) / dE))
    exact = np.sin(a * y) ** 2 / (a * y) ** 2
    inside = np.abs(np.arange(n) - node) <= R
    np.testing.assert_allclose(out[inside], float(w) * exact[inside], rtol=1e-12)
    far = np.abs(y) > 20 * R * dE
    avg = float(w) * 0.5 / (a * y[far]) ** 2
    rel = np.abs(out[far] / avg - 1.0)
    assert rel.max() < 6 * (7.3 / np.abs(y[far]).min()) ** 3
    # the moment expansion is exact through 1/y^4: a flipped m3 would leave an O(eps/y) error
    assert rel.max() < 1e-6


def test_windowed_close_to_dense_in_aggregate(oracle):
    rng = np.random.default_rng(2)
    E_r, a, w = _lines(2000, rng)
    n, E0, dE = 1024, 1000.0, 48.0
    grid = E0 + dE * np.arange(n)
    dense = oracle.sincsq_reduce(E_r, a, w.astype(float), grid)
    win = oracle.window_reduce(E_r, a, w, E0, dE, n, 100.0)
    # the cycle-averaged tail moves integrated mass by O(1/x_c) of the line mass at most
    assert abs(win.sum() / dense.sum() - 1.0) < 2e-2
