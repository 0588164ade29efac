"""Offset-averaged second moment of a line's bin content (learned-importance label)."""
import numpy as np
import pytest

from pyrite_amd.vr.importance import line_second_moment


def brute_moments(a_dE, n_off=4001, half_bins=400):
    """Mean and second moment of X = sum_k sinc^2(a dE (k - r)) over r uniform in [0, 1)."""
    r = (np.arange(n_off) + 0.5) / n_off
    k = np.arange(-half_bins, half_bins + 1)[:, None]
    x = a_dE * (k - r[None, :])
    with np.errstate(invalid="ignore", divide="ignore"):
        s = np.where(x == 0, 1.0, np.sin(x) / x) ** 2
    X = s.sum(axis=0)
    return X.mean(), (X * X).mean()


@pytest.mark.parametrize("a_dE", [30.0, 100.0, 300.0])
def test_narrow_lines_follow_the_sinc4_integral(a_dE):
    mean, m2 = brute_moments(a_dE)
    assert mean == pytest.approx(np.pi / a_dE, rel=2e-2)
    assert float(line_second_moment(1.0, np.array([a_dE]))[0]) == pytest.approx(m2, rel=3e-2)


@pytest.mark.parametrize("a_dE", [0.05, 0.2])
def test_wide_lines_deposit_their_mean(a_dE):
    mean, m2 = brute_moments(a_dE, half_bins=4000)
    assert m2 == pytest.approx(mean * mean, rel=2e-2)
    assert float(line_second_moment(1.0, np.array([a_dE]))[0]) == pytest.approx(m2, rel=3e-2)


def test_torch_and_numpy_agree():
    torch = pytest.importorskip("torch")
    a = np.array([0.1, 1.0, 4.0, 50.0])
    w = np.array([2.0, 0.5, 1.0, 3.0])
    t = line_second_moment(torch.tensor(w), torch.tensor(a)).numpy()
    np.testing.assert_allclose(t, line_second_moment(w, a), rtol=1e-15)
