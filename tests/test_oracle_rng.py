"""Oracle RNG pinned against the reference's known answers.

Vectors: the SplitMix64 inputs and the independent pure-Python finalizer of
tests/montecarlo/transport/test_transport_per_electron.py:52-97 and the K=0
table of docs/repo-design/compute/transport-jit-kernel-walkthrough.md:63-67.
"""

import numpy as np
import pytest

U64 = (1 << 64) - 1


def ref_splitmix64(x):
    """Independent pure-Python SplitMix64 from the published constants."""
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & U64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & U64
    return x ^ (x >> 31)


def ref_uniform(key, counter):
    z = ref_splitmix64((key + 0x9E3779B97F4A7C15 * (counter + 1)) & U64)
    return (z >> 11) * 2.0**-53


@pytest.mark.parametrize("x", [0, 1, 2, 3, 2**32 - 1, 2**63 + 7, 2**64 - 1])
def test_splitmix64_matches_published_reference(oracle, x):
    assert oracle.splitmix64(x) == ref_splitmix64(x)


def test_walkthrough_key_zero_table(oracle):
    # transport-jit-kernel-walkthrough.md:63-67
    z = [ref_splitmix64((0x9E3779B97F4A7C15 * (c + 1)) & U64) for c in range(3)]
    assert [f"{v:016x}" for v in z] == ["e220a8397b1dcdaf", "6e789e6aa1b965f4", "06c45d188009454f"]
    assert [v >> 11 for v in z] == [7956156453446585, 3886858653415212, 238094247788840]
    u = [oracle.stream_uniform(0, c) for c in range(3)]
    assert u == pytest.approx([0.8833108082136426, 0.4315279970485100, 0.0264337715925977], abs=1e-16)


def test_stream_uniform_matches_pure_python(oracle):
    for elec in (0, 1, 4095):
        key = oracle.stream_key(3, elec)
        assert key == ref_splitmix64((3 + 0x9E3779B97F4A7C15 * (elec + 1)) & U64)
        for counter in (0, 1, 2, 997, 10**6):
            assert oracle.stream_uniform(key, counter) == ref_uniform(key, counter)


def test_stream_keys_block_prefix(oracle):
    full = oracle.stream_keys(7, 512)
    assert np.array_equal(oracle.stream_keys(7, 100, start=200), full[200:300])
    assert [int(k) for k in full[:5]] == [oracle.stream_key(7, e) for e in range(5)]


def test_stream_draws_uniform_moments(oracle):
    key = oracle.stream_key(0, 0)
    draws = np.array([oracle.stream_uniform(key, i) for i in range(100_000)])
    assert draws.min() >= 0.0 and draws.max() < 1.0
    assert abs(draws.mean() - 0.5) < 0.01
    assert abs(draws.var() - 1.0 / 12.0) < 0.005


def test_urban_key_domain(oracle):
    # straggling.py:36-54, restated independently
    key = oracle.stream_key(11, 5)
    uk = ref_splitmix64(key ^ 0xD6E8FEB86659FD93)
    assert oracle.lib().or_urban_stream_key(key) == uk
    for flight, sub in ((0, 0), (3, 1), (2**31, 7)):
        comb = ((flight << 32) + sub) & U64
        want = ref_splitmix64((uk + 0x9E3779B97F4A7C15 * comb) & U64)
        assert oracle.lib().or_urban_flight_key(uk, flight, sub) == want
