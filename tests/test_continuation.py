"""Host bookkeeping of the continuation driver (CPU, stand-in launch).

The stand-in obeys the launch contract of ``hr_transport_continue``: lane i
emits at most ``rows_per_round`` rows of its history per launch, writes them
to its slots, adds them to the record's row count and reports suspended (6)
or the history's exit code.  Histories have known lengths, so every row must
land at its electron-major position exactly once, whatever the lane count,
round length, refill order and tail packing.
"""

import numpy as np
import pytest
import torch

from pyrite_amd.transport.continuation import ContinuationDriver, canonical_rows
from pyrite_amd.transport.core import TransportControl


def _lengths(first, n):
    h = np.arange(first, first + n)
    T = (h * 7919) % 61
    T[h % 17 == 0] = 0  # not entered
    T[h % 13 == 0] = 400  # long tracks reach the packed tail
    return T


def _exit_for(h, T):
    return 5 if T == 0 else int(h % 4)


class StandIn:
    def __init__(self, drv, first, T):
        self.d, self.first, self.T = drv, first, T
        self.launches = []

    def __call__(self, n_lanes, R):
        d = self.d
        self.launches.append((n_lanes, R))
        for i in range(n_lanes):
            h = int(d.history[i])
            if h < 0:
                d.seg_count[i] = 0
                d.exit_code[i] = -1
                continue
            T = int(self.T[h - self.first])
            left = T - int(d.nseg[i])
            k = min(R, left)
            for j in range(k):
                d.rows["seg_id"][i * R + j] = h
                d.rows["seg_len"][i * R + j] = float(int(d.nseg[i]) + j)
            d.seg_count[i] = k
            d.nseg[i] += k
            d.exit_code[i] = 6 if left > R else _exit_for(h, T)


@pytest.mark.parametrize("lanes,R,sort_lanes", [(16, 5, False), (7, 64, False), (64, 1, False), (16, 5, True),
                                                (64, 1, True)])
def test_rows_land_once_at_electron_major_positions(lanes, R, sort_lanes):
    first, n = 1000, 300
    T = _lengths(first, n)
    offsets = torch.as_tensor(np.concatenate([[0], np.cumsum(T)[:-1]]), dtype=torch.int64)
    drv = ContinuationDriver(None, TransportControl(), lanes, R, "cpu", 5.0, sort_lanes=sort_lanes)
    stand = StandIn(drv, first, T)
    drv._launch = stand
    seen = np.zeros(T.sum(), dtype=np.int64)
    got_id = np.full(T.sum(), -1, dtype=np.int64)
    got_k = np.full(T.sum(), -1.0)

    def fill(start, m, pos, dirs, E, clock, keys):
        E[:m] = torch.arange(start, start + m, dtype=torch.float64)

    def consume(rnd):
        slot, canon = canonical_rows(rnd, offsets, first)
        c = canon.numpy()
        np.add.at(seen, c, 1)
        got_id[c] = rnd.rows["seg_id"][slot].numpy()
        got_k[c] = rnd.rows["seg_len"][slot].numpy()

    res = drv.run(first, n, fill, consume)
    assert np.all(seen == 1)
    want_id = np.repeat(np.arange(first, first + n), T)
    want_k = np.concatenate([np.arange(t) for t in T])
    np.testing.assert_array_equal(got_id, want_id)
    np.testing.assert_array_equal(got_k, want_k)
    assert res.n_segments == T.sum()
    h = np.arange(first, first + n)
    codes = np.array([_exit_for(x, t) for x, t in zip(h, T)])
    assert res.counts["n_not_entered"] == int((codes == 5).sum())
    assert res.counts["n_cutoff_stopped"] == int((codes == 0).sum())
    assert res.counts["n_backscattered"] == int((codes == 1).sum())
    assert res.counts["n_transmitted"] == int((codes == 2).sum())
    assert res.counts["n_side_exited"] == int((codes == 3).sum())
    # the packed tail takes longer rounds on fewer lanes, within the slot grid
    assert all(nl * r <= lanes * R for nl, r in stand.launches)
    if lanes * R > 64:
        assert any(r > R for _, r in stand.launches)


def test_refill_gives_each_primary_its_entry_state():
    first, n = 50, 40
    T = np.full(n, 3)
    drv = ContinuationDriver(None, TransportControl(), 8, 2, "cpu", 5.0)
    stand = StandIn(drv, first, T)
    seen = {}

    def launch(n_lanes, R):
        for i in range(n_lanes):
            h = int(drv.history[i])
            if h >= 0 and int(drv.nseg[i]) == 0:
                seen[h] = float(drv.E[i])
                assert float(drv.tau[i]) == -1.0 and int(drv.steps[i]) == 0 and float(drv.stragg[i]) == 0.0
        stand(n_lanes, R)

    drv._launch = launch

    def fill(start, m, pos, dirs, E, clock, keys):
        E[:m] = torch.arange(start, start + m, dtype=torch.float64) * 10.0

    drv.run(first, n, fill)
    assert seen == {h: 10.0 * h for h in range(first, first + n)}


def test_empty_run_and_more_lanes_than_histories():
    drv = ContinuationDriver(None, TransportControl(), 8, 3, "cpu", 5.0)
    stand = StandIn(drv, 0, np.array([], dtype=int))
    drv._launch = stand
    res = drv.run(0, 0, lambda *a: None)
    assert res.n_segments == 0 and res.rounds == 0 and stand.launches == []
    T = np.array([5, 0, 2])
    drv = ContinuationDriver(None, TransportControl(), 8, 3, "cpu", 5.0)
    stand = StandIn(drv, 10, T)
    drv._launch = stand

    def fill(start, m, pos, dirs, E, clock, keys):
        E[:m] = 1.0

    res = drv.run(10, 3, fill)
    assert res.n_segments == 7 and res.counts["n_not_entered"] == 1


def test_driver_refuses_empty_populations():
    with pytest.raises(ValueError):
        ContinuationDriver(None, TransportControl(), 0, 3, "cpu", 5.0)
    with pytest.raises(ValueError):
        ContinuationDriver(None, TransportControl(), 4, 0, "cpu", 5.0)
