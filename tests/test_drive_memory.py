"""Memory bounds of the capacity-replay batch driver (CPU tensors, a fake core).

``drive_batches`` (``pyrite_amd/transport/core.py``) follows the reference's
driver (``transport/batching.py:428-628``): a sub-batch whose longest history
outgrows the slot capacity is replayed at 1.25x the measured maximum, with
fewer electrons so the grid stays inside the scratch budget.  These tests run
it on host tensors with a core that only reports row counts, so they check the
planning, the replay and the memory held, not the physics.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from pyrite_amd.transport.core import (
    ARENA_SLACK,
    MIN_SCRATCH_BUDGET,
    PerElectronTransportConfig,
    RowArena,
    TransportControl,
    _ROW_BYTES_MIDPOINT,
    drive_batches,
)

LONG = 8192  # rows of the rare long histories


def row_count(ids):
    """Rows of each global electron id: 41 for most, 8192 for one id in 997."""
    return np.where(ids % 997 == 3, LONG, 41)


def fake_core(counts_seen):
    def core(run, state, scratch, out, straggling, stream=None, block_size=0):
        e_start, e_count, cap, _keys = run
        seg_count, exit_code = out
        ids = np.arange(e_start, e_start + e_count)
        seg_count.copy_(torch.as_tensor(row_count(ids), dtype=torch.int64))
        exit_code.fill_(2)
        # touch every slot the grid claims, so the grid really has m*cap rows
        assert scratch["seg_len"].shape[0] == e_count * cap
        counts_seen.append((e_start, e_count, cap))

    return core


def state(n):
    z = torch.zeros(n, dtype=torch.float64)
    return (torch.ones(n, dtype=torch.uint8), z.clone(), torch.zeros((n, 3), dtype=torch.float64),
            torch.zeros((n, 3), dtype=torch.float64), z.clone() + 5000.0, z.clone() + 5.0)


def two_arenas(nbytes):
    """The pipeline's allocation policy: two arenas used in turn, one per sub-batch."""
    arenas = [RowArena(nbytes, "cpu"), RowArena(nbytes, "cpu")]
    turn = {"k": 1, "e": None}

    def alloc(e, m, cap, midpoint, device):
        if e != turn["e"]:
            turn["e"] = e
            turn["k"] ^= 1
        return arenas[turn["k"]].rows(m, cap, midpoint)

    return arenas, alloc


def test_arena_grid_never_outgrows_two_arenas():
    Ne = 20_000
    arena = 64 << 20
    cfg = PerElectronTransportConfig(seg_capacity=256, scratch_budget_bytes=arena - ARENA_SLACK)
    arenas, alloc = two_arenas(arena)
    seen = []
    consumed = []

    def consume(e0, m, cap, rows, seg_count):
        assert rows["seg_len"].shape[0] == m * cap
        assert int(seg_count.max()) <= cap
        consumed.append((e0, m, cap, seg_count.clone()))

    res = drive_batches(fake_core(seen), torch.zeros(Ne, dtype=torch.uint64), state(Ne),
                        torch.zeros(Ne, dtype=torch.float64), TransportControl(), Ne, "cpu", cfg,
                        compact=False, alloc=alloc, consume=consume)
    # every grid the core saw fit one arena's budget
    for _e, m, cap in seen:
        assert m * cap * _ROW_BYTES_MIDPOINT <= arena - ARENA_SLACK
    # a long history forced a replay at 1.25 x its rows, and the later sub-batches keep that capacity
    assert res.replays >= 1
    assert res.capacity == int(np.ceil(LONG * 1.25))
    # sub-batches tile [0, Ne) in order, and the reported counts are the true ones
    starts = [c[0] for c in consumed]
    assert starts == sorted(starts) and starts[0] == 0
    assert sum(c[1] for c in consumed) == Ne
    np.testing.assert_array_equal(res.seg_count.numpy(), row_count(np.arange(Ne)))
    assert res.slots is None  # consumed sub-batches are not held
    assert res.n_segments == int(row_count(np.arange(Ne)).sum())
    # the memory the driver's grids live in is exactly the two arenas
    assert sum(a.nbytes for a in arenas) == 2 * arena


def test_fresh_grids_are_dropped_before_a_replay_allocates():
    """Without arenas every grid is fresh: the replayed grid must not coexist with
    the one it replaces, so at most one grid is alive while a sub-batch runs."""
    import gc
    import weakref

    Ne = 1_200
    cfg = PerElectronTransportConfig(seg_capacity=64, scratch_budget_bytes=96 << 20)
    alive = []

    def alloc(e, m, cap, midpoint, device):
        gc.collect()
        live = [r for r in alive if r() is not None]
        assert not live, f"{len(live)} earlier grids still referenced when a new one is allocated"
        rows = {"seg_len": torch.empty(m * cap, dtype=torch.float64)}
        alive.append(weakref.ref(rows["seg_len"]))
        return rows

    seen = []
    res = drive_batches(fake_core(seen), torch.zeros(Ne, dtype=torch.uint64), state(Ne),
                        torch.zeros(Ne, dtype=torch.float64), TransportControl(), Ne, "cpu", cfg,
                        compact=False, alloc=alloc, consume=lambda *a: None)
    assert res.replays >= 1


def test_out_of_memory_replays_at_half_the_budget():
    """A device OOM while a sub-batch's grid is allocated replays it at half the
    scratch budget (``_backend.py:179-205``), and the run still covers every electron."""
    Ne = 6_000
    cfg = PerElectronTransportConfig(seg_capacity=128, scratch_budget_bytes=512 << 20)
    limit = 150 << 20  # what the "device" can hold for one grid

    def alloc(e, m, cap, midpoint, device):
        if m * cap * _ROW_BYTES_MIDPOINT > limit:
            raise torch.cuda.OutOfMemoryError("HIP out of memory. Tried to allocate a slot grid")
        return {"seg_len": torch.empty(m * cap, dtype=torch.float64)}

    seen = []
    res = drive_batches(fake_core(seen), torch.zeros(Ne, dtype=torch.uint64), state(Ne),
                        torch.zeros(Ne, dtype=torch.float64), TransportControl(), Ne, "cpu", cfg,
                        compact=False, alloc=alloc, consume=lambda *a: None)
    assert res.oom_replays >= 1
    np.testing.assert_array_equal(res.seg_count.numpy(), row_count(np.arange(Ne)))
    for _e, m, cap in seen:
        assert m * cap * _ROW_BYTES_MIDPOINT <= limit


def test_out_of_memory_below_the_floor_is_raised():
    Ne = 100
    cfg = PerElectronTransportConfig(seg_capacity=64, scratch_budget_bytes=MIN_SCRATCH_BUDGET)

    def alloc(e, m, cap, midpoint, device):
        raise torch.cuda.OutOfMemoryError("HIP out of memory")

    with pytest.raises(torch.cuda.OutOfMemoryError):
        drive_batches(fake_core([]), torch.zeros(Ne, dtype=torch.uint64), state(Ne),
                      torch.zeros(Ne, dtype=torch.float64), TransportControl(), Ne, "cpu", cfg,
                      compact=False, alloc=alloc, consume=lambda *a: None)


def test_other_errors_are_not_replayed():
    Ne = 100

    def alloc(e, m, cap, midpoint, device):
        raise ValueError("bad grid")

    with pytest.raises(ValueError):
        drive_batches(fake_core([]), torch.zeros(Ne, dtype=torch.uint64), state(Ne),
                      torch.zeros(Ne, dtype=torch.float64), TransportControl(), Ne, "cpu",
                      PerElectronTransportConfig(), compact=False, alloc=alloc, consume=lambda *a: None)


def test_arena_views_are_disjoint_and_aligned():
    a = RowArena(8 << 20, "cpu")
    rows = a.rows(100, 50, True)
    spans = []
    for name, t in rows.items():
        start = t.data_ptr() - a.buf.data_ptr()
        assert start % 256 == 0, name
        spans.append((start, start + t.numel() * t.element_size()))
        assert t.shape[0] == 5000
    spans.sort()
    for (s0, e0), (s1, _e1) in zip(spans, spans[1:]):
        assert e0 <= s1
    with pytest.raises(ValueError):
        a.rows(10_000, 100, True)
