"""Bounded continuation rounds on the hip backend.

The reference's continuation driver (``montecarlo/transport/continuation.py``,
module docstring ``:1-27``, ``_drive_continuation_rounds`` ``:506``) keeps a
fixed population of active histories, launches each for at most
``rows_per_round`` rows, and between launches retires finished histories and
refills their lanes from the pending primaries, so live row scratch is
``lanes x rows_per_round`` whatever the track length.  A resumed history
continues exactly as one uninterrupted run.

Here the launch is ``hr_transport_continue``: lane state, the loop-carried
record (open optical depth, stream counter, flight/substep/step counters) and
the slot rows stay on the device between rounds, and each round's rows are
handed to a consumer (a scorer) before the next launch overwrites them.  That
is the streamed transport -> scoring route: a run of 1e8 histories at ~300
rows each never holds more than one round of rows.  Refill order, lane count
and round length change only performance: every row, exit code and final
state is the one ``hr_transport_run`` produces for the same history.

Once no primaries are pending and the live lanes have dropped below a quarter
of the population, the live lanes are packed to the front and take
proportionally longer rounds (the slot grid is reused), so the tail of long
histories costs few launches.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

from .. import _lib
from .core import (
    EXIT_BACKSCATTERED,
    EXIT_CUTOFF_STOPPED,
    EXIT_NOT_ENTERED,
    EXIT_SIDE,
    EXIT_STEP_LIMITED,
    EXIT_TRANSMITTED,
    DeviceTables,
    TransportControl,
    alloc_rows,
    transport_args,
)

EXIT_SUSPENDED = 6


@dataclass
class ContinuationRound:
    """One launch's output, valid until the next launch.

    Lane ``i < n_lanes`` wrote ``seg_count[i]`` rows to slots
    ``[i*rows_per_round, i*rows_per_round + seg_count[i])`` of ``rows``; they are
    rows ``row_start[i] ...`` of history ``history[i]``.
    """

    index: int
    n_lanes: int
    rows_per_round: int
    rows: dict
    seg_count: object
    history: object
    row_start: object
    exit_code: object


@dataclass
class ContinuationResult:
    n_histories: int
    n_segments: int
    rounds: int
    counts: dict = field(default_factory=dict)


class ContinuationDriver:
    """Fixed lane population with refill (``continuation.py:506``).

    ``fill(start, n, pos, dirs, E_keV, clock, keys)`` writes the entry state and
    stream key of global electrons ``[start, start+n)`` into the first ``n``
    rows of the given device tensors.  ``launch`` defaults to
    ``hr_transport_continue``; tests may pass a stand-in with the same
    contract.
    """

    def __init__(self, tables: DeviceTables | None, control: TransportControl, lanes: int, rows_per_round: int,
                 device, E_cut_keV: float, launch=None, tail_fraction: float = 0.25, max_tail_rows: int = 4096,
                 sort_lanes: bool = False):
        import torch

        if lanes < 1 or rows_per_round < 1:
            raise ValueError("continuation needs lanes >= 1 and rows_per_round >= 1")
        self.tables, self.control = tables, control
        self.lanes, self.R = int(lanes), int(rows_per_round)
        self.dev = device
        self.tail_fraction = float(tail_fraction)
        self.max_tail_rows = max(int(max_tail_rows), self.R)
        # order the active lanes by energy before every launch, so a wavefront
        # holds histories of similar step cost (empty lanes last)
        self.sort_lanes = bool(sort_lanes)
        n = self.lanes
        f64 = dict(dtype=torch.float64, device=device)
        i64 = dict(dtype=torch.int64, device=device)
        self.key = torch.zeros(n, dtype=torch.uint64, device=device)
        self.pos = torch.zeros((n, 3), **f64)
        self.dirs = torch.zeros((n, 3), **f64)
        self.E = torch.zeros(n, **f64)
        self.clock = torch.zeros(n, **f64)
        self.E_cut = torch.full((n,), float(E_cut_keV), **f64)
        self.alive = torch.ones(n, dtype=torch.uint8, device=device)
        self.stragg = torch.zeros(n, **f64)
        self.history = torch.full((n,), -1, **i64)
        self.tau = torch.full((n,), -1.0, **f64)
        self.draws = torch.zeros(n, dtype=torch.uint64, device=device)
        self.nseg = torch.zeros(n, **i64)
        self.flight = torch.zeros(n, **i64)
        self.substep = torch.zeros(n, **i64)
        self.steps = torch.zeros(n, dtype=torch.int32, device=device)
        self.seg_count = torch.zeros(n, **i64)
        self.exit_code = torch.full((n,), EXIT_SUSPENDED, dtype=torch.int8, device=device)
        self.rows = alloc_rows(n, self.R, control.energy_model == 1, device)
        self._launch = launch or self._hip_launch
        self._lib = None

    # -- the device launch -------------------------------------------------
    def _hip_launch(self, n_lanes, rows_per_round):
        import torch

        if self._lib is None:
            self._lib = _lib.load()
        a = transport_args(self.tables, self.control, (0, n_lanes, rows_per_round, self.key),
                           (self.alive, self.clock, self.pos, self.dirs, self.E, self.E_cut), self.rows,
                           (self.seg_count, self.exit_code), (self.control.straggle_on, self.stragg))
        c = _lib.Continuation()
        c.history, c.tau, c.draws = self.history.data_ptr(), self.tau.data_ptr(), self.draws.data_ptr()
        c.nseg, c.flight, c.substep = self.nseg.data_ptr(), self.flight.data_ptr(), self.substep.data_ptr()
        c.steps = self.steps.data_ptr()
        stream = torch.cuda.current_stream().cuda_stream
        _lib.check(self._lib.hr_transport_continue(C.byref(a), C.byref(c), C.c_void_p(stream)),
                   "hr_transport_continue")

    # -- lane bookkeeping --------------------------------------------------
    def _refill(self, fill, next_id, end_id, n_active):
        """Give pending primaries to empty lanes among the first ``n_active``."""
        import torch

        free = torch.nonzero(self.history[:n_active] < 0).squeeze(1)
        n_new = min(int(free.numel()), end_id - next_id)
        if n_new <= 0:
            return 0
        lanes = free[:n_new]
        st = self._stage(n_new)
        fill(next_id, n_new, st["pos"], st["dirs"], st["E"], st["clock"], st["key"])
        self.pos.index_copy_(0, lanes, st["pos"][:n_new])
        self.dirs.index_copy_(0, lanes, st["dirs"][:n_new])
        self.E.index_copy_(0, lanes, st["E"][:n_new])
        self.clock.index_copy_(0, lanes, st["clock"][:n_new])
        self.key.view(torch.int64).index_copy_(0, lanes, st["key"][:n_new].view(torch.int64))
        self.history.index_copy_(0, lanes, torch.arange(next_id, next_id + n_new, dtype=torch.int64,
                                                         device=self.dev))
        self.tau.index_fill_(0, lanes, -1.0)
        self.draws.view(torch.int64).index_fill_(0, lanes, 0)
        for t in (self.nseg, self.flight, self.substep, self.steps, self.stragg):
            t.index_fill_(0, lanes, 0)
        self.alive.index_fill_(0, lanes, 1)
        return n_new

    def _stage(self, n):
        import torch

        st = getattr(self, "_staging", None)
        if st is None or st["E"].shape[0] < n:
            m = self.lanes
            st = {"pos": torch.empty((m, 3), dtype=torch.float64, device=self.dev),
                  "dirs": torch.empty((m, 3), dtype=torch.float64, device=self.dev),
                  "E": torch.empty(m, dtype=torch.float64, device=self.dev),
                  "clock": torch.empty(m, dtype=torch.float64, device=self.dev),
                  "key": torch.empty(m, dtype=torch.uint64, device=self.dev)}
            self._staging = st
        return st

    _LANE_STATE = ("key", "pos", "dirs", "E", "clock", "E_cut", "alive", "stragg", "history", "tau", "draws",
                   "nseg", "flight", "substep", "steps")

    def _sort(self, n_active):
        """Permute the first ``n_active`` lanes into descending energy, empty lanes last."""
        import torch

        E = self.E[:n_active].clone()
        E.masked_fill_(self.history[:n_active] < 0, -1.0)
        order = torch.argsort(E, descending=True, stable=True)
        for name in self._LANE_STATE:
            t = getattr(self, name)
            v = t.view(torch.int64) if t.dtype == torch.uint64 else t
            v[:n_active] = v[:n_active].index_select(0, order)

    def _pack(self, n_active):
        """Move the live lanes among the first ``n_active`` to the front; returns their count."""
        import torch

        live = torch.nonzero(self.history[:n_active] >= 0).squeeze(1)
        n_live = int(live.numel())
        if n_live == n_active:
            return n_live
        for name in self._LANE_STATE:
            t = getattr(self, name)
            v = t.view(torch.int64) if t.dtype == torch.uint64 else t
            v[:n_live] = v.index_select(0, live)
        self.history[n_live:n_active] = -1
        return n_live

    # -- the round loop ----------------------------------------------------
    def run(self, first_id: int, n_histories: int, fill, consume=None) -> ContinuationResult:
        """Transport electrons ``[first_id, first_id + n_histories)``; ``consume(round)``
        sees every launch's rows before the next launch."""
        import torch

        next_id, end_id = int(first_id), int(first_id) + int(n_histories)
        self.history.fill_(-1)
        n_active, R = self.lanes, self.R
        codes = torch.zeros(8, dtype=torch.int64, device=self.dev)
        n_rows = torch.zeros((), dtype=torch.int64, device=self.dev)
        rounds = 0
        while True:
            next_id += self._refill(fill, next_id, end_id, n_active)
            if next_id >= end_id:
                n_live = int((self.history[:n_active] >= 0).sum())
                if n_live == 0:
                    break
                if n_live <= self.tail_fraction * n_active:
                    n_active = self._pack(n_active)
                    R = max(self.R, min(self.max_tail_rows, (self.lanes * self.R) // n_active))
            if self.sort_lanes:
                self._sort(n_active)
            self._launch(n_active, R)
            cnt = self.seg_count[:n_active]
            rnd = ContinuationRound(rounds, n_active, R, self.rows, cnt, self.history[:n_active],
                                    self.nseg[:n_active] - cnt, self.exit_code[:n_active])
            if consume is not None:
                consume(rnd)
            rounds += 1
            n_rows += cnt.sum()
            ec = self.exit_code[:n_active]
            done = (self.history[:n_active] >= 0) & (ec != EXIT_SUSPENDED)
            codes += torch.bincount(ec[done].to(torch.int64), minlength=8)[:8]
            self.history[:n_active].masked_fill_(done, -1)
        c = codes.cpu().numpy()
        counts = {
            "n_cutoff_stopped": int(c[EXIT_CUTOFF_STOPPED]),
            "n_backscattered": int(c[EXIT_BACKSCATTERED]),
            "n_transmitted": int(c[EXIT_TRANSMITTED]),
            "n_side_exited": int(c[EXIT_SIDE]),
            "n_step_limited": int(c[EXIT_STEP_LIMITED]),
            "n_not_entered": int(c[EXIT_NOT_ENTERED]),
        }
        return ContinuationResult(int(n_histories), int(n_rows.item()), rounds, counts)


def canonical_rows(rnd: ContinuationRound, offsets, first_id=0):
    """``(slot index, canonical index)`` of a round's rows, where
    ``offsets[h - first_id]`` is the electron-major position of history ``h``'s
    first row -- the order the full-history driver compacts to."""
    import torch

    dev = rnd.seg_count.device
    R = rnd.rows_per_round
    k = torch.arange(R, device=dev)
    live = rnd.history >= 0
    mask = (k[None, :] < rnd.seg_count[:, None]) & live[:, None]
    lane, kk = torch.nonzero(mask, as_tuple=True)
    slot = lane * R + kk
    canon = offsets[rnd.history[lane] - int(first_id)] + rnd.row_start[lane] + kk
    return slot, canon
