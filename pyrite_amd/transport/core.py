"""The ``hip`` transport core and its capacity-replay batch driver.

Mirrors the reference's per-electron core contract and driver:

* ``core((e_start, e_count, cap, keys), control, geometry, materials, lut,
  state, scratch, (seg_count, exit_code), (straggle_on, stragg_dE))`` --
  ``src/pyrite/montecarlo/transport/batching.py:547-555`` and the CUDA launcher
  signature ``_jit_launch.py:234-252``.  The core mutates ``pos``, ``dirs``,
  ``E_keV`` and ``clock`` in place, writes row ``s`` of local electron ``i`` to
  slot ``i*cap + s`` only while ``s < cap`` and reports the true count.
* ``drive_batches`` -- ``batching.py:477-628``: snapshot, launch, one
  device-to-host read of ``seg_count.max()``, replay at the measured capacity,
  compaction into electron-major / step-minor rows.

Device memory is torch.cuda tensors; the kernel is ``hr_transport_run`` of
libhiprite.so.  There is no CPU fallback: without the library or a GPU the
core raises.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from .. import _lib
from ..backend import require_device
from .tables import LayerTables, TransportLUT

EXIT_CUTOFF_STOPPED = 0
EXIT_BACKSCATTERED = 1
EXIT_TRANSMITTED = 2
EXIT_SIDE = 3
EXIT_STEP_LIMITED = 4
EXIT_NOT_ENTERED = 5

ROW_FIELDS = (
    "seg_dir", "seg_mid", "seg_len", "seg_E", "seg_t0", "seg_id", "seg_lay",
    "seg_E_end", "seg_t_end", "seg_flight", "seg_substep", "seg_event",
)


class TransportStepLimitError(RuntimeError):
    """A history reached ``max_steps`` (``transport/events.py:19-31``)."""


@dataclass(frozen=True)
class PerElectronTransportConfig:
    """Batching policy (``batching.py:228-290``), sized for 288 GB of HBM.

    No field changes results.  ``seg_capacity`` is the first guess of slots per
    electron; ``scratch_budget_bytes`` bounds the ``(batch, cap)`` row grid;
    ``capacity_headroom`` scales the measured maximum on replay.
    """

    seg_capacity: int = 256
    scratch_budget_bytes: int = 16 * 1024**3
    capacity_headroom: float = 1.25
    max_batch: int | None = 1 << 21
    probe_electrons: int = 0
    kernel_threads: int = 256


DEFAULT_CONFIG = PerElectronTransportConfig()

# float64 dir(3)+mid(3)+len+E+t0, int64 id, int16 layer  (batching.py:398-400)
_ROW_BYTES_FROZEN = 9 * 8 + 8 + 2
# + float64 E_end, t_end; int64 flight, substep; int8 event
_ROW_BYTES_MIDPOINT = _ROW_BYTES_FROZEN + 8 + 8 + 8 + 8 + 1
_STATUS_BYTES = 8 + 1


def _ptr(t) -> int:
    return 0 if t is None else int(t.data_ptr())


class DeviceTables:
    """Layer, element and LUT tables resident on the device (uploaded once)."""

    def __init__(self, layers: LayerTables, lut: TransportLUT, device):
        import torch

        def f64(a):
            return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=device)

        self.layers = layers
        self.lut = lut
        self.internal_bounds = f64(layers.internal_bounds)
        self.L_nel = torch.as_tensor(np.ascontiguousarray(layers.L_nel, dtype=np.int32), device=device)
        self.L_top = f64(layers.L_top)
        self.L_bot = f64(layers.L_bot)
        self.L_Z = f64(layers.L_Z)
        self.L_J = f64(layers.L_J)
        self.L_k = f64(layers.L_k)
        self.L_coeff = f64(layers.L_coeff)
        self.L_E_cross = f64(layers.L_E_cross)
        self.total_rate = f64(lut.total_rate)
        self.dEds = f64(lut.dEds)
        self.inv_beta = f64(lut.inv_beta)
        self.cdf = f64(lut.cdf)
        self.alpha = f64(lut.alpha)
        # FP32 copies for the mixed transport precision (transport/_jit_precision.py)
        self.total_rate_f = self.total_rate.float()
        self.dEds_f = self.dEds.float()
        self.inv_beta_f = self.inv_beta.float()
        self.cdf_f = self.cdf.float()
        self.alpha_f = self.alpha.float()
        # tabulated elastic angles (hr_elsepa), FP64 in both transport precisions
        self.elsepa_args = None
        g = getattr(layers, "elsepa", None)
        if g is not None:
            self._elsepa = {
                "start": torch.as_tensor(np.ascontiguousarray(g.start, dtype=np.int64), device=device),
                "length": torch.as_tensor(np.ascontiguousarray(g.length, dtype=np.int64), device=device),
                "logE": f64(g.logE), "mu": f64(g.mu), "cdf": f64(g.cdf), "pdf": f64(g.pdf),
                "guide": torch.as_tensor(np.ascontiguousarray(g.guide, dtype=np.int32), device=device),
            }
            t = _lib.ElsepaArgs()
            t.on, t.n_mu, t.guide_cells = 1, int(g.mu.size), int(g.guide_cells)
            for k, v in self._elsepa.items():
                setattr(t, k, int(v.data_ptr()))
            self.elsepa_args = t


@dataclass
class TransportControl:
    """``control`` tuple of the core (``cores.py:184``)."""

    max_steps: int = 400_000
    energy_model: int = 1  # 1 midpoint, 0 frozen
    max_dE_frac: float = 0.0
    straggle_on: bool = False
    finite_footprint: bool = False
    width_ang: float = 0.0
    height_ang: float = 0.0
    precision: str = "fp64"  # "fp64" or "mixed" (FP32 table lookups, FP64 state)


def transport_args(tables: DeviceTables, control: TransportControl, run, state, scratch, out, straggling,
                   block_size=0) -> "_lib.TransportArgs":
    """``hr_transport_args`` for the core's positional contract (pointers only)."""
    L, lut = tables.layers, tables.lut
    e_start, e_count, cap, keys = run
    alive, clock, pos, dirs, E_keV, E_cut = state
    seg_count, exit_code = out
    straggle_on, stragg_dE = straggling
    a = _lib.TransportArgs()
    a.e_start, a.e_count, a.cap = int(e_start), int(e_count), int(cap)
    a.stream_key = _ptr(keys)
    a.max_steps = int(control.max_steps)
    a.energy_model = int(control.energy_model)
    a.max_dE_frac = float(control.max_dE_frac)
    a.straggle_on = int(bool(straggle_on))
    a.n_layers = int(L.n_layers)
    a.internal_bounds = _ptr(tables.internal_bounds)
    a.z_total = float(L.z_total)
    a.finite_footprint = int(bool(control.finite_footprint))
    a.width_ang = float(control.width_ang)
    a.height_ang = float(control.height_ang)
    a.L_nel, a.L_top, a.L_bot = _ptr(tables.L_nel), _ptr(tables.L_top), _ptr(tables.L_bot)
    a.max_el = int(L.max_el)
    a.L_Z, a.L_J, a.L_k = _ptr(tables.L_Z), _ptr(tables.L_J), _ptr(tables.L_k)
    a.L_coeff, a.L_E_cross = _ptr(tables.L_coeff), _ptr(tables.L_E_cross)
    a.lut_log_E_min = float(lut.log_E_min)
    a.lut_inv_dlogE = float(lut.inv_dlogE)
    a.lut_n_energy = int(lut.n_energy)
    a.lut_total_rate, a.lut_dEds = _ptr(tables.total_rate), _ptr(tables.dEds)
    a.lut_inv_beta, a.lut_cdf, a.lut_alpha = _ptr(tables.inv_beta), _ptr(tables.cdf), _ptr(tables.alpha)
    a.alive, a.clock, a.pos, a.dirs = _ptr(alive), _ptr(clock), _ptr(pos), _ptr(dirs)
    a.E_keV, a.E_cut, a.stragg_dE = _ptr(E_keV), _ptr(E_cut), _ptr(stragg_dE)
    for name in ROW_FIELDS:
        setattr(a, name, _ptr(scratch.get(name)))
    a.work = _ptr(scratch.get("work"))  # the launch's claim counter (driver-owned scratch)
    a.seg_count, a.exit_code = _ptr(seg_count), _ptr(exit_code)
    a.block_size = int(block_size)
    if control.precision not in ("fp64", "mixed"):
        raise ValueError(f"transport precision {control.precision!r} is not 'fp64' or 'mixed'")
    a.precision = 1 if control.precision == "mixed" else 0
    a.lut_total_rate_f, a.lut_dEds_f = _ptr(tables.total_rate_f), _ptr(tables.dEds_f)
    a.lut_inv_beta_f, a.lut_cdf_f, a.lut_alpha_f = _ptr(tables.inv_beta_f), _ptr(tables.cdf_f), _ptr(tables.alpha_f)
    if tables.elsepa_args is not None:
        a.elsepa = tables.elsepa_args
    return a


def make_hip_transport_core(tables: DeviceTables, control: TransportControl, hard=None):
    """Return ``core(run, state, scratch, out, straggling, stream=None)``.

    ``run = (e_start, e_count, cap, keys)``; ``state = (alive, clock, pos,
    dirs, E_keV, E_cut)`` device tensors indexed by global electron id;
    ``scratch`` a dict of row tensors sized ``e_count*cap`` (``ROW_FIELDS``)
    plus ``"work"``, the launch's 8-byte claim counter (the driver owns every
    buffer; the core allocates nothing);
    ``out = (seg_count, exit_code)``; ``straggling = (straggle_on, stragg_dE)``.
    ``hard`` (a ``hard_events.DeviceHardModes``) switches on the hard-event
    modes; its row columns then ride in ``scratch`` too.
    """
    lib = _lib.load()
    own = {}

    def core(run, state, scratch, out, straggling, stream=None, block_size=0):
        a = transport_args(tables, control, run, state, scratch, out, straggling, block_size)
        if hard is not None and hard.on:
            hard.fill(a.hard, scratch)
        if not a.work:  # a scratch dict without a counter: the core's own (calls of one core run in order)
            if "work" not in own:
                import torch

                own["work"] = torch.zeros(1, dtype=torch.int64, device=tables.total_rate.device)
            a.work = own["work"].data_ptr()
        _lib.check(lib.hr_transport_run(C.byref(a), C.c_void_p(stream)), "hr_transport_run")

    return core


def alloc_rows(m: int, cap: int, midpoint: bool, device, extra=()):
    """Scratch row grid ``(m*cap)`` in the reference's column dtypes, plus
    ``extra`` ``(name, dtype name)`` columns (the hard-event modes' rows)."""
    import torch

    n = int(m) * int(cap)
    rows = {
        "seg_dir": torch.empty((n, 3), dtype=torch.float64, device=device),
        "seg_mid": torch.empty((n, 3), dtype=torch.float64, device=device),
        "seg_len": torch.empty(n, dtype=torch.float64, device=device),
        "seg_E": torch.empty(n, dtype=torch.float64, device=device),
        "seg_t0": torch.empty(n, dtype=torch.float64, device=device),
        "seg_id": torch.empty(n, dtype=torch.int64, device=device),
        "seg_lay": torch.empty(n, dtype=torch.int16, device=device),
    }
    if midpoint:
        rows.update(
            seg_E_end=torch.empty(n, dtype=torch.float64, device=device),
            seg_t_end=torch.empty(n, dtype=torch.float64, device=device),
            seg_flight=torch.empty(n, dtype=torch.int64, device=device),
            seg_substep=torch.empty(n, dtype=torch.int64, device=device),
            seg_event=torch.empty(n, dtype=torch.int8, device=device),
        )
    for name, dt in extra:
        rows[name] = torch.empty(n, dtype=getattr(torch, dt), device=device)
    return rows


_COLUMN_SPEC = (  # (name, dtype name, width) in the reference's column dtypes (batching.py:398-425)
    ("seg_dir", "float64", 3), ("seg_mid", "float64", 3), ("seg_len", "float64", 1), ("seg_E", "float64", 1),
    ("seg_t0", "float64", 1), ("seg_id", "int64", 1), ("seg_lay", "int16", 1),
)
_MIDPOINT_SPEC = (
    ("seg_E_end", "float64", 1), ("seg_t_end", "float64", 1), ("seg_flight", "int64", 1),
    ("seg_substep", "int64", 1), ("seg_event", "int8", 1),
)
_ALIGN = 256
# alignment padding an arena reserves on top of the grid bytes (one pad per column)
ARENA_SLACK = _ALIGN * (len(_COLUMN_SPEC) + len(_MIDPOINT_SPEC))


class RowArena:
    """One flat device byte buffer that every slot-row grid of a driver is carved from.

    The grid of a sub-batch, ``(m, cap)`` slots in the reference's column
    dtypes, is a set of views into the buffer (columns 256-byte aligned), so
    a driver that owns two arenas and uses them in turn holds exactly their
    bytes however long its histories grow: a longer capacity means fewer
    electrons per sub-batch, never a larger allocation.  ``budget_bytes`` is
    what ``_batch_electrons`` may plan against."""

    def __init__(self, nbytes: int, device):
        import torch

        self.nbytes = int(nbytes)
        self.buf = torch.empty(self.nbytes, dtype=torch.uint8, device=device)

    @property
    def budget_bytes(self) -> int:
        return self.nbytes - ARENA_SLACK

    def rows(self, m: int, cap: int, midpoint: bool):
        import torch

        n = int(m) * int(cap)
        spec = _COLUMN_SPEC + (_MIDPOINT_SPEC if midpoint else ())
        out, off = {}, 0
        for name, dt, width in spec:
            dtype = getattr(torch, dt)
            size = torch.empty(0, dtype=dtype).element_size()
            nb = n * width * size
            if off + nb > self.nbytes:
                raise ValueError(f"row grid of {m} x {cap} slots does not fit a {self.nbytes}-byte arena")
            col = self.buf[off:off + nb].view(dtype)
            out[name] = col.view(n, width) if width > 1 else col
            off += (nb + _ALIGN - 1) // _ALIGN * _ALIGN
        return out


def _capacity_for(seen_max, config):
    """``batching.py:462-469``: never below ``seen_max``."""
    return max(1, seen_max, int(np.ceil(seen_max * config.capacity_headroom)))


def _batch_electrons(e, cap, Ne, config, row_bytes, budget=None):
    """``batching.py:431-480``: electrons whose ``(batch, cap)`` grid fits the budget."""
    per = cap * int(row_bytes) + _STATUS_BYTES
    n = int((config.scratch_budget_bytes if budget is None else budget) // per)
    limit = (1 << 62) if config.max_batch is None else int(config.max_batch)
    n = min(max(1, min(n, limit)), Ne - e)
    if e == 0 and config.probe_electrons > 0:
        n = min(n, int(config.probe_electrons))
    return n


# Smallest scratch budget an out-of-memory replay halves down to.
MIN_SCRATCH_BUDGET = 64 << 20


@dataclass
class DriveResult:
    rows: dict  # compacted, electron-major / step-minor device tensors (None with compact=False)
    n_segments: int
    counts: dict  # exit-code tallies
    seg_count: object  # per-electron true segment counts (device int64)
    exit_code: object  # per-electron exit codes (device int8)
    replays: int
    slots: list = None  # compact=False: [(e0, m, cap, slot rows, seg_count)] per sub-batch
    capacity: int = 0  # capacity learned from the largest count seen (a hint for the next call)
    oom_replays: int = 0  # sub-batches replayed at half the scratch budget after a device OOM


def drive_batches(core, keys, state, stragg_dE, control: TransportControl, Ne: int, device,
                  config: PerElectronTransportConfig = DEFAULT_CONFIG, stream=None, compact=True,
                  cap_hint=None, alloc=None, consume=None, extra_columns=()) -> DriveResult:
    """Capacity-replayed batches (``batching.py:477-628``), rows kept on device.

    ``compact=False`` skips the electron-major compaction and returns the slot
    rows of each sub-batch (a scorer can gather them itself); ``cap_hint``
    starts from a capacity learned by an earlier call; ``alloc(e, m, cap,
    midpoint, device)`` (default: fresh tensors) supplies the slot rows of the
    sub-batch starting at electron ``e``, e.g. from a persistent arena.
    ``consume(e, m, cap, rows, seg_count)`` (with ``compact=False``) takes each
    sub-batch as soon as it is transported, before the next one is allocated,
    and the result then holds no slot rows: a driver that recycles two arenas
    keeps its memory bounded whatever the batch size and capacity.

    A device out-of-memory error while a sub-batch is allocated or launched
    (``backend.is_oom_error``) is replayed at half the scratch budget, the
    reference's OOM replay (``_backend.py:179-205``); below
    ``MIN_SCRATCH_BUDGET`` it is raised."""
    import torch

    from ..backend import is_oom_error

    alive, clock, pos, dirs, E_keV, E_cut = state
    midpoint = control.energy_model == 1
    row_bytes = (_ROW_BYTES_MIDPOINT if midpoint else _ROW_BYTES_FROZEN) + sum(
        {"float64": 8, "int16": 2, "int64": 8, "int8": 1}[dt] for _, dt in extra_columns)
    cap = max(1, int(config.seg_capacity if cap_hint is None else cap_hint))
    budget = int(config.scratch_budget_bytes)
    work = torch.zeros(1, dtype=torch.int64, device=device)  # the core's claim counter (this call's own)
    seen_max = 0
    e = 0
    batches = []
    slots = []
    counts_all = []
    exit_all = []
    replays = 0
    oom_replays = 0
    while e < Ne:
        m = _batch_electrons(e, cap, Ne, config, row_bytes, budget)
        sl = slice(e, e + m)
        snap = (pos[sl].clone(), dirs[sl].clone(), E_keV[sl].clone(), clock[sl].clone())
        rows = None
        while True:
            rows = None  # drop a replayed grid before its successor is allocated
            try:
                rows = (alloc_rows(m, cap, midpoint, device, extra_columns) if alloc is None
                        else alloc(e, m, cap, midpoint, device))
                rows.setdefault("work", work)
                seg_count = torch.zeros(m, dtype=torch.int64, device=device)
                exit_code = torch.zeros(m, dtype=torch.int8, device=device)
                core((e, m, cap, keys), state, rows, (seg_count, exit_code),
                     (control.straggle_on, stragg_dE), stream=stream, block_size=config.kernel_threads)
                needed = int(seg_count.max().item())  # the one device sync per batch
            except Exception as exc:  # noqa: BLE001 - only OOM is replayed
                if not is_oom_error(exc) or budget // 2 < MIN_SCRATCH_BUDGET:
                    raise
                rows = None
                if torch.cuda.is_available():
                    torch.cuda.empty_cache()
                budget //= 2
                oom_replays += 1
                pos[sl], dirs[sl], E_keV[sl], clock[sl] = snap
                m = _batch_electrons(e, cap, Ne, config, row_bytes, budget)
                sl = slice(e, e + m)
                snap = tuple(x[:m] for x in snap)
                continue
            seen_max = max(seen_max, needed)
            if needed <= cap:
                break
            replays += 1
            cap = _capacity_for(seen_max, config)
            pos[sl], dirs[sl], E_keV[sl], clock[sl] = snap
            m = _batch_electrons(e, cap, Ne, config, row_bytes, budget)
            sl = slice(e, e + m)
            snap = tuple(x[:m] for x in snap)
        if compact:
            keep = (torch.arange(cap, device=device)[None, :] < seg_count[:, None]).reshape(-1)
            idx = torch.nonzero(keep).squeeze(1)  # one stream compaction shared by every column
            batches.append({k: v.index_select(0, idx) for k, v in rows.items() if k != "work"})
        elif consume is not None:
            consume(e, m, cap, rows, seg_count)
        else:
            slots.append((e, m, cap, rows, seg_count))
        rows = None
        counts_all.append(seg_count)
        exit_all.append(exit_code)
        e += m
        if seen_max > 0:
            cap = _capacity_for(seen_max, config)
    if not compact:
        joined = None
    elif len(batches) == 1:
        joined = batches[0]
    elif batches:
        joined = {k: torch.cat([b[k] for b in batches]) for k in batches[0]}
    else:
        joined = alloc_rows(0, 1, midpoint, device, extra_columns)
    seg_count = torch.cat(counts_all) if counts_all else torch.zeros(0, dtype=torch.int64, device=device)
    exit_code = torch.cat(exit_all) if exit_all else torch.zeros(0, dtype=torch.int8, device=device)
    codes = torch.bincount(exit_code.to(torch.int64), minlength=6).cpu().numpy()
    counts = {
        "n_cutoff_stopped": int(codes[EXIT_CUTOFF_STOPPED]),
        "n_backscattered": int(codes[EXIT_BACKSCATTERED]),
        "n_transmitted": int(codes[EXIT_TRANSMITTED]),
        "n_side_exited": int(codes[EXIT_SIDE]),
        "n_step_limited": int(codes[EXIT_STEP_LIMITED]),
        "n_not_entered": int(codes[EXIT_NOT_ENTERED]),
    }
    n_seg = int(joined["seg_len"].shape[0]) if compact else int(seg_count.sum())
    return DriveResult(joined, n_seg, counts, seg_count, exit_code, replays,
                       None if (consume is not None and not compact) else slots, cap, oom_replays)
