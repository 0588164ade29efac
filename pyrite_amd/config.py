"""Run-config knobs of the reference, lowered onto the hip backend.

The reference groups its transport and scoring launch knobs in
``ExecutionConfig`` / ``TransportKernelTuning``
(``src/pyrite/_numerics_groups.py:65-200``).  The same field names are kept
here so a run config written for the reference reads unchanged;
``lower()`` turns one into this backend's objects: the per-electron batching
policy, the transport precision, the schedule (batched full histories or
streamed continuation rounds) and the streamed block size.  Fields with no
hip counterpart are accepted only at their defaults.
"""

from __future__ import annotations

from dataclasses import dataclass, field

SCHEDULES = ("full-history", "continuation", "auto")
THREADS = (32, 64, 128, 256, 512, 1024)
TRANSPORT_CORES = ("auto", "hip", "per-electron")


def _positive(name, value):
    if value is not None and (type(value) is not int or value <= 0):
        raise ValueError(f"{name}={value!r}: expected a positive int or None")


@dataclass(frozen=True)
class TransportKernelTuning:
    """Transport launch and schedule knobs; none of them changes a transport row."""

    schedule: str | None = None
    continuation_rows: int | None = None
    electron_batch: int | None = None
    scratch_mib: int | None = None
    kernel_threads: int | None = None
    precision: str | None = None

    def __post_init__(self):
        if self.schedule not in (None, *SCHEDULES):
            raise ValueError(f"schedule={self.schedule!r}: expected one of {SCHEDULES} or None")
        if self.precision not in (None, "fp64", "mixed"):
            raise ValueError(f"precision={self.precision!r}: expected 'fp64', 'mixed' or None")
        if self.kernel_threads is not None and (type(self.kernel_threads) is not int
                                                or self.kernel_threads not in THREADS):
            raise ValueError(f"kernel_threads={self.kernel_threads!r}: expected one of {THREADS} or None")
        for name in ("continuation_rows", "electron_batch", "scratch_mib"):
            _positive(name, getattr(self, name))
        if self.continuation_rows is not None and self.schedule not in ("continuation", "auto"):
            raise ValueError("continuation_rows needs schedule='continuation' or 'auto'")


@dataclass(frozen=True)
class ExecutionConfig:
    """Transport-core selection and launch sizes of one run (reference field names)."""

    transport_core: str = "auto"
    cxr_chunk: int | None = None
    brem_chunk: int | None = None
    transport: TransportKernelTuning = field(default_factory=TransportKernelTuning)
    cxr_threads: int | None = None
    cxr_energies_per_block: int | None = None
    stream_block_electrons: int | None = None
    stream_memory_mib: int | None = None
    stream_spill_dir: str | None = None
    stream_spill_mib: int | None = None
    characteristic_chunk: int | None = None
    math: str | None = None

    def __post_init__(self):
        if self.transport_core not in TRANSPORT_CORES:
            raise ValueError(f"transport_core={self.transport_core!r}: the hip backend runs {TRANSPORT_CORES}")
        if not isinstance(self.transport, TransportKernelTuning):
            raise TypeError("transport must be a TransportKernelTuning")
        for name in ("cxr_chunk", "brem_chunk", "stream_block_electrons", "stream_memory_mib",
                     "characteristic_chunk"):
            _positive(name, getattr(self, name))
        if self.stream_block_electrons is not None and self.stream_memory_mib is not None:
            raise ValueError("stream_block_electrons and stream_memory_mib exclude each other")
        if self.math not in (None, "reproducible", "fast"):
            raise ValueError(f"math={self.math!r}: expected 'reproducible', 'fast' or None")
        if self.cxr_threads is not None and self.cxr_threads not in THREADS:
            raise ValueError(f"cxr_threads={self.cxr_threads!r}: expected one of {THREADS} or None")
        if self.cxr_energies_per_block not in (None, 1, 2, 3, 4):
            raise ValueError("cxr_energies_per_block must be 1-4 or None")
        # knobs of reference subsystems this backend does not have
        for name in ("stream_spill_dir", "stream_spill_mib", "characteristic_chunk"):
            if getattr(self, name) is not None:
                raise NotImplementedError(f"{name}: the hip backend has no line-grid spill or characteristic "
                                          "lines; leave it None")


@dataclass(frozen=True)
class Lowered:
    """What a run config means on this backend."""

    transport_config: object  # transport.core.PerElectronTransportConfig
    precision: str
    schedule: str  # "full-history" (batched, CasePipeline.step) or "continuation" (CasePipeline.stream)
    rows_per_round: int
    stream_block_electrons: int | None


def lower(cfg: ExecutionConfig, streamed_default_rows: int = 32) -> Lowered:
    """Lower ``cfg`` onto the backend's batching policy, precision and schedule.

    ``kernel_threads`` rounds up to whole 64-lane waves (32 runs as 64: the
    device's wave is 64 lanes); ``scratch_mib`` becomes the slot-row budget;
    ``electron_batch`` caps electrons per launch; ``schedule="auto"`` runs
    full histories (the batched route) unless a streamed block size is set;
    ``math="fast"`` runs the reproducible kernels (the backend has no
    separate fast transport)."""
    from .transport.core import DEFAULT_CONFIG, PerElectronTransportConfig

    t = cfg.transport
    base = DEFAULT_CONFIG
    threads = base.kernel_threads if t.kernel_threads is None else max(64, min(256, t.kernel_threads))
    tc = PerElectronTransportConfig(
        seg_capacity=base.seg_capacity,
        scratch_budget_bytes=base.scratch_budget_bytes if t.scratch_mib is None else t.scratch_mib << 20,
        capacity_headroom=base.capacity_headroom,
        max_batch=base.max_batch if t.electron_batch is None else t.electron_batch,
        probe_electrons=base.probe_electrons,
        kernel_threads=threads,
    )
    schedule = t.schedule or "auto"
    if schedule == "auto":
        schedule = "continuation" if cfg.stream_block_electrons is not None else "full-history"
    rows = streamed_default_rows if t.continuation_rows is None else t.continuation_rows
    return Lowered(tc, t.precision or "fp64", schedule, rows, cfg.stream_block_electrons)
