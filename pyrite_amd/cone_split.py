"""Detector-cone splitting of elastic events on the hip transport.

The reference's opt-in variance reduction for detector-bound line yield
(``src/pyrite/_cone_splitting.py`` the case policy,
``src/pyrite/montecarlo/transport/cone_split.py`` the estimator, its copy
records and statistics, ``cores.py:1189-1226`` the call site): at every
elastic event above an energy gate a copy may represent the outcomes inside
the detector's cone ``C = {Omega : Omega . n >= cos theta_c}``, with a
likelihood-ratio weight and a roulette that bounds the copy count; the
primary keeps its analog draw and ends when that draw lands in ``C``.  The
copies are then transported as one analog generation (no further
splitting) and their rows are scored with their weights, so every tally
keeps its expectation while the rare in-cone histories are sampled often.

Here the analog angular law is this transport's screened Rutherford law (the
reference splits ELSEPA tabulated events only); its band probability and
inverse are closed forms (``csrc/cone.hpp``).  The per-history sample unit
is the primary plus its weighted copies (``history_sums``).

``ConeSplitting`` mirrors the reference's case record (field names,
defaults, model string); ``ConeSplit`` its transport request;
``ConeSplitCore`` wraps the hip transport core so each launch also writes
and collects its copies; ``transport_copies`` runs the copy generation.
"""

from __future__ import annotations

import ctypes as C
import math
import warnings
from collections.abc import Mapping
from dataclasses import asdict, dataclass, fields

import numpy as np

from . import _lib

CONE_SPLIT_MODEL = "elastic-cone-split-ladder-v1"
LOW_GAIN_ENERGY_KEV = 3000.0  # below it the reference measured near-unit gain (_cone_splitting.py:30-31)
EXIT_CONE_SPLIT = 6  # exit code of a primary ended at a cone-entering event
_MC2_KEV = 510.99895


class ConeSplitLowGainWarning(UserWarning):
    """Detector-cone splitting requested where the reference measured no gain."""


def _finite_positive(name, value):
    if isinstance(value, bool) or not isinstance(value, (int, float, np.floating, np.integer)) \
            or not math.isfinite(float(value)) or float(value) <= 0.0:
        raise ValueError(f"cone split {name} must be positive and finite")


@dataclass(frozen=True)
class ConeSplitting:
    """Case record of detector-cone splitting (``_cone_splitting.py:57-141`` fields).

    The cone half-angle is ``cone_factor / gamma0`` about the case's detector
    direction; elastic events below ``min_energy_fraction * E0`` stay analog;
    ``angle_scale`` (``u_ref``) turns on the angle roulette (``None``: off);
    ``target_weight`` is the copy roulette target; ``max_copies_per_electron``
    bounds the copies one primary may make (more raise); copy weights are
    snapped to powers of two by the ladder roulette."""

    cone_factor: float = 8.0
    min_energy_fraction: float = 0.4
    angle_scale: float | None = 1.0
    target_weight: float = 1e-3
    max_copies_per_electron: int = 1024
    model: str = CONE_SPLIT_MODEL

    def __post_init__(self):
        if self.model != CONE_SPLIT_MODEL:
            raise ValueError(f"unknown cone split model {self.model!r}")
        for name in ("cone_factor", "target_weight"):
            _finite_positive(name, getattr(self, name))
            object.__setattr__(self, name, float(getattr(self, name)))
        if self.angle_scale is not None:
            _finite_positive("angle_scale", self.angle_scale)
            object.__setattr__(self, "angle_scale", float(self.angle_scale))
        f = self.min_energy_fraction
        if isinstance(f, bool) or not isinstance(f, (int, float, np.floating)) or not (0.0 <= float(f) < 1.0):
            raise ValueError("cone split min_energy_fraction must lie in [0, 1)")
        object.__setattr__(self, "min_energy_fraction", float(f))
        cap = self.max_copies_per_electron
        if isinstance(cap, bool) or not isinstance(cap, (int, np.integer)) or int(cap) < 1:
            raise ValueError("cone split max_copies_per_electron must be a positive integer")
        object.__setattr__(self, "max_copies_per_electron", int(cap))

    def to_dict(self) -> dict:
        return asdict(self)

    @classmethod
    def from_dict(cls, payload: Mapping) -> "ConeSplitting":
        if not isinstance(payload, Mapping):
            raise TypeError("cone_split must be a mapping")
        known = {f.name for f in fields(cls)}
        extra = sorted(set(payload) - known)
        if extra:
            raise ValueError(f"unknown cone split field(s) {extra}")
        return cls(**payload)

    def half_angle_rad(self, E0_keV: float) -> float:
        angle = self.cone_factor / (1.0 + float(E0_keV) / _MC2_KEV)
        if not 0.0 < angle < math.pi / 2:
            raise ValueError(f"cone split half-angle {angle:.3g} rad at {E0_keV:g} keV is not below pi/2; "
                             "lower cone_factor")
        return angle

    def request(self, n_hat, E0_keV: float) -> "ConeSplit":
        """The transport request toward detector direction ``n_hat`` at beam energy ``E0_keV``."""
        return ConeSplit(tuple(float(v) for v in n_hat), self.half_angle_rad(E0_keV), self.target_weight,
                         max_copies_per_electron=self.max_copies_per_electron,
                         min_energy_keV=self.min_energy_fraction * float(E0_keV),
                         angle_scale=math.inf if self.angle_scale is None else self.angle_scale,
                         weight_ladder=True)

    def validate_case(self, cfg) -> None:
        """Refuse case settings whose tallies cannot carry per-history copy weights."""
        refused = {
            "coherent": "the coherent route squares a sum of amplitudes over correlated copies",
            "line_roulette": "line roulette is not composed with copy weights",
        }
        for key, why in refused.items():
            v = getattr(cfg, key, None)
            if v is not None and v is not False:
                raise ValueError(f"detector-cone splitting does not support {key}: {why}")
        if int(getattr(cfg, "depth_bins", 0) or 0) > 0:
            raise ValueError("detector-cone splitting does not support depth_bins: the depth-dose tally takes "
                             "no row weights")
        self.half_angle_rad(cfg.E0_keV)


def warn_low_gain(E0_keV: float) -> None:
    if float(E0_keV) < LOW_GAIN_ENERGY_KEV:
        warnings.warn(f"detector-cone splitting at {float(E0_keV):g} keV: below {LOW_GAIN_ENERGY_KEV:g} keV the "
                      "detector-cone population is not rare and the copies cost transport time without reducing "
                      "variance", ConeSplitLowGainWarning, stacklevel=3)


@dataclass(frozen=True)
class ConeSplit:
    """Transport request (``cone_split.py:61-117``): detector direction, cone
    half-angle, roulette target weight, copy bound, energy gate, angle-roulette
    scale (``inf``: off) and the power-of-two weight ladder."""

    n_hat: tuple
    half_angle_rad: float
    target_weight: float
    max_copies_per_electron: int = 64
    min_energy_keV: float = 0.0
    angle_scale: float = math.inf
    weight_ladder: bool = False

    def __post_init__(self):
        n = np.asarray(self.n_hat, dtype=np.float64)
        norm = float(np.linalg.norm(n)) if n.shape == (3,) else float("nan")
        if n.shape != (3,) or not math.isfinite(norm) or norm == 0.0:
            raise ValueError("n_hat must be a finite nonzero three-vector")
        object.__setattr__(self, "n_hat", tuple(float(v) for v in n / norm))
        if not 0.0 < self.half_angle_rad < math.pi / 2:
            raise ValueError("half_angle_rad must lie in (0, pi/2)")
        if not (math.isfinite(self.target_weight) and self.target_weight > 0.0):
            raise ValueError("target_weight must be positive and finite")
        if type(self.max_copies_per_electron) is not int or self.max_copies_per_electron < 1:
            raise ValueError("max_copies_per_electron must be a positive integer")
        if not (math.isfinite(self.min_energy_keV) and self.min_energy_keV >= 0.0):
            raise ValueError("min_energy_keV must be finite and non-negative")
        if not self.angle_scale > 0.0:
            raise ValueError("angle_scale must be positive (inf disables it)")
        if type(self.weight_ladder) is not bool:
            raise ValueError("weight_ladder must be a bool")

    def fill(self, c: "_lib.ConeSplitArgs") -> None:
        """The request's scalar fields of ``hr_cone_split`` (slots are the caller's)."""
        c.on = 1
        c.nx, c.ny, c.nz = self.n_hat
        c.cos_c, c.sin_c = math.cos(self.half_angle_rad), math.sin(self.half_angle_rad)
        c.target_weight = float(self.target_weight)
        c.min_energy_keV = float(self.min_energy_keV)
        c.angle_scale = 0.0 if math.isinf(self.angle_scale) else float(self.angle_scale)


def history_sums(primary, copies, copy_parent, copy_weight, n_histories):
    """Per-history tally ``X_i = x_i + sum_{c of i} w_c x_c`` (``cone_split.py:338-352``)."""
    total = np.array(primary, dtype=np.float64, copy=True)
    if total.shape != (int(n_histories),):
        raise ValueError("primary needs one value per history")
    if np.size(copies):
        total += np.bincount(np.asarray(copy_parent, dtype=np.int64),
                             weights=np.asarray(copy_weight, dtype=np.float64) * np.asarray(copies, dtype=np.float64),
                             minlength=int(n_histories))
    return total


def weighted_history_statistics(values):
    """Mean, relative standard error, largest share and Kish effective sample
    size of per-history sums (``cone_split.py:355-378``); copies of one history
    are correlated, so the history is the sample unit."""
    x = np.asarray(values, dtype=np.float64)
    n = x.size
    total = float(x.sum())
    mean = total / n if n else 0.0
    sd = float(x.std(ddof=1)) if n > 1 else 0.0
    sq = float(np.dot(x, x))
    return {
        "n": int(n),
        "mean": mean,
        "relative_se": sd / (math.sqrt(n) * abs(mean)) if n > 1 and mean else None,
        "max_share": float(np.max(np.abs(x)) / abs(total)) if total else None,
        "effective_sample_size": total * total / sq if sq else 0.0,
    }


def _stream():
    import torch

    return torch.cuda.current_stream().cuda_stream


def split_keys(stream_keys, out=None):
    """Split-decision keys ``splitmix64(stream_key ^ 0xC2B2AE3D27D4EB4F)`` on the device."""
    import torch

    if out is None:
        out = torch.empty_like(stream_keys)
    _lib.call("hr_cone_split_keys", stream_keys.data_ptr(), int(stream_keys.numel()), out.data_ptr(), _stream())
    return out


class ConeSplitCore:
    """``core(run, state, scratch, out, straggling, stream=None, block_size=0)``
    with detector-cone splitting on: each launch writes its copies into
    per-electron slots and collects them (ladder, copy keys) into compact
    records kept per sub-batch start (a replay of a sub-batch replaces its
    record).  When a primary makes more copies than the slots hold, the
    launch is rerun from its entry state with room for them (up to
    ``max_copies_per_electron``; beyond it the call raises, as the
    reference's copy buffer does).  ``keys`` must be the stream keys the
    launch reads (global electron index)."""

    def __init__(self, tables, control, cone: ConeSplit, device, slots_per_electron: int = 16):
        from .transport.core import transport_args

        self._args = transport_args
        self.tables, self.control, self.cone, self.dev = tables, control, cone, device
        self.k = max(1, min(int(slots_per_electron), cone.max_copies_per_electron))
        self.records = {}
        self._work = None
        self._skeys = None
        self._slots = None
        self.reruns = 0

    def reset(self):
        self.records = {}

    def _slot_arrays(self, m):
        import torch

        need = m * self.k
        if self._slots is None or self._slots["E_keV"].numel() < need or self._slots["count"].numel() < m:
            n = max(need, 1)
            f64 = dict(dtype=torch.float64, device=self.dev)
            self._slots = {
                "pos": torch.empty((n, 3), **f64), "dir": torch.empty((n, 3), **f64),
                "E_keV": torch.empty(n, **f64), "clock": torch.empty(n, **f64), "weight": torch.empty(n, **f64),
                "flight": torch.empty(n, dtype=torch.int64, device=self.dev),
                "count": torch.empty(max(m, 1), dtype=torch.int64, device=self.dev),
                "killed": torch.empty(max(m, 1), dtype=torch.uint8, device=self.dev),
            }
        return self._slots

    def _cone_args(self, s, skeys):
        c = _lib.ConeSplitArgs()
        self.cone.fill(c)
        c.split_key, c.cap = skeys.data_ptr(), self.k
        for name in ("count", "killed", "pos", "dir", "E_keV", "clock", "weight", "flight"):
            setattr(c, name, s[name].data_ptr())
        return c

    def __call__(self, run, state, scratch, out, straggling, stream=None, block_size=0):
        import torch

        e_start, e_count, cap, keys = run
        e_start, e_count = int(e_start), int(e_count)
        alive, clock, pos, dirs, E_keV, E_cut = state
        lib = _lib.load()
        if self._skeys is None or self._skeys.numel() < keys.numel():
            self._skeys = torch.empty_like(keys)
        skeys = self._skeys[:keys.numel()]
        _lib.call("hr_cone_split_keys", keys.data_ptr(), int(keys.numel()), skeys.data_ptr(), stream)
        if self._work is None:
            self._work = torch.zeros(1, dtype=torch.int64, device=self.dev)
        sl = slice(e_start, e_start + e_count)
        snap = (pos[sl].clone(), dirs[sl].clone(), E_keV[sl].clone(), clock[sl].clone())
        while True:
            s = self._slot_arrays(e_count)
            a = self._args(self.tables, self.control, run, state, scratch, out, straggling, block_size)
            if not a.work:
                a.work = self._work.data_ptr()
            a.cone = self._cone_args(s, skeys)
            _lib.check(lib.hr_transport_run(C.byref(a), C.c_void_p(stream)), "hr_transport_run")
            cnt = s["count"][:e_count]
            stats = torch.stack([cnt.clamp(max=self.k).sum(), cnt.max()]).cpu() if e_count else torch.zeros(2)
            bound, biggest = int(stats[0]), int(stats[1])
            if biggest > self.cone.max_copies_per_electron:
                raise RuntimeError(f"detector-cone splitting made {biggest} copies for one electron, above "
                                   f"max_copies_per_electron={self.cone.max_copies_per_electron}; raise it or "
                                   "the target weight")
            if biggest <= self.k:
                break
            self.reruns += 1
            self.k = min(self.cone.max_copies_per_electron, 1 << (biggest - 1).bit_length())
            pos[sl], dirs[sl], E_keV[sl], clock[sl] = snap
        f64 = dict(dtype=torch.float64, device=self.dev)
        n = max(bound, 1)
        rec = {
            "parent": torch.empty(n, dtype=torch.int64, device=self.dev),
            "flight": torch.empty(n, dtype=torch.int64, device=self.dev),
            "pos": torch.empty((n, 3), **f64), "dir": torch.empty((n, 3), **f64),
            "E_keV": torch.empty(n, **f64), "clock": torch.empty(n, **f64), "weight": torch.empty(n, **f64),
            "stream_key": torch.empty(n, dtype=torch.uint64, device=self.dev),
            "n_out": torch.zeros(2, dtype=torch.int64, device=self.dev),
            "killed": s["killed"][:e_count].clone(),
        }
        ca = _lib.ConeCollectArgs()
        ca.cone = C.pointer(a.cone)
        ca.e_start, ca.e_count = e_start, e_count
        ca.ladder, ca.out_cap = int(self.cone.weight_ladder), bound
        for name in ("parent", "flight", "pos", "dir", "E_keV", "clock", "weight", "stream_key", "n_out"):
            setattr(ca, name, rec[name].data_ptr())
        nbytes = int(lib.hr_cone_collect_workspace_bytes(e_count))
        ws = torch.empty(max(nbytes, 256) // 8, dtype=torch.int64, device=self.dev)
        ca.ws.ptr, ca.ws.bytes = ws.data_ptr(), ws.numel() * 8
        _lib.check(lib.hr_cone_collect(C.byref(ca), C.c_void_p(stream)), "hr_cone_collect")
        rec["e_start"] = e_start  # parent is local to the sub-batch; copies() makes it global
        rec["_ws"] = ws
        self.records[e_start] = rec

    def copies(self):
        """Every collected copy of the calls since ``reset()``, in (electron, event)
        order; ``parent`` indexes the primaries of the driven run (each
        sub-batch's local parent plus the sub-batch start)."""
        import torch

        recs = [self.records[k] for k in sorted(self.records)]
        if not recs:
            return None
        n = torch.stack([r["n_out"][0] for r in recs]).cpu().numpy()
        parts = {k: [] for k in ("parent", "flight", "pos", "dir", "E_keV", "clock", "weight", "stream_key")}
        for r, k in zip(recs, n):
            k = int(k)
            parts["parent"].append(r["parent"][:k] + r["e_start"])
            for name in ("flight", "pos", "dir", "E_keV", "clock", "weight", "stream_key"):
                parts[name].append(r[name][:k])
        out = {name: torch.cat(v) for name, v in parts.items()}
        out["killed"] = torch.cat([r["killed"] for r in recs])
        return out


def transport_copies(copies, tables, control, E_cut_keV, device, config=None, consume=None, compact=False):
    """The copy generation: every copy transported once, analog, from its
    event state with its own stream key and its parent's cutoff
    (``cone_split.py:447-513``).  Returns the ``drive_batches`` result;
    ``consume(e0, m, cap, rows, seg_count)`` takes each sub-batch's slot rows
    (copy ``e0 + i`` of the sub-batch carries ``copies["weight"][e0 + i]``)."""
    import torch

    from .transport.core import DEFAULT_CONFIG, drive_batches, make_hip_transport_core

    n = int(copies["weight"].numel())
    core = make_hip_transport_core(tables, control)
    state = (torch.ones(n, dtype=torch.uint8, device=device), copies["clock"].clone(), copies["pos"].clone(),
             copies["dir"].clone(), copies["E_keV"].clone(),
             torch.full((n,), float(E_cut_keV), dtype=torch.float64, device=device))
    stragg = torch.zeros(n, dtype=torch.float64, device=device)
    return drive_batches(core, copies["stream_key"], state, stragg, control, n, device, config or DEFAULT_CONFIG,
                         stream=_stream(), compact=compact, consume=consume)


__all__ = [
    "CONE_SPLIT_MODEL",
    "ConeSplit",
    "ConeSplitCore",
    "ConeSplitLowGainWarning",
    "ConeSplitting",
    "EXIT_CONE_SPLIT",
    "history_sums",
    "split_keys",
    "transport_copies",
    "warn_low_gain",
    "weighted_history_statistics",
]
