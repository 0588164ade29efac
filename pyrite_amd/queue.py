"""Many beam/target configurations through one persistent work-queue kernel.

Host side of ``hr_queue_run`` / ``hr_queue_graph_*``.  The reference runs a
campaign as a loop of per-case transports (``montecarlo/runner/case_driver.py:
142-152``); ``TransportQueue`` uploads every configuration's tables once, lays
the histories of all configurations end to end as one queue, and runs them in
one launch (optionally a captured hipGraph replayed per campaign).
``autotune`` picks the launch block size and the histories a lane claims at
once by timing a short sample of the queue.
"""

from __future__ import annotations

import ctypes as C
import json
import time
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from . import _lib
from .spectrum.geometry import tilted_geometry
from .transport.core import DeviceTables, TransportControl, transport_args
from .transport.tables import build_layer_tables, build_transport_lut

MATERIALS = {
    "graphite": [("C", 0.1136)],
    "Si": [("Si", 0.04994)],
    "Ge": [("Ge", 0.0442)],
    "W": [("W", 0.0632)],
}

TALLY_COLUMNS = ("cutoff", "backscattered", "transmitted", "side", "step_limited", "not_entered", "suspended",
                 "rows", "rows2", "E_transmitted_keV", "E_backscattered_keV", "E_deposited_keV")


@dataclass
class QueueCase:
    """One configuration of the campaign: a pencil beam on a slab."""

    Ne: int
    thickness_ang: float
    E0_keV: float
    material: str
    tilt_deg: float = 0.0
    seed: int = 1
    E_cut_keV: float = 5.0
    straggling: bool = True
    composition: list = field(default=None)

    def __post_init__(self):
        if self.composition is None:
            self.composition = list(MATERIALS[self.material])


def campaign(n_cfg=256, seed=2024, scale=1.0):
    """Uneven campaign: Ne log-uniform 1e4-1e7 (times ``scale``), thickness
    log-uniform 10 um - 2 mm, E0 uniform 1-30 MeV, material from graphite /
    Si / Ge / W, tilt uniform 0-3 deg; one seed per configuration."""
    rng = np.random.default_rng(seed)
    out = []
    mats = list(MATERIALS)
    for c in range(n_cfg):
        Ne = int(round(np.exp(rng.uniform(np.log(1e4), np.log(1e7))) * scale))
        out.append(QueueCase(
            Ne=max(1, Ne),
            thickness_ang=float(np.exp(rng.uniform(np.log(1e5), np.log(2e7)))),
            E0_keV=float(rng.uniform(1000.0, 30000.0)),
            material=mats[int(rng.integers(0, len(mats)))],
            tilt_deg=float(rng.uniform(0.0, 3.0)),
            seed=int(rng.integers(1, 2**31)),
        ))
    return out


class TransportQueue:
    """Device-resident campaign: per-configuration tables, config array, offsets, tallies."""

    def __init__(self, cases, device, max_steps=400_000, precision="fp64"):
        import torch

        if not 1 <= len(cases) <= 1024:
            raise ValueError("a queue holds 1..1024 configurations")
        self.cases = list(cases)
        self.device = device
        self.lib = _lib.load()
        self._keep = []
        arr = (_lib.QueueConfig * len(cases))()
        for i, q in enumerate(self.cases):
            layers = build_layer_tables([(0.0, q.thickness_ang, list(q.composition))])
            lut = build_transport_lut(q.E_cut_keV, q.E0_keV, layers)
            tabs = DeviceTables(layers, lut, device)
            ctl = TransportControl(max_steps=max_steps, energy_model=1, straggle_on=q.straggling,
                                   precision=precision)
            self._keep.append(tabs)
            t = transport_args(tabs, ctl, (0, 0, 0, None), (None,) * 6, {}, (None, None), (q.straggling, None))
            arr[i].t = t
            arr[i].seed = int(q.seed)
            arr[i].E0_keV = float(q.E0_keV)
            arr[i].E_cut_keV = float(q.E_cut_keV)
            bd, _ = tilted_geometry(np.pi / 2, np.deg2rad(q.tilt_deg))
            for k in range(3):
                arr[i].beam_dir[k] = float(bd[k])
        raw = bytes(arr)
        self.configs = torch.frombuffer(bytearray(raw), dtype=torch.uint8).to(device)
        counts = np.array([q.Ne for q in self.cases], dtype=np.int64)
        self.offsets_host = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
        self.offsets = torch.as_tensor(self.offsets_host, device=device)
        self.total = int(self.offsets_host[-1])
        self.tally = torch.zeros((len(cases), _lib.QUEUE_TALLY), dtype=torch.float64, device=device)
        self.block_size = 256
        self.per_claim = 1
        self.precision = 1 if precision == "mixed" else 0
        self.counter = torch.zeros(1, dtype=torch.int64, device=device)  # claim counter of eager launches

    def launch_args(self, offsets=None, total=None, tally=None, block_size=None, per_claim=None, accumulate=0,
                    counter=None):
        L = _lib.QueueLaunch()
        L.configs = self.configs.data_ptr()
        L.counter = (self.counter if counter is None else counter).data_ptr()
        L.n_cfg = len(self.cases)
        L.offsets = (self.offsets if offsets is None else offsets).data_ptr()
        L.total = self.total if total is None else int(total)
        L.tally = (self.tally if tally is None else tally).data_ptr()
        L.accumulate = int(accumulate)
        L.block_size = int(self.block_size if block_size is None else block_size)
        L.per_claim = int(self.per_claim if per_claim is None else per_claim)
        L.precision = int(self.precision)
        return L

    def run(self, **kw):
        """Every history of every configuration, one launch; returns the tally (device)."""
        import torch

        L = self.launch_args(**kw)
        s = torch.cuda.current_stream().cuda_stream
        _lib.check(self.lib.hr_queue_run(C.byref(L), C.c_void_p(s)), "hr_queue_run")
        return self.tally

    def graph(self, **kw):
        """A captured hipGraph of the whole campaign (reset + queue kernel) with its
        own claim counter, so a replay may run beside an eager launch."""
        import torch

        counter = torch.zeros(1, dtype=torch.int64, device=self.device)
        return QueueGraph(self, self.launch_args(counter=counter, **kw), counter)

    def sample_offsets(self, fraction, minimum=64):
        """Offsets of the queue's leading ``fraction`` of each configuration."""
        counts = np.diff(self.offsets_host)
        sub = np.minimum(counts, np.maximum(minimum, np.ceil(counts * fraction).astype(np.int64)))
        return np.concatenate([[0], np.cumsum(sub)]).astype(np.int64)

    def autotune(self, fraction=0.02, blocks=(64, 128, 256), claims=(1, 2, 4, 8), cache=None):
        """Time each (block size, per-claim) on a sample of the queue; keep the fastest."""
        import torch

        key = None
        if cache is not None:
            props = torch.cuda.get_device_properties(self.device)
            key = f"{props.name}|{len(self.cases)}|{self.total}|{fraction}"
            path = Path(cache)
            if path.exists():
                table = json.loads(path.read_text())
                if key in table:
                    self.block_size, self.per_claim = table[key]["best"]
                    return table[key]
        off = self.sample_offsets(fraction)
        off_d = torch.as_tensor(off, device=self.device)
        tally = torch.zeros_like(self.tally)
        results = []
        self.run(offsets=off_d, total=int(off[-1]), tally=tally)  # warm
        for b in blocks:
            for k in claims:
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                self.run(offsets=off_d, total=int(off[-1]), tally=tally, block_size=b, per_claim=k)
                torch.cuda.synchronize()
                results.append(((b, k), time.perf_counter() - t0))
        best = min(results, key=lambda r: r[1])[0]
        self.block_size, self.per_claim = best
        rec = {"best": list(best), "timings": [[list(bk), t] for bk, t in results], "sample_histories": int(off[-1])}
        if cache is not None:
            path = Path(cache)
            table = json.loads(path.read_text()) if path.exists() else {}
            table[key] = rec
            path.write_text(json.dumps(table, indent=1))
        return rec


class QueueGraph:
    def __init__(self, queue: TransportQueue, launch, counter):
        self.queue = queue
        self.launch_struct = launch
        self.counter = counter  # the graph's claim counter; held as long as the graph
        self.handle = C.c_void_p()
        _lib.check(queue.lib.hr_queue_graph_create(C.byref(launch), C.byref(self.handle)), "hr_queue_graph_create")

    def replay(self):
        import torch

        s = torch.cuda.current_stream().cuda_stream
        _lib.check(self.queue.lib.hr_queue_graph_launch(self.handle, C.c_void_p(s)), "hr_queue_graph_launch")

    def close(self):
        if self.handle:
            _lib.check(self.queue.lib.hr_queue_graph_destroy(self.handle), "hr_queue_graph_destroy")
            self.handle = C.c_void_p()

    def __del__(self):  # pragma: no cover - best effort
        try:
            self.close()
        except Exception:
            pass
