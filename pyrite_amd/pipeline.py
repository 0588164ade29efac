"""Transport -> scoring pipeline for one beam/target case, device resident.

The reference runs one case as ``_transport_case`` then ``_spectrum_case``
(``montecarlo/runner/case_driver.py:142-152``) with the segment arrays kept on
the device between them (``keep_segments_on_device``).  ``CasePipeline`` is
that contract on the hip backend: tables are built and uploaded once, every
batch of histories is transported by ``hr_transport_run`` and scored by
``hr_cxr_lines`` + ``hr_sincsq_reduce`` and ``hr_brem_reduce`` into FP64 device
tallies, and histories are addressed by global electron id, so batches
``[s*Ne, (s+1)*Ne)`` are slices of one large run.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .backend import require_device
from .spectrum.geometry import cone_detector, tilted_geometry
from .spectrum.lines import BremScorer, LineScorer, line_tables, segments_fp32
from .transport.beam import BeamSpec, fill_entry, upper_beam_energy_keV
from .transport.core import (
    DeviceTables,
    PerElectronTransportConfig,
    TransportControl,
    drive_batches,
    make_hip_transport_core,
)
from .transport.tables import build_layer_tables, build_transport_lut


def bin_centers(lo_eV, hi_eV, n):
    edges = np.linspace(lo_eV, hi_eV, n + 1)
    return 0.5 * (edges[:-1] + edges[1:])


@dataclass
class CaseConfig:
    """One plain-method case (defaults: 5 MeV pencil beam on 100 um HOPG, (002)).

    ``silicon_1mm_10MeV()`` builds the 1 mm Si (111) case with a Gaussian
    Twiss beam, 16 reflections and 4096 bins."""

    E0_keV: float = 5000.0
    Ne: int = 100_000
    thickness_ang: float = 1.0e6
    composition: list = field(default_factory=lambda: [("C", 0.1136)])
    crystal: str = "hopg"
    hkl_list: list = field(default_factory=lambda: [(0, 0, 2)])
    tilt_deg: float = 0.5
    theta_obs_deg: float = 90.0
    cone_half_angle_rad: float = 0.01
    n_side: int = 3
    n_bins: int = 1024
    E_lo_eV: float = 1000.0
    E_hi_eV: float = 50000.0
    E_cut_keV: float = 5.0
    seed: int = 1
    straggling: bool = True
    max_steps: int = 400_000
    beam: BeamSpec = field(default_factory=BeamSpec)
    precision: str = "fp64"  # transport precision: "fp64" or "mixed"
    line_route: str = "dense"  # "dense" sinc^2 over every bin, or "window" (support window + far tail)
    x_c: float = 100.0  # support cutoff of the windowed route
    line_roulette: dict | None = None  # LineRoulette policy fields (per-primary line roulette), or None
    depth_bins: int = 0  # > 0: also tally the depth-dose profile over this many bins of the slab
    coherent: bool = False  # sum line fields before squaring (the coherent route); brem stays incoherent
    coherent_grouped: bool = False  # also keep the decoherence blend's per-electron grouped term (batched route)
    cone_split: dict | None = None  # ConeSplitting policy fields (detector-cone splitting of elastic events), or None
    elastic_model: str = "sr"  # "sr" or "mott" (SRD 64 tables from mott_tables_dir / PYRITE_MOTT_TABLES_DIR)
    mott_tables_dir: str | None = None
    atomic_electron_deflection: str | None = None  # "kawrakow": elastic rate x (1 + 1/Z), or None
    elastic_tables: list | None = None  # elastic_model="elsepa": one ELSEPA native-grid table per element

    def grid(self):
        return bin_centers(self.E_lo_eV, self.E_hi_eV, self.n_bins)

    @staticmethod
    def silicon_1mm_10MeV(Ne=100_000, **kw):
        """10 MeV Gaussian beam (sigma 1 mm, divergence 0.5 mrad) on 1 mm Si with
        (111) tilted 1 deg; PXR/CBS over the 16 shortest allowed reflections
        and brem into 4096 bins over 1-100 keV; mixed transport precision and
        the windowed line route."""
        from .spectrum.crystal import allowed_reflections

        base = dict(E0_keV=10_000.0, Ne=Ne, thickness_ang=1.0e7, composition=[("Si", 0.04994)],
                    crystal="silicon", hkl_list=allowed_reflections("silicon", 16), tilt_deg=1.0,
                    n_bins=4096, E_lo_eV=1000.0, E_hi_eV=100_000.0,
                    beam=BeamSpec.gaussian_waist(1.0, 5.0e-4, 10_000.0), precision="mixed",
                    line_route="window")
        base.update(kw)
        return CaseConfig(**base)


class CaseTables:
    """Host-side inputs of a case: layer/LUT tables, geometry, detector rays, line tables."""

    def __init__(self, cfg: CaseConfig):
        self.cfg = cfg
        self.layers = build_layer_tables([(0.0, cfg.thickness_ang, list(cfg.composition))], cfg.elastic_model,
                                         cfg.mott_tables_dir, cfg.atomic_electron_deflection,
                                         elastic_tables=cfg.elastic_tables)
        # tables span the cutoff to the beam's strict upper energy (E0 with any spread)
        E_top = upper_beam_energy_keV(cfg.E0_keV, cfg.beam.energy_spread_frac)
        if cfg.elastic_model == "mott":
            from .tables.mott import check_mott_coverage

            check_mott_coverage(self.layers.compositions, cfg.E_cut_keV, E_top, cfg.mott_tables_dir)
        if cfg.elastic_model == "elsepa":
            from .tables.elsepa import check_coverage

            check_coverage(self.layers.elsepa, cfg.E_cut_keV, E_top)
        self.lut = build_transport_lut(cfg.E_cut_keV, E_top, self.layers)
        th, tilt = np.deg2rad(cfg.theta_obs_deg), np.deg2rad(cfg.tilt_deg)
        self.beam_dir, self.n_hat = tilted_geometry(th, tilt)
        self.n_hats, self.weights = cone_detector(th, tilt, cfg.cone_half_angle_rad, cfg.n_side)
        self.omega = float(self.weights.sum())
        self.E_grid = cfg.grid()
        self.line_tables = line_tables(cfg.crystal, cfg.hkl_list, self.E_grid, cfg.composition)


# Device bytes all live pipelines may hold in slot-row arenas together, as a
# fraction of the device's memory; each pipeline's two arenas take at most
# ROW_ARENA_MAX_BYTES each.
ROW_ARENA_DEVICE_FRACTION = 0.30
ROW_ARENA_MAX_BYTES = 24 << 30
_ARENA_LIVE = {"bytes": 0}


def _arena_release(nbytes):
    _ARENA_LIVE["bytes"] = max(0, _ARENA_LIVE["bytes"] - int(nbytes))


class CasePipeline(CaseTables):
    def __init__(self, cfg: CaseConfig, transport_config: PerElectronTransportConfig | None = None,
                 arena_bytes: int | None = None):
        import torch

        if cfg.coherent_grouped and not cfg.coherent:
            raise ValueError("coherent_grouped needs the coherent route (coherent=True)")
        super().__init__(cfg)
        self.dev = require_device().torch_device
        dev = self.dev
        # Slot rows come from two byte arenas used in turn, one sub-batch each:
        # 2048 slots per history to start with (a 1e5-history batch of slots
        # is 23.6 GB), so capacity replays are rare, and a batch whose grid
        # outgrows an arena is cut into sub-batches, each scored before the
        # arena it sits in is refilled.  Every live pipeline's arenas together
        # stay under ROW_ARENA_DEVICE_FRACTION of the device's memory.  The
        # arena size changes no transport row; it can move where sub-batches
        # split a batch, and with them the FP32 line groups of the tallies
        # (last-bit differences), so bitwise comparisons pass equal
        # ``arena_bytes``.
        total = int(torch.cuda.get_device_properties(dev).total_memory)
        room = int(ROW_ARENA_DEVICE_FRACTION * total) - _ARENA_LIVE["bytes"]
        want = ROW_ARENA_MAX_BYTES if arena_bytes is None else int(arena_bytes)
        self.arena_bytes = max(256 << 20, min(want, room // 2))
        self._reserved = 2 * self.arena_bytes  # held against the device-wide arena budget until release()
        _ARENA_LIVE["bytes"] += self._reserved
        self.tcfg = transport_config or PerElectronTransportConfig(seg_capacity=2048,
                                                                   scratch_budget_bytes=self.arena_bytes)
        self.tables = DeviceTables(self.layers, self.lut, dev)
        self.control = TransportControl(max_steps=cfg.max_steps, energy_model=1, straggle_on=cfg.straggling,
                                        precision=cfg.precision)
        self.core = make_hip_transport_core(self.tables, self.control)
        self.lines = LineScorer(self.line_tables, self.E_grid, cfg.thickness_ang, self.n_hats, self.weights, dev)
        self.brem = BremScorer(cfg.composition, self.E_grid, cfg.thickness_ang, self.n_hat, dev)
        self.coherent = None
        if cfg.coherent:
            if cfg.line_roulette is not None:
                raise ValueError("line roulette weights line intensities; the coherent route sums fields")
            from .spectrum.coherent import CoherentScorer

            self.coherent = CoherentScorer(self.line_tables, cfg.crystal, self.E_grid, cfg.thickness_ang,
                                           self.n_hats, self.weights, dev, grouped=cfg.coherent_grouped)
        self.cone_core = None
        if cfg.cone_split is not None:
            from .cone_split import ConeSplitCore, ConeSplitting

            policy = ConeSplitting.from_dict(cfg.cone_split)
            policy.validate_case(cfg)
            self.cone_request = policy.request(self.n_hat, cfg.E0_keV)
            self.cone_core = ConeSplitCore(self.tables, self.control, self.cone_request, dev)
            self.cone_stats = {"copies": 0, "killed": 0, "copy_rows": 0}
            self.copy_pool = 1 << 16  # copies pooled before a copy generation runs
            self._pool, self._pool_n, self._copy_cap_hint = [], 0, 256
        self.cone_roulette = None  # a fitted vr.cone_roulette.LearnedConeRoulette thins the copies
        n = cfg.Ne
        self.keys = torch.empty(n, dtype=torch.uint64, device=dev)
        self.pos = torch.empty((n, 3), dtype=torch.float64, device=dev)
        self.dirs = torch.empty((n, 3), dtype=torch.float64, device=dev)
        self.E = torch.empty(n, dtype=torch.float64, device=dev)
        self.clock = torch.empty(n, dtype=torch.float64, device=dev)
        self.E_cut = torch.full((n,), cfg.E_cut_keV, dtype=torch.float64, device=dev)
        self.alive = torch.ones(n, dtype=torch.uint8, device=dev)
        self.stragg = torch.zeros(n, dtype=torch.float64, device=dev)
        self.beam_d = torch.as_tensor(self.beam_dir, device=dev)
        self.spec_lines = torch.zeros(self.E_grid.size, dtype=torch.float64, device=dev)
        self.spec_brem = torch.zeros(self.E_grid.size, dtype=torch.float64, device=dev)
        self.spec_dose = torch.zeros(max(int(cfg.depth_bins), 1), dtype=torch.float64, device=dev)
        self.histories = 0
        self.segments = 0
        self.timers = {}
        self._cap_hint = None
        self._arenas = [None, None]
        self._arena_turn = 0
        self._arena_cur = 0
        self._arena_e = None
        self._arena_free = [None, None]
        self._rec = None
        self._ones = torch.ones(n, dtype=torch.float64, device=dev)

    def __del__(self):
        try:
            self.release()
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass

    def release(self):
        """Drop the slot-row arenas (their bytes return to the other pipelines' budget)."""
        for k in range(len(getattr(self, "_arenas", ()))):
            self._arenas[k] = None
        _arena_release(getattr(self, "_reserved", 0))
        self._reserved = 0

    def _arena(self, idx):
        from .backend import is_oom_error
        from .transport.core import RowArena

        a = self._arenas[idx]
        if a is None:
            nbytes = self.arena_bytes
            while True:
                try:
                    a = RowArena(nbytes, self.dev)
                    break
                except Exception as exc:  # noqa: BLE001 - halve on device OOM only
                    if not is_oom_error(exc) or nbytes <= (256 << 20):
                        raise
                    import torch

                    torch.cuda.empty_cache()
                    nbytes //= 2
            self._arenas[idx] = a
        return a

    def _arena_rows(self, e, m, cap, midpoint, device):
        """Slot rows of the sub-batch starting at electron ``e`` of the current call,
        carved from the next arena in turn (a replay of the same sub-batch reuses
        its arena); waits, on the device, until the scoring that last read that
        arena has finished."""
        import torch

        if e != self._arena_e:
            self._arena_e = e
            self._arena_cur = self._arena_turn
            self._arena_turn ^= 1
            ev = self._arena_free[self._arena_cur]
            if ev is not None:
                torch.cuda.current_stream().wait_event(ev)
        return self._arena(self._arena_cur).rows(m, cap, midpoint)

    def transport(self, batch_index: int, compact: bool = True, consume=None, core=None):
        """Transport histories ``[b*Ne, (b+1)*Ne)``; returns the device segment rows
        (compacted; with ``compact=False`` the slot rows of each sub-batch, or,
        with ``consume``, nothing: ``consume(e0, m, cap, rows, seg_count)`` takes
        each sub-batch as it is done, and must queue its readers of ``rows``
        before it returns, then call ``_scored()``).

        Slot rows live in two arenas used in turn, one sub-batch each, so the
        rows of a sub-batch stay valid until the second sub-batch after it is
        transported; a sub-batch waits (on the device) until the scoring that
        last read its arena has finished.  Without ``consume`` a
        ``compact=False`` call may use at most two sub-batches (both stay
        referenced in the result): a batch whose grid needs more raises."""
        cfg = self.cfg
        n = cfg.Ne
        stream = __import__("torch").cuda.current_stream().cuda_stream
        _lib.call("hr_stream_keys", int(cfg.seed), int(batch_index) * n, n, self.keys.data_ptr(), stream)
        fill_entry(cfg.beam, cfg.seed, int(batch_index) * n, n, cfg.E0_keV, np.deg2rad(cfg.tilt_deg),
                   self.pos, self.dirs, self.E, self.clock, stream=stream)
        self._arena_e = None
        held = []

        def alloc(e, m, cap, mid, dev):
            if not compact and consume is None and e != self._arena_e:
                held.append(e)
                if len(held) > 2:
                    raise RuntimeError(f"batch of {n} histories needs more than two slot-row sub-batches at "
                                       f"capacity {cap}; pass consume= (or compact=True)")
            return self._arena_rows(e, m, cap, mid, dev)

        res = drive_batches(self.core if core is None else core, self.keys,
                            (self.alive, self.clock, self.pos, self.dirs, self.E, self.E_cut),
                            self.stragg, self.control, n, self.dev, self.tcfg, stream=stream, compact=compact,
                            cap_hint=None if self._cap_hint is None else max(self._cap_hint, self.tcfg.seg_capacity),
                            alloc=alloc, consume=consume)
        self._cap_hint = res.capacity
        return res

    def _scored(self):
        """Mark the rows of the current arena as read by the work queued on the
        current stream (the arena may be refilled after it)."""
        import torch

        ev = torch.cuda.Event()
        ev.record(torch.cuda.current_stream())
        self._arena_free[self._arena_cur] = ev

    def records(self, res, first_id=None):
        """FP32 scoring records gathered straight from the slot rows of a
        ``compact=False`` transport (no FP64 compaction pass); with
        ``first_id`` (the batch's first global electron id) also each row's
        primary id as ``"prim"``."""
        return self.slot_records(res.slots, int(res.n_segments), first_id)

    def slot_records(self, slots, n_rows, first_id=None, weights=None):
        """FP32 scoring records of slot rows ``[(e0, m, cap, rows, seg_count)]``
        holding ``n_rows`` written rows in total; with ``weights`` (FP64, one per
        particle of a single sub-batch) the records also carry ``"w"``, each
        row's particle weight."""
        import ctypes as C

        import torch

        n_rows = int(n_rows)
        if self._rec is None or self._rec["L"].shape[0] < n_rows:
            m = max(n_rows, 1)
            self._rec = {"r_mid": torch.empty((m, 3), dtype=torch.float32, device=self.dev),
                         "v_hat": torch.empty((m, 3), dtype=torch.float32, device=self.dev),
                         "L": torch.empty(m, dtype=torch.float32, device=self.dev),
                         "E": torch.empty(m, dtype=torch.float32, device=self.dev),
                         "w": torch.empty(m, dtype=torch.float32, device=self.dev),
                         "prim": torch.empty(m, dtype=torch.int64, device=self.dev)}
        rec = self._rec
        lib = _lib.load()
        o = 0
        for e0, m, cap, rows, seg_count in slots:
            cnt = seg_count.clamp(max=cap)
            off = torch.cumsum(cnt, 0) - cnt + o
            g = _lib.VrGatherArgs()
            g.n, g.rows = m, cap
            g.n_out = n_rows if len(slots) == 1 else int(cnt.sum())  # one thread per written row
            g.seg_count, g.row_offset = seg_count.data_ptr(), off.data_ptr()
            if weights is not None and len(slots) != 1:
                raise ValueError("particle weights are taken for one sub-batch at a time")
            g.weight = (self._ones[:m] if weights is None else weights).data_ptr()
            g.seg_mid, g.seg_dir = rows["seg_mid"].data_ptr(), rows["seg_dir"].data_ptr()
            g.seg_len, g.seg_E, g.seg_E_end = rows["seg_len"].data_ptr(), rows["seg_E"].data_ptr(), \
                rows["seg_E_end"].data_ptr()
            g.r_mid, g.v_hat = rec["r_mid"].data_ptr(), rec["v_hat"].data_ptr()
            g.L, g.E, g.row_weight = rec["L"].data_ptr(), rec["E"].data_ptr(), rec["w"].data_ptr()
            if first_id is not None:
                prim = torch.arange(int(first_id) + e0, int(first_id) + e0 + m, dtype=torch.int64, device=self.dev)
                g.prim, g.row_prim = prim.data_ptr(), rec["prim"].data_ptr()
            _lib.check(lib.hr_vr_gather(C.byref(g), C.c_void_p(torch.cuda.current_stream().cuda_stream)),
                       "hr_vr_gather")
            o += int(cnt.sum()) if len(slots) > 1 else 0
        keys = ("r_mid", "v_hat", "L", "E") + (("prim",) if first_id is not None else ()) + \
            (("w",) if weights is not None else ())
        return {k: rec[k][:n_rows] for k in keys}

    def score_dose(self, res, Ne):
        """Depth-dose of a ``compact=False`` transport's slot rows into ``spec_dose``
        (keV per electron of ``Ne`` per bin)."""
        from .transport.tally import depth_dose

        nb = int(self.cfg.depth_bins)
        if nb <= 0:
            return
        dz = float(self.cfg.thickness_ang) / nb
        for _e0, m, cap, rows, seg_count in res.slots:
            depth_dose(self.spec_dose, 0.0, dz, rows, m, slots=cap, seg_count=seg_count, scale=1.0 / float(Ne))

    def score_records(self, seg, Ne, line_weight=None, row_weight=None):
        """Line and brem tallies of FP32 scoring records; ``line_weight`` (FP32 per
        row) weights the line tally only (the line-roulette strata),
        ``row_weight`` (FP32 per row, the particle weight of a copy) both."""
        if line_weight is not None and row_weight is not None:
            raise ValueError("line_weight and row_weight are not combined")
        lw = row_weight if line_weight is None else line_weight
        if self.cfg.line_route == "window":
            self.spec_lines += self.lines.window_spectra(seg, Ne, x_c=self.cfg.x_c, row_weight=lw).sum(0)
        elif self.cfg.line_route == "dense":
            self.lines.spectrum(seg, Ne, self.spec_lines, row_weight=lw)
        else:
            raise ValueError(f"unknown line route {self.cfg.line_route!r}")
        if row_weight is None:
            self.brem.spectrum(seg, Ne, self.spec_brem)
        else:
            self.brem.spectrum(seg, Ne, self.spec_brem, row_weight=row_weight)

    def line_screen(self, n_electrons):
        """The case's per-primary line-roulette screen over ``n_electrons`` primaries, or None."""
        if self.cfg.line_roulette is None:
            return None
        from .line_roulette import LineRoulette, LineScreen

        scr = getattr(self, "_screen", None)
        if scr is None or scr.n != int(n_electrons):
            policy = LineRoulette.from_dict(self.cfg.line_roulette)
            policy.validate_case(self.cfg, self.control.max_dE_frac)
            scr = self._screen = LineScreen(policy, self.cfg.seed, n_electrons, self.dev)
        return scr

    def score(self, rows, Ne):
        seg = segments_fp32({
            "r_mid": rows["seg_mid"], "v_hat": rows["seg_dir"], "L_ang": rows["seg_len"],
            "E_keV": rows["seg_E"], "E_repr_keV": 0.5 * (rows["seg_E"] + rows["seg_E_end"]),
        }, self.dev)
        self.score_records(seg, Ne)
        return seg

    def _score_slot(self, e0, m, cap, rows, seg_count, batch_index, n_total, weights=None):
        """Every tally of one sub-batch of slot rows (queued on the current stream);
        ``weights`` (FP64 per particle) weights every tally of its rows (copies)."""
        slot = [(e0, m, cap, rows, seg_count)]
        n_rows = int(seg_count.clamp(max=cap).sum())
        if weights is not None:
            seg = self.slot_records(slot, n_rows, weights=weights)
            self.score_records(seg, n_total, row_weight=seg["w"])
            self._scored()
            self.segments += n_rows
            return
        nb = int(self.cfg.depth_bins)
        if nb > 0:
            from .transport.tally import depth_dose

            depth_dose(self.spec_dose, 0.0, float(self.cfg.thickness_ang) / nb, rows, m, slots=cap,
                       seg_count=seg_count, scale=1.0 / float(n_total))
        screen = self.line_screen(n_total)
        if self.coherent is not None:
            self.coherent.add_slots(slot, self.cfg.Ne if e0 == 0 else 0)
            self.brem.spectrum(self.slot_records(slot, n_rows), n_total, self.spec_brem)
        elif screen is None:
            self.score_records(self.slot_records(slot, n_rows), n_total)
        else:
            seg = self.slot_records(slot, n_rows, first_id=int(batch_index) * self.cfg.Ne)
            self.score_records(seg, n_total, screen.row_weights(self.lines, seg, seg["prim"], n_total))
        self._scored()
        self.segments += n_rows

    def step(self, batch_index: int, n_batches_total: int = 1):
        """One batch of ``Ne`` histories through transport and scoring.

        Tallies are per incident electron of the whole run when
        ``n_batches_total`` batches are summed.
        """
        n_total = self.cfg.Ne * n_batches_total

        def consume(e0, m, cap, rows, seg_count):
            self._score_slot(e0, m, cap, rows, seg_count, batch_index, n_total)

        if self.cone_core is not None:
            return self._step_cone(batch_index, n_total, consume)
        res = self.transport(batch_index, compact=False, consume=consume)
        self.histories += self.cfg.Ne
        return res

    def _step_cone(self, batch_index, n_total, consume, copy_consumer=None):
        """A batch with detector-cone splitting: the primaries (rows weight 1,
        copies collected per launch), their copies pooled for one later copy
        generation (``flush_copies``) once ``copy_pool`` copies are waiting."""
        self.cone_core.reset()
        res = self.transport(batch_index, compact=False, consume=consume, core=self.cone_core)
        self.histories += self.cfg.Ne
        cp = self.cone_core.copies()
        if cp is not None and self.cone_roulette is not None:
            cp = self.cone_roulette.apply(cp)  # learned importance roulette of the copies
        n_copies = 0 if cp is None else int(cp["weight"].numel())
        self.cone_stats["copies"] += n_copies
        self.cone_stats["killed"] += 0 if cp is None else int(cp["killed"].sum())
        if self._pool and self._pool[0][1] != n_total:
            self.flush_copies(copy_consumer)
        if n_copies:
            self._pool.append((cp, n_total))
            self._pool_n += n_copies
        if self._pool_n >= self.copy_pool:
            self.flush_copies(copy_consumer)
        res.cone_copies = n_copies
        return res

    def flush_copies(self, copy_consumer=None):
        """Transport every pooled copy as one analog generation and score its
        rows with its weight (``cone_split.py:447-513``).  Copies are
        independent of each other, so pooling several batches' copies into one
        launch changes no tally beyond FP64 summation order, while the launch's
        latency (the longest copy history, often one running along the slab
        face) is paid once per pool instead of once per batch.
        ``copy_consumer(weights, n_total)`` returns the ``consume`` callback of
        the copies' slot rows (the two-stream overlap of ``run_steps``); by
        default they are scored on the current stream."""
        import torch

        from .cone_split import transport_copies

        if not self._pool:
            return
        n_total = self._pool[0][1]
        keys = ("flight", "pos", "dir", "E_keV", "clock", "weight", "stream_key")
        cp = {k: torch.cat([c[k] for c, _ in self._pool]) for k in keys}
        self._pool, self._pool_n = [], 0
        w = cp["weight"]

        def consume_copies(e0, m, cap, rows, seg_count):
            self.cone_stats["copy_rows"] += int(seg_count.clamp(max=cap).sum())
            self._score_slot(e0, m, cap, rows, seg_count, 0, n_total, weights=w[e0:e0 + m])

        if copy_consumer is not None:
            consume_copies = copy_consumer(w, n_total)

        budget = min(self.arena_bytes, 8 << 30)
        res = transport_copies(cp, self.tables, self.control, self.cfg.E_cut_keV, self.dev,
                               PerElectronTransportConfig(seg_capacity=max(256, self._copy_cap_hint),
                                                          scratch_budget_bytes=budget,
                                                          kernel_threads=self.tcfg.kernel_threads),
                               consume=consume_copies)
        self._copy_cap_hint = max(self._copy_cap_hint, int(res.capacity))
        self.cone_stats["generations"] = self.cone_stats.get("generations", 0) + 1

    def run_steps(self, first_batch: int, n_steps: int, n_batches_total: int | None = None, overlap: bool = True):
        """Batches ``first_batch ... first_batch + n_steps - 1`` through transport and scoring.

        With ``overlap`` the transport of batch s+1 runs on one HIP stream while
        batch s is scored on another: the transport kernel is latency bound
        and leaves issue slots that the scoring kernels fill.  Each sub-batch's
        slot rows are scored in order on the scoring stream, so tallies are
        those of ``step()`` run back to back.
        """
        import torch

        n_total_b = int(n_steps if n_batches_total is None else n_batches_total)
        if not overlap or self.cfg.line_roulette is not None:
            for b in range(int(n_steps)):
                self.step(int(first_batch) + b, n_total_b)
            return
        n_total = self.cfg.Ne * n_total_b
        main = torch.cuda.current_stream()
        if getattr(self, "_streams", None) is None:
            self._streams = (torch.cuda.Stream(device=self.dev), torch.cuda.Stream(device=self.dev))
        sa, sb = self._streams
        sa.wait_stream(main)
        sb.wait_stream(main)

        def copy_consumer(w, n_tot):
            """Copies' slot rows (fresh tensors, not arena rows) scored on the
            scoring stream; their memory is held for it by record_stream."""

            def consume_copies(e0, m, cap, rows, seg_count):
                self.cone_stats["copy_rows"] += int(seg_count.clamp(max=cap).sum())
                done = torch.cuda.Event()
                done.record(sa)
                wi = w[e0:e0 + m]
                for t in (*rows.values(), seg_count, wi):
                    t.record_stream(sb)
                with torch.cuda.stream(sb):
                    sb.wait_event(done)
                    self._score_slot(e0, m, cap, rows, seg_count, 0, n_tot, weights=wi)

            return consume_copies

        for b in range(int(n_steps)):
            batch = int(first_batch) + b

            def consume(e0, m, cap, rows, seg_count, batch=batch):
                done = torch.cuda.Event()
                done.record(sa)
                seg_count.record_stream(sb)
                with torch.cuda.stream(sb):
                    sb.wait_event(done)
                    self._score_slot(e0, m, cap, rows, seg_count, batch, n_total)

            with torch.cuda.stream(sa):
                if self.cone_core is None:
                    self.transport(batch, compact=False, consume=consume)
                    self.histories += self.cfg.Ne
                else:
                    self._step_cone(batch, n_total, consume, copy_consumer)
        if self.cone_core is not None:
            with torch.cuda.stream(sa):
                self.flush_copies(copy_consumer)
        main.wait_stream(sa)
        main.wait_stream(sb)

    def stream(self, first_id: int, n_histories: int, n_total: int | None = None, lanes: int = 1 << 20,
               rows_per_round: int = 32, timings: dict | None = None, sort_lanes: bool = False):
        """Streamed transport -> scoring over electrons ``[first_id, first_id + n_histories)``.

        Continuation rounds (``hr_transport_continue``) keep ``lanes`` histories
        in flight, refilled from the pending primaries; every round's rows are
        gathered into FP32 records and scored before the next launch, so device
        memory holds one round of rows (``lanes x rows_per_round``) however many
        histories run.  Tallies are per electron of ``n_total`` (default
        ``n_histories``).  Rows and tallies equal the batched route's up to the
        FP64 summation order of the tallies.
        """
        import ctypes as C

        import torch

        from .transport.continuation import ContinuationDriver

        if self.coherent is not None and self.coherent.grouped is not None:
            raise ValueError("the grouped coherent term squares whole histories, which continuation rounds "
                             "split; run it on the batched route")

        cfg = self.cfg
        if cfg.line_roulette is not None:
            from .line_roulette import LineRoulette

            LineRoulette.from_dict(cfg.line_roulette).validate_case(cfg, self.control.max_dE_frac, route="streamed")
        n_total = int(n_histories if n_total is None else n_total)
        lanes = max(1, min(int(lanes), int(n_histories)))
        drv = getattr(self, "_stream_drv", None)
        if drv is None or drv.lanes != lanes or drv.R != int(rows_per_round) or drv.sort_lanes != bool(sort_lanes):
            self._stream_drv = None
            drv = ContinuationDriver(self.tables, self.control, lanes, int(rows_per_round), self.dev,
                                     cfg.E_cut_keV, sort_lanes=sort_lanes)
            self._stream_drv = drv
            m = lanes * int(rows_per_round)
            self._srec = {"r_mid": torch.empty((m, 3), dtype=torch.float32, device=self.dev),
                          "v_hat": torch.empty((m, 3), dtype=torch.float32, device=self.dev),
                          "L": torch.empty(m, dtype=torch.float32, device=self.dev),
                          "E": torch.empty(m, dtype=torch.float32, device=self.dev),
                          "w": torch.empty(m, dtype=torch.float32, device=self.dev)}
            self._sones = torch.ones(lanes, dtype=torch.float64, device=self.dev)
        rec, lib = self._srec, _lib.load()
        tilt = np.deg2rad(cfg.tilt_deg)
        tim = timings if timings is not None else {}

        def fill(start, n, pos, dirs, E, clock, keys):
            fill_entry(cfg.beam, cfg.seed, start, n, cfg.E0_keV, tilt, pos, dirs, E, clock)
            _lib.call("hr_stream_keys", int(cfg.seed), int(start), int(n), keys.data_ptr(), None)

        def consume(rnd):
            cnt = rnd.seg_count
            incl = torch.cumsum(cnt, 0)
            n_rows = int(incl[-1].item()) if rnd.n_lanes else 0
            if n_rows == 0:
                return
            off = incl - cnt
            g = _lib.VrGatherArgs()
            g.n, g.rows, g.n_out = rnd.n_lanes, rnd.rows_per_round, n_rows
            g.seg_count, g.row_offset = cnt.data_ptr(), off.data_ptr()
            g.weight = self._sones.data_ptr()
            rows = rnd.rows
            g.seg_mid, g.seg_dir = rows["seg_mid"].data_ptr(), rows["seg_dir"].data_ptr()
            g.seg_len, g.seg_E, g.seg_E_end = rows["seg_len"].data_ptr(), rows["seg_E"].data_ptr(), \
                rows["seg_E_end"].data_ptr()
            g.r_mid, g.v_hat = rec["r_mid"].data_ptr(), rec["v_hat"].data_ptr()
            g.L, g.E, g.row_weight = rec["L"].data_ptr(), rec["E"].data_ptr(), rec["w"].data_ptr()
            _lib.check(lib.hr_vr_gather(C.byref(g), C.c_void_p(torch.cuda.current_stream().cuda_stream)),
                       "hr_vr_gather")
            seg = {k: rec[k][:n_rows] for k in ("r_mid", "v_hat", "L", "E")}
            if self.coherent is not None:  # fields add over rounds; squared when the spectrum is read
                self.coherent.add_rows(rnd.rows, cnt, rnd.n_lanes, rnd.rows_per_round)
                self.brem.spectrum(seg, n_total, self.spec_brem)
            else:
                self.score_records(seg, n_total)
            if cfg.depth_bins > 0:
                from .transport.tally import depth_dose

                depth_dose(self.spec_dose, 0.0, float(cfg.thickness_ang) / cfg.depth_bins, rnd.rows, rnd.n_lanes,
                           slots=rnd.rows_per_round, seg_count=cnt, scale=1.0 / float(n_total))
            tim["rows"] = tim.get("rows", 0) + n_rows

        res = drv.run(int(first_id), int(n_histories), fill, consume)
        if self.coherent is not None:
            self.coherent.n_electrons += int(n_histories)
        self.histories += int(n_histories)
        self.segments += res.n_segments
        return res

    def spectra(self):
        """``(lines [photons/eV/electron], brem [photons/eV/electron] over the cone)``.

        On the coherent route the lines are the squared fields of every
        history added since ``reset_coherent()``, per electron of them.  With
        cone splitting, pooled copies are transported and scored first."""
        if self.cone_core is not None:
            self.flush_copies()
        if self.coherent is not None:
            lines = self.coherent.spectrum().cpu().numpy()
        else:
            lines = self.spec_lines.cpu().numpy()
        return lines, self.spec_brem.cpu().numpy() * self.omega

    def reset_coherent(self):
        """Clear the coherent field planes (a new independent estimate starts)."""
        if self.coherent is not None:
            self.coherent.reset()

    def depth_dose(self):
        """Deposited energy per depth bin [keV/electron] (``depth_bins`` uniform bins over the slab)."""
        return self.spec_dose.cpu().numpy()[: max(int(self.cfg.depth_bins), 0)]
