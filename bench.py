"""Throughput of the plain-method hot path on one MI355X: histories/s.

Workload (one "step"): one batch of Ne = 1e5 histories of the HOPG case --
5 MeV pencil beam at normal incidence on 100 um graphite with the c-axis tilted
0.5 deg, per-electron LUT transport (screened-Rutherford elastic, spliced
stopping, Urban straggling, midpoint rule), then PXR/CBS (002) line scoring over
a 10 mrad detector cone at 90 deg (3x3 rays) and BH+Elwert brem into 1024 FP64
bins over 1-50 keV.  Inputs are synthetic (a seeded beam), tables are resident
in HBM before the timed region, and every step transports new histories
(global electron ids [s*Ne, (s+1)*Ne)).  Successive batches overlap: the
transport of batch s+1 runs on one HIP stream while batch s is scored on
another (`--no-overlap` runs them back to back on one stream).

Prints ONE JSON line.  `roofline` is for the dominant kernel (the dense sinc^2
reduction: one hr_sincsq_reduce call = sincsq_uniform_kernel plus its fixed-order
partial-sum kernel), timed with HIP events around its launches inside the timed
region; `cpu_baseline` times the FP64 C oracle (the CPU restatement of the
reference path, OpenMP over histories and bins) on a bounded sample of the same
workload.  `fom` compares the PXR-peak figure of merit 1/(rel_err^2 t) of
the analog run, detector-cone splitting with the reference policy's defaults,
and cone splitting with the learned importance roulette of its copies (48
groups of 16 batches of 65,536 histories per arm, disjoint, with bootstrap
intervals over groups; after the timed region, not part of `value`).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=5)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--ne", type=int, default=100_000)
    p.add_argument("--n-side", type=int, default=3)
    p.add_argument("--line-route", choices=("dense", "window"), default="dense",
                   help="dense: every (line, bin) pair, the reference's device default for per-row lines; "
                        "window: the support-windowed route with x_c = 100 (the reference's NumPy default)")
    p.add_argument("--cpu-sample", type=int, default=8000,
                   help="histories per timed run of the all-core CPU baseline (median of 3)")
    p.add_argument("--cpu-sample-one", type=int, default=500,
                   help="histories per timed run of the single-thread CPU baseline (median of 3)")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--breakdown", action="store_true", help="print per-phase times to stderr")
    p.add_argument("--no-overlap", action="store_true", help="run transport and scoring of successive batches "
                                                              "back to back on one stream")
    p.add_argument("--no-fom", action="store_true", help="skip the learned-importance figure of merit")
    p.add_argument("--fom-groups", type=int, default=96,
                   help="FOM sample units per arm (groups of batches); 96 x 16 x 65536 = 1.0e8 histories per arm")
    p.add_argument("--fom-train-batches", type=int, default=153,
                   help="batches whose copies train the learned roulette (153 x 65536 = 1.0e7 histories)")
    p.add_argument("--fom-group-batches", type=int, default=16, help="batches per FOM group")
    p.add_argument("--fom-histories", type=int, default=65536, help="histories per FOM batch")
    p.add_argument("--fom-sequential", action="store_true", help="FOM arms back to back on one stream")
    p.add_argument("--coherent-grouped", action="store_true",
                   help="coherent workload: also accumulate the decoherence blend's per-electron grouped term")
    p.add_argument("--workload", choices=("hopg", "si-stream", "pixels", "coherent", "queue"), default="hopg",
                   help="hopg: the bench line (default); si-stream: the 1 mm Si 10 MeV case streamed through "
                        "continuation rounds; pixels: windowed scoring of synthetic segments over 32x32 pixels; "
                        "coherent: the HOPG case on the coherent line route; queue: 256 uneven configurations "
                        "in one graph-replayed work-queue kernel")
    p.add_argument("--queue-scale", type=float, default=1e-3, help="queue: history-count scale of the campaign")
    p.add_argument("--queue-replays", type=int, default=3, help="queue: timed graph replays (the fastest counts)")
    p.add_argument("--histories", type=int, default=1 << 21, help="si-stream: histories")
    p.add_argument("--lanes", type=int, default=1 << 19, help="si-stream: lanes in flight")
    p.add_argument("--rows-per-round", type=int, default=32, help="si-stream: rows per launch")
    p.add_argument("--log2-segments", type=int, default=20, help="pixels: log2 of the segment count")
    p.add_argument("--copy-calibration", action="store_true",
                   help="also run one 1 GiB device copy (HBM counter calibration under rocprofv3 --pmc)")
    return p.parse_args()


def _fom_arm(pipe, window, dE, first, n_groups, per_group, label, overlap=True):
    """Per-group PXR-peak tallies of ``n_groups`` groups of ``per_group`` batches
    (a group is the sample unit: a cone-split pipeline transports the copies of
    a group in one generation), and the wall time.  With ``overlap`` a group's
    batches run through ``run_steps`` on two streams (transport and copy
    generations beside scoring), else back to back on one."""
    import torch

    vals = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for g in range(n_groups):
        pipe.spec_lines.zero_()
        pipe.spec_brem.zero_()
        if overlap:
            pipe.run_steps(first + g * per_group, per_group, per_group, overlap=True)
        else:
            for b in range(per_group):
                pipe.step(first + g * per_group + b, per_group)
        if pipe.cone_core is not None:
            pipe.flush_copies()
        vals.append(pipe.spec_lines[window].sum() * dE)
        if g % 16 == 15:
            print(f"fom: {label} {g + 1}/{n_groups} groups, {time.perf_counter() - t0:.1f} s", file=sys.stderr,
                  flush=True)
    vals = torch.stack(vals).cpu().numpy()
    torch.cuda.synchronize()
    return vals, time.perf_counter() - t0


def learned_fom(pipe_cfg, n_groups=96, per_group=16, Ne=65536, overlap=True, train_batches=153):
    """PXR-peak figure of merit 1/(rel_err^2 t) of three estimators of the same case:
    analog; detector-cone splitting with the reference policy's defaults; and
    cone splitting without the angle roulette, its copies thinned by the learned
    importance roulette (the BF16 network trained on the copies' own futures).
    Every arm runs the same schedule (``overlap``: transport beside scoring on
    two streams, as the bench's timed steps; else back to back), over disjoint
    histories; gains carry 68% bootstrap intervals over groups."""
    import torch

    from pyrite_amd.pipeline import CaseConfig, CasePipeline
    from pyrite_amd.vr.cone_roulette import LearnedConeRoulette
    from pyrite_amd.vr.importance import LearnedImportance, figure_of_merit

    base = {**pipe_cfg.__dict__, "Ne": Ne}
    # first batch index of each arm (disjoint electron ranges; training from batch 30,000)
    arms = (("analog", None, 0), ("cone_split", {}, 10_000), ("learned", {"angle_scale": None}, 20_000))
    rng = np.random.default_rng(0)
    out = {"histories_per_batch": Ne, "batches_per_group": per_group, "groups": n_groups,
           "schedule": "two streams" if overlap else "back to back"}
    boots = {}
    window = None
    for mode, cs, first in arms:
        pipe = CasePipeline(CaseConfig(**{**base, "cone_split": cs}))  # the timed steps' slot-row arenas
        if window is None:
            li = LearnedImportance(pipe)
            window = torch.as_tensor(li.window_bins, device=pipe.dev)
            dE = li.dE
            out["window_eV"] = list(li.window)
        extra = {}
        if mode == "learned":
            t0 = time.perf_counter()
            lr = LearnedConeRoulette(pipe, keep_quantile=0.9)
            x, y, w = lr.training_set(train_batches, first_batch=30_000)
            lr.fit(x, y, w)
            torch.cuda.synchronize()
            pipe.cone_roulette = lr
            extra = {"training_s": time.perf_counter() - t0, "training_copies": int(x.shape[0]),
                     "training_histories": int(train_batches) * int(Ne), "keep_quantile": lr.keep_quantile}
        pipe.step(10**6 + first, 1)  # warm-up
        pipe.spectra()
        vals, t = _fom_arm(pipe, window, dE, first, n_groups, per_group, mode, overlap)
        mean, rel, fom = figure_of_merit(vals, t)
        boot = []
        for _ in range(1000):
            sm = rng.choice(vals, vals.size)
            r = sm.std(ddof=1) / np.sqrt(sm.size) / sm.mean()
            boot.append(1.0 / (r * r * t))
        boots[mode] = np.asarray(boot)
        out[mode] = {"histories": n_groups * per_group * Ne, "mean_photons_per_electron": mean, "rel_err": rel,
                     "time_s": t, "fom": fom, "fom_boot68": [float(v) for v in np.quantile(boot, [0.16, 0.84])],
                     "max_over_median": float(vals.max() / np.median(vals)), **extra}
        if pipe.cone_core is not None:
            out[mode]["copies"] = int(pipe.cone_stats["copies"])
            out[mode]["copy_rows"] = int(pipe.cone_stats["copy_rows"])
        if pipe.cone_roulette is not None:
            out[mode]["roulette_kept"] = pipe.cone_roulette.stats["kept"] / max(1, pipe.cone_roulette.stats["seen"])
        pipe.release()
        del pipe
        torch.cuda.empty_cache()
    for mode in ("cone_split", "learned"):
        out[f"gain_{mode}_vs_analog"] = out[mode]["fom"] / out["analog"]["fom"]
        out[f"gain_{mode}_boot68"] = [float(v) for v in np.quantile(boots[mode] / boots["analog"], [0.16, 0.84])]
    # the learned roulette against the method without it: both arms' tails are bounded
    # by the copies' weights, so this ratio does not hang on the analog arm's rare events
    out["gain_learned_vs_cone_split"] = out["learned"]["fom"] / out["cone_split"]["fom"]
    out["gain_learned_vs_cone_split_boot68"] = [float(v) for v in
                                               np.quantile(boots["learned"] / boots["cone_split"], [0.16, 0.84])]
    return out


def si_stream(args):
    """1 mm Si (111), 10 MeV Twiss beam, 16 reflections x 9 rays, 4096 bins, streamed."""
    import torch

    from pyrite_amd.pipeline import CaseConfig, CasePipeline

    cfg = CaseConfig.silicon_1mm_10MeV(Ne=65536)
    pipe = CasePipeline(cfg)
    pipe.stream(10**9, 65536, lanes=args.lanes, rows_per_round=args.rows_per_round)  # warm-up
    torch.cuda.synchronize()
    pipe.spec_lines.zero_()
    pipe.spec_brem.zero_()
    t0 = time.perf_counter()
    res = pipe.stream(0, args.histories, lanes=args.lanes, rows_per_round=args.rows_per_round)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    return {"metric": "histories/s, streamed transport -> scoring", "value": args.histories / t,
            "unit": "histories/s", "n_gpus": 1, "seconds": t, "rows_per_second": res.n_segments / t,
            "rows_per_history": res.n_segments / args.histories, "rounds": res.rounds, "counts": res.counts,
            "dtype": "mixed transport / f64 line phase / f64 tallies", "data": "synthetic (seeded Twiss beam)",
            "config": {"workload": "1 mm Si (111) tilted 1 deg, 10 MeV Twiss beam (sigma 1 mm, 0.5 mrad), "
                                   "16 reflections x 9 rays (windowed route) + brem, 4096 bins 1-100 keV",
                       "histories": args.histories, "lanes": args.lanes, "rows_per_round": args.rows_per_round}}


def pixel_scoring(args):
    """Windowed line scoring of synthetic FP32 segments: 64 reflections, 32x32 pixels, 4096 bins."""
    import torch

    from pyrite_amd.backend import require_device
    from pyrite_amd.spectrum.geometry import cone_detector
    from pyrite_amd.spectrum.lines import LineScorer, synthetic_line_tables

    dev = require_device().torch_device
    rng = np.random.default_rng(2026)
    n = 1 << args.log2_segments
    n_E = 4096
    E0, dE = 1000.0 + 0.5 * 99000.0 / n_E, 99000.0 / n_E
    E_grid = E0 + dE * np.arange(n_E)
    tables = synthetic_line_tables(64, E_grid, rng)
    n_hats, weights = cone_detector(np.pi / 2, 0.0, 0.01, 32)
    r = np.stack([rng.uniform(-1e6, 1e6, n), rng.uniform(-1e6, 1e6, n), rng.uniform(0, 5e6, n)], 1)
    d = np.stack([rng.normal(0, 1e-3, n), rng.normal(0, 1e-3, n), np.ones(n)], 1)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    E = np.exp(rng.uniform(np.log(1000.0), np.log(20000.0), n))
    L = rng.exponential(1.0e4, n)
    seg = {k: torch.as_tensor(np.ascontiguousarray(v, dtype=np.float32), device=dev)
           for k, v in (("r_mid", r), ("v_hat", d), ("L", L), ("E", E))}
    sc = LineScorer(tables, E_grid, 5.0e6, n_hats, weights, dev)
    sc.window_spectra({k: v[:4096] for k, v in seg.items()}, 1.0)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = sc.window_spectra(seg, 1.0)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    lines = n * 64 * n_hats.shape[0]
    o = out.cpu().numpy()
    return {"metric": "(segment, reflection, pixel) lines/s, windowed route", "value": lines / t,
            "unit": "lines/s", "n_gpus": 1, "seconds": t, "lines": lines,
            "full_workload_seconds": (2**26 * 64 * 1024) / (lines / t),
            "finite": bool(np.isfinite(o).all()), "nonzero_bins": int(np.count_nonzero(o)),
            "dtype": "f32 records / f64 prologue and phase / f64 tallies", "data": "synthetic segments (seed 2026)",
            "config": {"workload": "2^%d segments x 64 reflections x 32x32 pixels x 4096 bins 1-100 keV"
                                   % args.log2_segments}}


def coherent_case(args):
    """The HOPG case with coherent line scoring (fields summed over every history, squared at the end)."""
    import torch

    from pyrite_amd import roofline
    from pyrite_amd.pipeline import CaseConfig, CasePipeline

    cfg = CaseConfig(Ne=args.ne, n_side=args.n_side, coherent=True, coherent_grouped=args.coherent_grouped)
    pipe = CasePipeline(cfg)
    pipe.run_steps(1000, args.warmup, overlap=not args.no_overlap)  # warm-up through the timed path
    torch.cuda.synchronize()
    pipe.reset_coherent()
    pipe.spec_brem.zero_()
    pipe.histories = pipe.segments = 0
    sc = pipe.coherent
    sc.events = []
    sc.record_events = True
    t0 = time.perf_counter()
    pipe.run_steps(0, args.steps, overlap=not args.no_overlap)
    lines = sc.spectrum()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    sc.record_events = False
    ms = [ev[0].elapsed_time(ev[1]) for ev, _ in sc.events]
    # live (line, row) pairs of the timed field launches, recomputed after the
    # timed region (the records are deterministic per batch) so that counting
    # adds no work inside it
    counts = []
    for s in range(args.steps):
        pipe.transport(s, compact=False, consume=lambda e0, m, cap, rows, cnt: (
            sc.add_rows(rows, cnt, m, cap, live_counts=counts), pipe._scored()))
    assert len(counts) == len(ms), (len(counts), len(ms))
    cells = [float(c) * cfg.n_bins for c in counts]
    avg_ms = sum(ms) / len(ms)
    achieved = roofline.COHERENT_CELL_OPS * (sum(cells) / len(cells)) / (avg_ms * 1e-3) / 1e12
    out = {
        "metric": "histories/s, coherent line route", "value": pipe.histories / dt, "unit": "histories/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 transport / f64 records and phase + f32 profile / f64 fields",
        "data": "synthetic (seeded pencil beam; analytic tables)",
        "config": {"workload": "HOPG 100 um, 5 MeV pencil, (002) coherent PXR/CBS fields over a 10 mrad cone "
                               "(3x3 rays) + BH brem, 1024 bins 1-50 keV",
                   "histories_per_step": cfg.Ne, "segments_per_history": pipe.segments / max(pipe.histories, 1),
                   "grouped_term": bool(args.coherent_grouped)},
        "roofline": {
            "kernel": "hr_coherent_field: coherent_field_kernel + field_partials_kernel",
            "bound": "valu", "pipe": "vector ALU, FP32 peak (FP64 phase reduction counted as FP32 work)",
            "achieved": achieved, "peak": roofline.PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / roofline.PEAK_FP32_TFLOPS, "traffic": None,
            "ops_per_cell": roofline.COHERENT_CELL_OPS, "live_cells_per_launch": sum(cells) / len(cells),
            "avg_launch_ms": avg_ms, "launches": len(ms),
            "algorithmic_bytes_per_launch": (sum(cells) / len(cells)) / cfg.n_bins * 104 + 9 * 4 * cfg.n_bins * 8,
        },
        "spectrum_integral": float(lines.sum().item()),
    }
    if not args.no_cpu:
        from oracle import case as ocase

        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        n = args.cpu_sample
        ocase.run_case(pipe, 10_000_000, 32, threads=threads, coherent=True)  # warm-up
        tim = {}
        t0 = time.perf_counter()
        ocase.run_case(pipe, 0, n, threads=threads, timings=tim, coherent=True)
        dtc = time.perf_counter() - t0
        out["cpu_baseline"] = {
            "value": n / dtc, "unit": "histories/s", "cores": threads, "kind": "port",
            "sample": f"{n} histories of the same case on the coherent route, FP64 C oracle, {threads} OpenMP "
                      f"threads, {dtc:.2f} s; transport {tim['transport']:.2f} s, lines {tim['lines']:.2f} s, "
                      f"brem {tim['brem']:.2f} s"}
        out["speedup_vs_cpu_baseline"] = out["value"] / out["cpu_baseline"]["value"]
    return out


def queue_campaign(args):
    """256 configurations of uneven size (histories log-uniform 1e4-1e7 times the
    scale, thickness 10 um-2 mm, 1-30 MeV, graphite/Si/Ge/W, tilt 0-3 deg) as one
    graph-replayed work-queue launch, against one launch per configuration."""
    import ctypes as C

    import torch

    from pyrite_amd import _lib
    from pyrite_amd.backend import require_device
    from pyrite_amd.queue import TransportQueue, campaign

    dev = require_device().torch_device
    cases = campaign(256, seed=2024, scale=args.queue_scale)
    tq = TransportQueue(cases, dev)
    tuned = tq.autotune(fraction=0.02)
    g = tq.graph()
    t0 = time.perf_counter()
    g.replay()
    torch.cuda.synchronize()
    print(f"queue: {tq.total} histories, warm-up replay {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    ts = []
    for r in range(max(1, int(args.queue_replays))):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        g.replay()
        torch.cuda.synchronize()
        ts.append(time.perf_counter() - t0)
        print(f"queue: replay {r + 1}: {ts[-1]:.1f} s", file=sys.stderr, flush=True)
    t_q = min(ts)
    tally = tq.tally.cpu().numpy()
    rows = float(tally[:, 7].sum())
    offs = tq.offsets_host
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(len(cases)):  # one launch per configuration, same kernel
        sub = torch.as_tensor(np.array([0, offs[i + 1] - offs[i]], dtype=np.int64), device=dev)
        L = tq.launch_args(offsets=sub, total=int(offs[i + 1] - offs[i]), tally=tq.tally[i:i + 1])
        L.configs = tq.configs.data_ptr() + i * C.sizeof(_lib.QueueConfig)
        L.n_cfg = 1
        _lib.check(tq.lib.hr_queue_run(C.byref(L), C.c_void_p(torch.cuda.current_stream().cuda_stream)),
                   "hr_queue_run")
        if i % 32 == 31:
            torch.cuda.synchronize()
            print(f"queue: serial launches {i + 1}/{len(cases)}, {time.perf_counter() - t0:.1f} s", file=sys.stderr,
                  flush=True)
    torch.cuda.synchronize()
    t_s = time.perf_counter() - t0
    same = bool(np.array_equal(tq.tally.cpu().numpy(), tally))  # every tally, energies included, bit for bit
    return {"metric": "histories/s, 256-configuration campaign in one work-queue graph", "value": tq.total / t_q,
            "unit": "histories/s", "n_gpus": 1, "seconds": t_q, "histories": int(tq.total), "rows": rows,
            "rows_per_second": rows / t_q, "serial_seconds": t_s, "speedup_vs_serial_launches": t_s / t_q,
            "serial_counts_equal": same, "autotune": tuned, "dtype": "f64 transport (analog, no rows)",
            "data": "synthetic campaign (seed 2024)",
            "config": {"workload": "256 configurations, histories log-uniform 1e4-1e7 x %g, thickness 10 um-2 mm, "
                                   "1-30 MeV, graphite/Si/Ge/W, tilt 0-3 deg" % args.queue_scale}}


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(case_tables, n_all, n_one, repeats=3):
    """The FP64 C oracle (the CPU restatement of the reference path) on the
    host cores: ``repeats`` timed runs after one warm-up, median reported, both
    on every core the process may use (OpenMP over histories and bins) and on
    one thread, each on a bounded sample of the bench workload's histories."""
    from oracle import case as ocase

    threads = int(os.environ.get("OMP_NUM_THREADS", 0)) or len(os.sched_getaffinity(0))

    def timed(n_hist, nthr, first):
        ocase.run_case(case_tables, 10_000_000, 32, threads=nthr)  # warm-up (page-in, thread pool)
        runs, tims = [], []
        for r in range(repeats):
            tim = {}
            t0 = time.perf_counter()
            ocase.run_case(case_tables, first + r * n_hist, n_hist, threads=nthr, timings=tim)
            runs.append(time.perf_counter() - t0)
            tims.append(tim)
        k = int(np.argsort(runs)[len(runs) // 2])
        return n_hist / runs[k], runs, tims[k]

    v_all, runs_all, tim = timed(n_all, threads, 0)
    v_one, runs_one, tim1 = timed(n_one, 1, 50_000)
    return {
        "value": v_all,
        "unit": "histories/s",
        "cores": threads,
        "kind": "port",
        "sample": f"median of {repeats} warm runs of {n_all} histories each (disjoint ids) of the same case "
                  f"(transport + lines over {len(case_tables.n_hats)} rays + brem, {case_tables.E_grid.size} bins), "
                  f"FP64 C oracle, {threads} OpenMP threads: {', '.join(f'{t:.2f}' for t in runs_all)} s; median "
                  f"run: transport {tim['transport']:.2f} s, lines {tim['lines']:.2f} s, brem {tim['brem']:.2f} s",
        "single_thread": {"value": v_one, "unit": "histories/s", "cores": 1,
                          "sample": f"median of {repeats} warm runs of {n_one} histories: "
                                    f"{', '.join(f'{t:.2f}' for t in runs_one)} s; median run: transport "
                                    f"{tim1['transport']:.2f} s, lines {tim1['lines']:.2f} s, brem "
                                    f"{tim1['brem']:.2f} s",
                          "us_per_history": 1e6 / v_one},
        "cpu_model": _cpu_model(),
        "nproc": os.cpu_count(),
        "affinity_cpus": len(os.sched_getaffinity(0)),
    }


def live_cells(pipe, steps, launches):
    """Live (nonzero-weight) line x bin cells of the timed sinc^2 launches.

    Lines of weight zero (rejected resonances: grazing rays whose in-medium root
    does not settle) are skipped by the kernel, so only live cells are work.
    Counting them inside the timed region would add tensor work there, so the
    timed batches are recomputed afterwards (transport and line prologue are
    deterministic per batch) and their lines counted block by block, in the
    order the reduction launched them."""
    import torch

    from pyrite_amd.spectrum.lines import LINE_BLOCK

    sc = pipe.lines
    per_seg = sc.n_dir * len(sc.tables.wm)
    step = max(1, LINE_BLOCK // per_seg)
    total = 0.0
    n = 0
    def consume(e0, m, cap, rows, seg_count):
        nonlocal total, n
        seg = pipe.slot_records([(e0, m, cap, rows, seg_count)], int(seg_count.clamp(max=cap).sum()))
        n_seg = int(seg["L"].shape[0])
        for j0 in range(0, n_seg, step):
            _E_r, _a, w = sc.lines(seg, 1.0, j0, min(step, n_seg - j0))
            total += float((w != 0).sum().item()) * sc.E_grid.size
            n += 1
        pipe._scored()

    for s in range(steps):
        pipe.transport(s, compact=False, consume=consume)
        torch.cuda.synchronize()
    assert n == launches, (n, launches)
    return total


# HBM bytes per sincsq_uniform_kernel launch of the default workload (mean of the
# step's two launches), measured with rocprofv3 --pmc and a copy calibration.
SINC_TRAFFIC_BYTES = 265.5e6


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}): several GPUs would split each "
              f"batch by global electron id and add their FP64 tallies once, which is not launched here",
              file=sys.stderr)
        sys.exit(2)
    if args.workload != "hopg":
        run = {"si-stream": si_stream, "pixels": pixel_scoring, "coherent": coherent_case,
               "queue": queue_campaign}[args.workload]
        print(json.dumps(run(args)))
        return
    import torch

    from pyrite_amd.pipeline import CaseConfig, CasePipeline
    from pyrite_amd import roofline

    cfg = CaseConfig(Ne=args.ne, n_side=args.n_side, line_route=args.line_route)
    pipe = CasePipeline(cfg)
    # warm-up through the timed path itself (its two streams and their allocator pools)
    pipe.run_steps(1000, args.warmup, overlap=not args.no_overlap)
    torch.cuda.synchronize()
    pipe.spec_lines.zero_()
    pipe.spec_brem.zero_()
    pipe.histories = pipe.segments = 0
    pipe.lines.events = []
    pipe.lines.record_events = True
    phase = {"transport": 0.0, "score": 0.0}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps if args.breakdown else 0):
        if args.breakdown:
            ta = time.perf_counter()
            res = pipe.transport(s, compact=False)
            torch.cuda.synchronize()
            tb = time.perf_counter()
            pipe.score_records(pipe.records(res), cfg.Ne * args.steps)
            torch.cuda.synchronize()
            phase["transport"] += tb - ta
            phase["score"] += time.perf_counter() - tb
            pipe.histories += cfg.Ne
            pipe.segments += res.n_segments
    if not args.breakdown:
        # transport of batch s+1 overlaps the scoring of batch s (two HIP streams)
        pipe.run_steps(0, args.steps, overlap=not args.no_overlap)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    pipe.lines.record_events = False
    dt = t1 - t0
    hist = pipe.histories
    value = hist / dt
    # dominant kernel: sinc^2 reduction, event-timed per launch
    ms = [ev[0].elapsed_time(ev[1]) for ev, _ in pipe.lines.events] or [float("nan")]
    cells = live_cells(pipe, args.steps, len(ms))  # live (nonzero-weight) lines x bins of the timed launches
    avg_ms = sum(ms) / len(ms)
    flops_per_launch = roofline.SINC_CELL_OPS * cells / len(ms)
    achieved = flops_per_launch / (avg_ms * 1e-3) / 1e12
    line_bytes = 20  # E_r f64 + a f64 + w f32 per line, read once per launch
    n_E = cfg.n_bins
    lines_per_launch = cells / len(ms) / n_E
    out = {
        "metric": "histories/s + PXR-peak FOM 1/(sigma^2 t), learned-importance vs analog, 1 GPU",
        "value": value,
        "unit": "histories/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 transport / f64 phase + f32 profile / f64 tallies",
        "data": "synthetic (seeded pencil beam; analytic tables)",
        "config": {
            "workload": "plain method, HOPG 100 um, 5 MeV pencil, c-axis tilt 0.5 deg, (002) PXR/CBS + "
                        "BH brem, 1024 bins 1-50 keV, 10 mrad cone at 90 deg",
            "histories_per_step": cfg.Ne,
            "detector_rays": len(pipe.n_hats),
            "segments_per_history": pipe.segments / max(hist, 1),
            "parallelism": "single GPU",
        },
        "roofline": {
            "kernel": "hr_sincsq_reduce: sincsq_uniform_kernel + sum_partials_kernel (dense sinc^2 lines)",
            "bound": "valu",
            "pipe": "vector ALU, FP32 peak (the listed fp32 dense figure equals the VALU FP32 peak)",
            "achieved": achieved,
            "peak": roofline.PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / roofline.PEAK_FP32_TFLOPS,
            "traffic": SINC_TRAFFIC_BYTES if (args.ne, args.n_side, args.line_route) == (100_000, 3, "dense") else None,
            "traffic_source": "rocprofv3 --pmc FETCH_SIZE (x2.000) and WRITE_SIZE (x1.000), factors from a 1 GiB "
                              "copy; profiles/sincsq_uniform_hopg_5mev_pmc.md",
            "ops_per_cell": roofline.SINC_CELL_OPS,
            "live_cells_per_launch": cells / len(ms),
            "avg_launch_ms": avg_ms,
            "launches": len(ms),
            "algorithmic_bytes_per_launch": lines_per_launch * line_bytes + n_E * 8,
        },
    }
    if args.breakdown:
        print(json.dumps({"phase_s": phase, "segments": pipe.segments}), file=sys.stderr)
    if args.copy_calibration:
        src = torch.empty(1 << 30, dtype=torch.uint8, device=pipe.dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
    # the bench pipeline's arenas go before the figure-of-merit leg builds its own
    from pyrite_amd.pipeline import CaseTables

    host_tables = CaseTables(cfg)
    pipe.release()
    del pipe
    torch.cuda.empty_cache()
    if not args.no_fom:
        out["fom"] = learned_fom(cfg, args.fom_groups, args.fom_group_batches, args.fom_histories,
                                 overlap=not args.fom_sequential, train_batches=args.fom_train_batches)
        torch.cuda.empty_cache()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(host_tables, args.cpu_sample, args.cpu_sample_one)
        out["speedup_vs_cpu_baseline"] = value / out["cpu_baseline"]["value"]
        out["speedup_vs_cpu_single_thread"] = value / out["cpu_baseline"]["single_thread"]["value"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
