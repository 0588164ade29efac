"""One plain-method case on the FP64 CPU oracle: transport, then line + brem scoring.

TEST INFRASTRUCTURE ONLY: the checker for the end-to-end parity tests and the
CPU baseline leg of bench.py.  It consumes the same host tables the device
path uploads (layer/LUT tables, line tables, detector rays) and runs every
step in the oracle's FP64 C code with OpenMP over histories / bins.
"""

from __future__ import annotations

import time

import numpy as np

from . import oracle


def _transport_all(pipe_like, keys, pos, dirs, E, E_cut, threads, clock=None, cone=None):
    cfg = pipe_like.cfg
    cap = 512
    while True:
        rows, cnt, code, st = oracle.run_transport(
            pipe_like.layers, pipe_like.lut, keys, pos, dirs, E, E_cut, cap=cap, clock=clock,
            max_steps=cfg.max_steps, energy_model=1, straggle_on=cfg.straggling, threads=threads, cone=cone)
        if cnt.max(initial=0) <= cap:
            break
        cap = int(cnt.max() * 1.25) + 1
    if cone is not None:  # copies beyond the slots: rerun with room (the product raises past its bound)
        while st["cone"]["count"].max(initial=0) > cone["cap"]:
            cone = dict(cone, cap=int(st["cone"]["count"].max()))
            rows, cnt, code, st = oracle.run_transport(
                pipe_like.layers, pipe_like.lut, keys, pos, dirs, E, E_cut, cap=cap, clock=clock,
                max_steps=cfg.max_steps, energy_model=1, straggle_on=cfg.straggling, threads=threads, cone=cone)
    return oracle.compact(rows, cnt, cap), st


def run_case(pipe_like, start, n, threads=0, timings=None, dose_bins=0, coherent=False, cone=None):
    """Histories ``[start, start+n)`` of the case described by ``pipe_like``.

    ``pipe_like`` carries ``cfg``, ``layers``, ``lut``, ``beam_dir``,
    ``n_hats``, ``weights``, ``n_hat``, ``E_grid``, ``line_tables`` (a
    ``CasePipeline`` or ``CaseTables``).  Returns ``(lines, brem_sum, n_rows)``:
    the solid-angle line spectrum summed over these histories (not divided by
    Ne) and the raw brem sum ``sum n L dsigma/dk T`` at the central direction
    (divide by ``4 pi Ne`` and multiply by Omega for the cone).  With
    ``coherent`` the line spectrum is the coherent one: per direction and
    reflection the fields of all these histories' segments are summed, then
    squared (``sum_d w_d sum_g wm_g |F|^2``, again not divided by Ne).
    ``cone`` (a ``cone_split.ConeSplit``) runs detector-cone splitting: the
    primaries split, their copies (ladder weights, powers of two) are
    transported as one analog generation and each weight stratum of copy rows
    is scored with the plain scorers and scaled by its weight
    (``runner/cone_split_strata.py:25-190``).
    """
    cfg = pipe_like.cfg
    t0 = time.perf_counter()
    keys = oracle.stream_keys(cfg.seed, n, start=start)
    from pyrite_amd.transport.beam import beam_args

    from . import beam

    args = beam_args(cfg.beam, cfg.seed, start, n, cfg.E0_keV, np.deg2rad(cfg.tilt_deg))
    pos, dirs = beam.entry(args, n, start)
    E = beam.energies(args, n, start)
    E_cut = np.full(n, cfg.E_cut_keV)
    if cone is None:
        r, _ = _transport_all(pipe_like, keys, pos, dirs, E, E_cut, threads)
        strata = [(1.0, r)]
    else:
        if coherent or dose_bins:
            raise ValueError("cone splitting runs the incoherent line and brem tallies only")
        cdict = dict(n_hat=cone.n_hat, half_angle_rad=cone.half_angle_rad, target_weight=cone.target_weight,
                     min_energy_keV=cone.min_energy_keV,
                     angle_scale=0.0 if np.isinf(cone.angle_scale) else cone.angle_scale, cap=16)
        r, st = _transport_all(pipe_like, keys, pos, dirs, E, E_cut, threads, cone=cdict)
        cp = oracle.cone_collect(st["cone"], ladder=cone.weight_ladder)
        if cp["max_count"] > cone.max_copies_per_electron:
            raise RuntimeError("cone split copies above max_copies_per_electron")
        strata = [(1.0, r)]
        nc = cp["weight"].size
        if nc:
            rc, _ = _transport_all(pipe_like, cp["stream_key"], cp["pos"], cp["dir"], cp["E_keV"],
                                   np.full(nc, cfg.E_cut_keV), threads, clock=cp["clock"])
            row_w = cp["weight"][rc["seg_id"]] if rc["seg_len"].size else np.zeros(0)
            for W in np.unique(row_w):
                sel = row_w == W
                strata.append((float(W), {k: v[sel] for k, v in rc.items()}))
    t1 = time.perf_counter()
    if timings is not None:
        timings["transport"] = timings.get("transport", 0.0) + (t1 - t0)
    lines = np.zeros(pipe_like.E_grid.size)
    brem = np.zeros(pipe_like.E_grid.size)
    n_rows = 0
    for W, rs in strata:
        ln, br, nr = _score_rows(pipe_like, rs, threads, coherent, timings, t1)
        lines += W * ln
        brem += W * br
        n_rows += nr
    if dose_bins:
        from . import tally

        dose = tally.depth_dose(r, 0.0, float(cfg.thickness_ang) / dose_bins, dose_bins)
        return lines, brem, n_rows, dose
    return lines, brem, n_rows


def _score_rows(pipe_like, r, threads, coherent, timings, t1):
    cfg = pipe_like.cfg
    seg = {
        "r_mid": r["seg_mid"].astype(np.float32), "v_hat": r["seg_dir"].astype(np.float32),
        "L": r["seg_len"].astype(np.float32),
        "E": (0.5 * (r["seg_E"] + r["seg_E_end"])).astype(np.float32),
    }
    from pyrite_amd.spectrum.crystal import polarization_pairs

    tab = pipe_like.line_tables.as_dict()
    lines = np.zeros(pipe_like.E_grid.size)
    if coherent:
        from pyrite_amd.spectrum.crystal import CRYSTALS, decrement

        seg64 = {"r_mid": r["seg_mid"], "v_hat": r["seg_dir"], "L": r["seg_len"],
                 "E": 0.5 * (r["seg_E"] + r["seg_E_end"]), "t0": r["seg_t0"]}
        E_grid = pipe_like.E_grid
        dom = decrement(CRYSTALS[cfg.crystal], E_grid) * E_grid / 1973.269804
    for d, n_hat in enumerate(pipe_like.n_hats):
        es, ep = polarization_pairs(n_hat, tab["g"])
        if coherent:
            rec, _good = oracle.coherent_lines(seg64, cfg.thickness_ang, n_hat, tab, es, ep, threads=threads)
            for g in range(rec.shape[1]):
                F = oracle.coherent_field(rec[:, g, :], E_grid, dom, threads=threads)
                lines += pipe_like.weights[d] * tab["wm"][g] * (F**2).sum(0)
            continue
        E_r, a, w, good = oracle.cxr_lines(seg, cfg.thickness_ang, n_hat, tab, es, ep, threads=threads)
        w = np.where(good, w * pipe_like.weights[d], 0.0)
        oracle.sincsq_reduce(E_r, a, w, pipe_like.E_grid, out=lines, threads=threads)
    t2 = time.perf_counter()
    from pyrite_amd.spectrum import attenuation
    from pyrite_amd.transport.elements import TRANSPORT_ELEMENTS

    comp = list(cfg.composition)
    mu = attenuation.mu_total_inv_ang(comp, pipe_like.E_grid).astype(np.float32)
    brem = oracle.brem_reduce(seg, cfg.thickness_ang, pipe_like.n_hat,
                              [float(TRANSPORT_ELEMENTS[e]["Z"]) for e, _ in comp],
                              [float(x) * 1e24 for _, x in comp], pipe_like.E_grid, mu, threads=threads)
    t3 = time.perf_counter()
    if timings is not None:
        timings["lines"] = timings.get("lines", 0.0) + (t2 - t1)
        timings["brem"] = timings.get("brem", 0.0) + (t3 - t2)
    return lines, brem, len(seg["L"])
