"""``simulate_trajectories`` on the hip backend.

The flat entry mirrors the reference's ``_simulate_trajectories``
(``src/pyrite/montecarlo/transport/api.py:19-75`` defaults: ``E_cut_keV=5``,
``max_steps=400000``, ``energy_model="midpoint"``, ``max_dE_frac=0``,
``straggling=False``) for the physics the hip LUT core implements:
``elastic_model="sr"`` and ``elastic_model="mott"`` (the LUT's Browning rate
and SRD 64-calibrated screening, ``pyrite_amd.tables.mott``; the tables come
from ``PYRITE_MOTT_TABLES_DIR`` or ``mott_tables_dir``, and an energy range
outside any element's table is refused before transport,
``host_pass.py:595-611``), and ``elastic_model="elsepa"`` with
``elastic_tables[L][i]`` (ELSEPA native-grid tables, e.g. from
``pyrite_amd.tables.elsepa.stack_panels``): LUT rates from their log-log
rates and tabulated angles in the step, energies outside a table refused.

The returned mapping uses the reference's segment keys (``host_pass.py:1048-1144``):
``r_mid``, ``v_hat``, ``L_ang``, ``E_keV``, ``E_end_keV``, ``E_repr_keV``,
``t_ang``, ``t_end_ang``, ``elec_id``, ``layer``, ``flight_id``,
``substep_id``, ``event_kind``, counts and ``Ne``/``thickness_ang``.
"""

from __future__ import annotations

import numpy as np

from .. import _lib
from ..backend import require_device
from .core import (
    DEFAULT_CONFIG,
    DeviceTables,
    PerElectronTransportConfig,
    TransportControl,
    TransportStepLimitError,
    drive_batches,
    make_hip_transport_core,
)
from .beam import ENERGY_SPREAD_CLAMP_SIGMA, beam_args, fill_entry, flat_beam_spec, upper_beam_energy_keV
from .tables import build_layer_tables, build_transport_lut

MM_TO_ANG = 1.0e7


def _layers(thickness_ang, element, n_atoms_per_ang3, composition, layers):
    if layers is not None:
        return [(float(t), float(b), list(c)) for t, b, c in layers]
    if composition is None:
        if element is None or n_atoms_per_ang3 is None:
            raise ValueError("give element and n_atoms_per_ang3, composition, or layers")
        composition = [(element, float(n_atoms_per_ang3))]
    return [(0.0, float(thickness_ang), list(composition))]


def normalized_beam_dir(beam_dir):
    """``beam_entry.py:328-343``."""
    if beam_dir is None:
        beam_dir = np.array([0.0, 0.0, 1.0])
    beam_dir = np.asarray(beam_dir, dtype=float)
    if beam_dir.shape != (3,) or not np.all(np.isfinite(beam_dir)):
        raise ValueError("beam_dir must be a finite three-vector")
    nrm = np.linalg.norm(beam_dir)
    if not np.isfinite(nrm) or nrm == 0.0:
        raise ValueError("beam_dir must be nonzero")
    beam_dir = beam_dir / nrm
    if beam_dir[2] <= 1e-6:
        raise ValueError("beam_dir must point into the slab (z component > 0)")
    return beam_dir


INELASTIC_MODELS = ("continuous", "shell-soft-hard")
RADIATIVE_MODELS = ("uncoupled", "analytic-soft-hard")


def _hard_modes(tables, lut, dev, inelastic_model, inelastic_cutoff_eV, radiative_model, radiative_cutoff_eV,
                E_cut_min_keV, energy_model):
    """Device tables of the requested hard-event modes, or None (both off).

    Fail-closed like the reference (``hard_inelastic.py:50-75``,
    ``hard_radiative.py:694-741``): unknown models, cutoffs without their
    mode, non-midpoint runs and a radiative cutoff above the electron cutoff
    raise."""
    from .hard_events import DeviceHardModes, build_radiative_tables, build_shell_tables

    if inelastic_model not in INELASTIC_MODELS:
        raise ValueError(f"inelastic_model must be one of {', '.join(INELASTIC_MODELS)}")
    if radiative_model == "bremslib-soft-hard":
        raise NotImplementedError("radiative_model='bremslib-soft-hard' needs the fetched BremsLib tables; the hip "
                                  "backend offers 'analytic-soft-hard' (substitute tables)")
    if radiative_model not in RADIATIVE_MODELS:
        raise ValueError(f"radiative_model must be one of {', '.join(RADIATIVE_MODELS)}")
    shell_on = inelastic_model == "shell-soft-hard"
    rad_on = radiative_model == "analytic-soft-hard"
    if not shell_on and inelastic_cutoff_eV is not None:
        raise ValueError("inelastic_cutoff_eV requires inelastic_model='shell-soft-hard'")
    if not rad_on and radiative_cutoff_eV is not None:
        raise ValueError("radiative_cutoff_eV requires radiative_model='analytic-soft-hard'")
    if not (shell_on or rad_on):
        return None
    if energy_model != "midpoint":
        raise ValueError("the hard-event modes require energy_model='midpoint'")
    rad = None
    if rad_on:
        cut = E_cut_min_keV * 1e3 if radiative_cutoff_eV is None else float(radiative_cutoff_eV)
        if cut > E_cut_min_keV * 1e3:
            raise ValueError("radiative_cutoff_eV must not exceed the electron cutoff")
        rad = build_radiative_tables(tables, cut)
    shell = None
    if shell_on:
        cut = 1000.0 if inelastic_cutoff_eV is None else float(inelastic_cutoff_eV)
        shell = build_shell_tables(tables, lut.E_grid, cut)
    return DeviceHardModes(tables, lut, dev, rad=rad, shell=shell)


def resolve_cutoffs_by_electron(E_cut_by_electrons, E_cut_keV, Ne):
    """Per-electron cutoff energies [keV] as a checked ``(Ne,)`` array
    (``arguments.py:124-138``): ``E_cut_keV`` for every electron unless
    ``E_cut_by_electrons`` gives one each; all finite and positive."""
    if E_cut_by_electrons is None:
        out = np.full(int(Ne), float(E_cut_keV))
    else:
        out = np.asarray(E_cut_by_electrons, dtype=np.float64)
        if out.shape != (int(Ne),):
            raise ValueError(f"E_cut_by_electrons must have shape ({int(Ne)},), got {out.shape}")
    if not np.all(np.isfinite(out)) or not np.all(out > 0.0):
        raise ValueError("electron cutoff energies must be finite and strictly positive")
    return np.ascontiguousarray(out)


def simulate_trajectories(
    E0_keV,
    Ne,
    thickness_ang=None,
    element=None,
    n_atoms_per_ang3=None,
    E_cut_keV=5.0,
    seed=0,
    max_steps=400_000,
    elastic_model="sr",
    beam_dir=None,
    composition=None,
    layers=None,
    crystal_width_mm=None,
    crystal_height_mm=None,
    energy_model="midpoint",
    max_dE_frac=0.0,
    straggling=False,
    per_electron_config: PerElectronTransportConfig = DEFAULT_CONFIG,
    keep_segments_on_device=True,
    initial_state=None,
    electron_start=0,
    raise_on_step_limit=True,
    inelastic_model="continuous",
    inelastic_cutoff_eV=None,
    radiative_model="uncoupled",
    radiative_cutoff_eV=None,
    mott_tables_dir=None,
    atomic_electron_deflection=None,
    elastic_tables=None,
    energy_spread_frac=None,
    E_cut_by_electrons=None,
    beam_fwhm_mm=None,
    beam_fwhm_y_mm=None,
    transverse_distribution=None,
    tilt_polar_rad=0.0,
    tilt_azim_rad=0.0,
    transport_precision="auto",
    transport_lut_config=None,
    _electron_start=None,
    _cone_split=None,
    collect_diagnostics=False,
    transport_core="auto",
    _energy_range_keV=None,
    stopping_tables=None,
    inelastic_materials=None,
    bremslib_tables=None,
    gdf_source=None,
    secondary_threshold_eV=None,
    pair_production_model=None,
    positron_transport=False,
    mixed_elastic=None,
    groove=None,
    radiative_moments="table",
):
    """Transport ``Ne`` electrons of ``E0_keV`` into a slab (hip core).

    ``initial_state=(pos, dirs, E_keV)`` (host arrays) overrides the pencil
    beam, which enters at the origin along ``beam_dir``.  ``electron_start``
    addresses electrons ``[start, start+Ne)`` of a larger run: keys are
    ``splitmix64(seed + PHI*(e+1))`` of the global id, so blocks reproduce
    the matching slice of one run (``kinematics.py:226-245``).

    Beam keywords as the reference's: ``beam_fwhm_mm`` / ``beam_fwhm_y_mm``
    (Gaussian spot), ``transverse_distribution`` (a resolved Courant-Snyder
    mapping, ``transverse.py:74-96``), ``tilt_polar_rad`` / ``tilt_azim_rad``
    (the spot carried onto the tilted entrance face), directions about
    ``beam_dir``; all drawn per electron on the device (``hr_beam_entry``).
    ``transport_precision``: ``"auto"`` (FP64 here), ``"fp64"`` or ``"mixed"``.
    ``transport_lut_config``: ``intervals_per_decade``, ``min_points``,
    ``max_points`` of the LUT.

    ``collect_diagnostics``: ``result["transport_diagnostics"]``, percentile
    summaries of each flight's fractional energy loss, relative elastic-hazard
    change, clock error estimate and cutoff overshoot (``diagnostics.py``).

    ``E_cut_by_electrons``: one cutoff [keV] per electron in place of
    ``E_cut_keV`` (the tables span the lowest).

    ``energy_spread_frac``: RMS relative energy spread of the sampled beam,
    ``E0 (1 + f z)`` with ``z`` clamped to ``|z| <= 6`` in the reference's
    ``spawn(6)[5]`` namespace, drawn per electron on the device
    (``api.py:161-166`` of the reference); the tables span to ``E0 (1 + 6|f|)``.

    Hard-event modes (the reference's ``inelastic_model`` and
    ``radiative_model`` switches, ``transport/api.py:56-60,212-250``):
    ``inelastic_model="shell-soft-hard"`` runs explicit hard collisions above
    ``inelastic_cutoff_eV`` (default 1 keV) with two-moment soft losses below;
    ``radiative_model="analytic-soft-hard"`` runs explicit bremsstrahlung
    photons above ``radiative_cutoff_eV`` (default: the electron cutoff) with
    soft radiative stopping below.  Both run on this module's analytic
    substitute tables (``hard_events``); ``"bremslib-soft-hard"`` needs the
    fetched BremsLib tables and is refused.  Mode rows add
    ``hard_inelastic_W_keV``/``hard_inelastic_channel`` and
    ``hard_radiative_k_eV``/``hard_radiative_Z``.
    """
    import torch

    if elastic_model not in ("sr", "mott", "elsepa"):
        raise NotImplementedError(
            f"elastic_model={elastic_model!r}: the hip LUT core implements 'sr', 'mott' and 'elsepa'"
        )
    if energy_model not in ("midpoint", "frozen"):
        raise ValueError("energy_model must be 'midpoint' or 'frozen'")
    Ne = int(Ne)
    stack = _layers(thickness_ang, element, n_atoms_per_ang3, composition, layers)
    shell_cut = (1000.0 if inelastic_cutoff_eV is None else float(inelastic_cutoff_eV)) \
        if inelastic_model == "shell-soft-hard" else None
    tables = build_layer_tables(stack, elastic_model, mott_tables_dir, atomic_electron_deflection, shell_cut,
                                elastic_tables=elastic_tables)
    finite = crystal_width_mm is not None and crystal_height_mm is not None

    if _electron_start is not None:
        electron_start = int(_electron_start)
    # the reference's core selector: every device core of this backend is the hip LUT core
    if transport_core not in ("auto", "per-electron", "cuda", "hip"):
        raise NotImplementedError(f"transport_core={transport_core!r}: this backend runs the per-electron hip core "
                                  "('auto', 'per-electron', 'cuda' and 'hip' select it)")
    # inputs that need fetched data or subsystems this backend does not build fail closed
    unsupported = {"stopping_tables": stopping_tables, "inelastic_materials": inelastic_materials,
                   "bremslib_tables": bremslib_tables, "gdf_source": gdf_source,
                   "secondary_threshold_eV": secondary_threshold_eV, "pair_production_model": pair_production_model,
                   "mixed_elastic": mixed_elastic, "groove": groove}
    for key, value in unsupported.items():
        if value is not None:
            raise NotImplementedError(f"{key} is not supported by the hip backend (fetched tables or a subsystem "
                                      "outside this build); pass None")
    if positron_transport:
        raise NotImplementedError("positron_transport is not supported by the hip backend")
    if radiative_moments not in ("table", "exact"):
        raise ValueError("radiative_moments must be 'table' or 'exact'")
    if transport_precision not in ("auto", "fp64", "mixed"):
        raise ValueError("transport_precision must be 'auto', 'fp64' or 'mixed'")
    spread = float(energy_spread_frac or 0.0)
    spec = flat_beam_spec(beam_fwhm_mm, beam_fwhm_y_mm, transverse_distribution, spread)
    if initial_state is not None and (spec.kind != "pencil" or spread or tilt_polar_rad):
        raise ValueError("energy_spread_frac, the spot, Courant-Snyder and tilt keywords apply to the sampled beam, "
                         "not to an initial_state")
    if initial_state is None:
        bd = normalized_beam_dir(beam_dir)
        E_h = np.full(Ne, float(E0_keV))
    else:
        pos_h, dirs_h, E_h = (np.ascontiguousarray(x, dtype=np.float64) for x in initial_state)
    E_cut_h = resolve_cutoffs_by_electron(E_cut_by_electrons, E_cut_keV, Ne)
    # with a spread the energies are drawn on the device below; the checks and the
    # tables use its strict bounds E0 (1 -/+ 6 |f|) (beam_entry.py:251-261)
    E_lo_h = E_h * (1.0 - ENERGY_SPREAD_CLAMP_SIGMA * abs(spread))
    E_top = float(E_h.max()) if not spread else upper_beam_energy_keV(E0_keV, spread)
    if np.any(E_cut_h >= E_lo_h):
        raise ValueError("each electron cutoff energy must be below its initial energy")
    if elastic_model == "mott":
        from ..tables.mott import check_mott_coverage

        check_mott_coverage(tables.compositions, float(E_cut_h.min()), E_top, mott_tables_dir)
    if elastic_model == "elsepa":
        from ..tables.elsepa import check_coverage

        check_coverage(tables.elsepa, float(E_cut_h.min()), E_top)
    if initial_state is None:  # refuses a spread that could go non-positive, before any device work
        beam_args(spec, seed, electron_start, Ne, E0_keV, tilt_polar_rad, tilt_azim_rad, bd)
    lut_kw = {} if transport_lut_config is None else {
        k: v for k, v in dict(transport_lut_config).items()
        if k in ("intervals_per_decade", "min_points", "max_points", "rel_tol")}
    lo, hi = float(E_cut_h.min()), E_top
    if _energy_range_keV is not None:  # a block's population range, so its tables equal the single call's
        r_lo, r_hi = (float(v) for v in _energy_range_keV)
        if r_lo > lo or r_hi < hi:
            raise ValueError("_energy_range_keV must cover every transported electron")
        lo, hi = r_lo, r_hi
    lut = build_transport_lut(lo, hi, tables, **lut_kw)
    # host-side refusals above run before any device work
    backend = require_device()
    dev = backend.torch_device
    hard = _hard_modes(tables, lut, dev, inelastic_model, inelastic_cutoff_eV, radiative_model,
                       radiative_cutoff_eV, float(E_cut_h.min()), energy_model)
    if hard is not None:
        from .hard_events import lut_for

        lut = lut_for(lut, hard.shell)
    dtabs = DeviceTables(tables, lut, dev)
    control = TransportControl(
        max_steps=int(max_steps),
        energy_model=1 if energy_model == "midpoint" else 0,
        max_dE_frac=float(max_dE_frac),
        straggle_on=bool(straggling),
        finite_footprint=finite,
        width_ang=float(crystal_width_mm) * MM_TO_ANG if finite else 0.0,
        height_ang=float(crystal_height_mm) * MM_TO_ANG if finite else 0.0,
        precision="fp64" if transport_precision == "auto" else transport_precision,
    )
    if _cone_split is not None:
        from ..cone_split import ConeSplit, ConeSplitCore

        if not isinstance(_cone_split, ConeSplit):
            raise TypeError("_cone_split must be a pyrite_amd.cone_split.ConeSplit request")
        if hard is not None:
            raise ValueError("detector-cone splitting is not combined with the hard-event modes")
        core = ConeSplitCore(dtabs, control, _cone_split, dev)
    else:
        core = make_hip_transport_core(dtabs, control, hard)
    if hard is not None:
        hard.set_keys(seed, electron_start, Ne)

    keys = torch.empty(Ne, dtype=torch.uint64, device=dev)
    _lib.call("hr_stream_keys", int(seed) & (2**64 - 1), int(electron_start), Ne, keys.data_ptr(),
              None)
    clock = torch.zeros(Ne, dtype=torch.float64, device=dev)
    if initial_state is None:  # the sampled beam, drawn per electron on the device (counter-addressed)
        pos = torch.empty((Ne, 3), dtype=torch.float64, device=dev)
        dirs = torch.empty((Ne, 3), dtype=torch.float64, device=dev)
        E = torch.empty(Ne, dtype=torch.float64, device=dev)
        fill_entry(spec, seed, electron_start, Ne, E0_keV, tilt_polar_rad, pos, dirs, E, clock, tilt_azim_rad,
                   beam_dir=bd)
    else:
        pos = torch.as_tensor(pos_h, device=dev).contiguous()
        dirs = torch.as_tensor(dirs_h, device=dev).contiguous()
        E = torch.as_tensor(E_h, device=dev).contiguous()
    E_cut = torch.as_tensor(E_cut_h, device=dev)
    alive = torch.ones(Ne, dtype=torch.uint8, device=dev)
    stragg = torch.zeros(Ne, dtype=torch.float64, device=dev)
    # before transport mutates the state in place
    initial_E = E.cpu().numpy() if initial_state is None else None
    initial_r, initial_v = pos.cpu().numpy(), dirs.cpu().numpy()
    res = drive_batches(core, keys, (alive, clock, pos, dirs, E, E_cut), stragg, control, Ne, dev,
                        per_electron_config, extra_columns=() if hard is None else hard.row_columns)
    out = _assemble(res, Ne, tables, control, finite, elastic_model, atomic_electron_deflection,
                    E_h.copy() if initial_E is None else initial_E, initial_r, initial_v, E, pos, dirs, stragg,
                    straggling, energy_model, hard, raise_on_step_limit, max_steps)
    if collect_diagnostics:  # batching.py:98-217: percentile summaries of the per-flight errors
        from .diagnostics import flight_diagnostic_summary

        host = lambda v: v.cpu().numpy() if hasattr(v, "cpu") else np.asarray(v)  # noqa: E731
        out["transport_diagnostics"] = flight_diagnostic_summary(host(out["E_keV"]), host(out["L_ang"]),
                                                                 host(out["electron_id"]), host(out["layer"]),
                                                                 E_cut_h, tables)
    if _cone_split is not None:
        rec = _copy_generation(core, electron_start, E_cut, dtabs, control, dev, per_electron_config, tables, finite,
                               elastic_model, atomic_electron_deflection, straggling, energy_model,
                               raise_on_step_limit, max_steps)
        if rec["copies"] is not None and not keep_segments_on_device:
            rec["copies"] = _to_host(rec["copies"])
        out["cone_split"] = rec
    return out if keep_segments_on_device else _to_host(out)


def _copy_generation(core, electron_start, E_cut, dtabs, control, dev, config, tables, finite, elastic_model,
                     atomic_electron_deflection, straggling, energy_model, raise_on_step_limit, max_steps):
    """The primaries' copy records and the copies transported once, analog, as
    one generation with their own stream keys and their parents' cutoffs
    (``cone_split.py:457-513``); ``parent`` indexes this call's primaries."""
    import torch

    cp = core.copies()
    if cp is None:
        return {"parent": np.zeros(0, dtype=np.int64), "copies": None, "electron_start": int(electron_start)}
    # copies() offsets each sub-batch's parents by the sub-batch start within this
    # call, so they already index this call's primaries (0 .. Ne-1)
    parent = cp["parent"]
    parent_h = parent.cpu().numpy()
    if parent_h.size and (parent_h.min() < 0 or parent_h.max() >= E_cut.numel()):
        raise RuntimeError("cone splitting: copy parents outside this call's primaries")
    rec = {"parent": parent_h, "flight": cp["flight"].cpu().numpy(),
           "weight": cp["weight"].cpu().numpy(), "stream_keys": cp["stream_key"].cpu().numpy(),
           "r_ang": cp["pos"].cpu().numpy(), "v_hat": cp["dir"].cpu().numpy(), "E_keV": cp["E_keV"].cpu().numpy(),
           "t_ang": cp["clock"].cpu().numpy(), "killed": cp["killed"].cpu().numpy().astype(bool),
           "electron_start": int(electron_start), "copies": None}
    n = int(parent.numel())
    if n == 0:
        return rec
    state = (torch.ones(n, dtype=torch.uint8, device=dev), cp["clock"].clone(), cp["pos"].clone(),
             cp["dir"].clone(), cp["E_keV"].clone(), E_cut[parent].contiguous())
    stragg = torch.zeros(n, dtype=torch.float64, device=dev)
    E0s, r0, v0 = rec["E_keV"].copy(), rec["r_ang"].copy(), rec["v_hat"].copy()
    res = drive_batches(make_hip_transport_core(dtabs, control), cp["stream_key"], state, stragg, control, n, dev,
                        config)
    rec["copies"] = _assemble(res, n, tables, control, finite, elastic_model, atomic_electron_deflection, E0s, r0, v0,
                              state[4], state[2], state[3], stragg, straggling, energy_model, None,
                              raise_on_step_limit, max_steps)
    return rec


def _assemble(res, Ne, tables, control, finite, elastic_model, atomic_electron_deflection, initial_E, initial_r,
              initial_v, E, pos, dirs, stragg, straggling, energy_model, hard, raise_on_step_limit, max_steps):
    """The result mapping of one transport (the reference's segment keys)."""
    if raise_on_step_limit and res.counts["n_step_limited"]:
        raise TransportStepLimitError(
            f"{res.counts['n_step_limited']} histories reached max_steps={max_steps}"
        )
    r = res.rows
    out = {
        "r_mid": r["seg_mid"],
        "v_hat": r["seg_dir"],
        "L_ang": r["seg_len"],
        "E_keV": r["seg_E"],
        "E_start_keV": r["seg_E"],
        "t_ang": r["seg_t0"],
        "t_start_ang": r["seg_t0"],
        "elec_id": r["seg_id"],
        "electron_id": r["seg_id"],
        "layer": r["seg_lay"],
        "Ne": Ne,
        "thickness_ang": float(tables.z_total),
        "crystal_width_ang": control.width_ang if finite else None,
        "crystal_height_ang": control.height_ang if finite else None,
        "n_layers": tables.n_layers,
        "elastic_model": elastic_model,
        "atomic_electron_deflection": atomic_electron_deflection,
        "initial_E_keV": initial_E,
        "initial_r_ang": initial_r,
        "initial_v_hat": initial_v,
        "final_E_keV": E,
        "final_pos": pos,
        "final_dirs": dirs,
        "exit_code": res.exit_code,
        "seg_count": res.seg_count,
        "replays": res.replays,
        **res.counts,
    }
    if straggling:
        out["straggle_dE_keV"] = stragg
    if energy_model == "midpoint":
        out["E_end_keV"] = r["seg_E_end"]
        out["t_end_ang"] = r["seg_t_end"]
        out["E_repr_keV"] = 0.5 * (r["seg_E"] + r["seg_E_end"])
        out["flight_id"] = r["seg_flight"]
        out["substep_id"] = r["seg_substep"]
        out["event_kind"] = r["seg_event"]
    if hard is not None and hard.shell is not None:
        out["hard_inelastic_W_keV"] = r["seg_hard_W"]
        out["hard_inelastic_channel"] = r["seg_hard_ch"]
        out["inelastic"] = {"model": "shell-soft-hard", "cutoff_eV": hard.shell.cutoff_eV,
                            "tables": "analytic-substitute"}
    if hard is not None and hard.rad is not None:
        out["hard_radiative_k_eV"] = r["seg_rad_k"]
        out["hard_radiative_Z"] = r["seg_rad_Z"]
        out["radiative"] = {"model": "analytic-soft-hard", "cutoff_eV": hard.rad.cutoff_eV,
                            "plasma_energy_eV_by_layer": tuple(float(x) for x in hard.rad.plasma)}
    return out


def _to_host(out):
    return {k: (v.cpu().numpy() if hasattr(v, "cpu") else v) for k, v in out.items()}
