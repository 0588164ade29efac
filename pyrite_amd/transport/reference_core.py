"""The reference's positional per-electron LUT core contract, on ``hr_transport_run``.

PyRITE's batch driver calls its per-electron LUT core positionally
(``transport/core_adapters.py:26-52``; driver ``batching.py:532-555``):

    core(run, control, geometry, lut_args, materials, state, segments, pe_out,
         straggling, inelastic_args=())

with ``run = (e_start, e_count, cap, stream_key)``, ``control = (max_steps,
max_segments, elastic_model_code, energy_model_code, max_dE_frac)``,
``geometry = (n_layers, internal_bounds, z_total, finite_footprint,
width_ang, height_ang, L_nel, L_top, L_bot)``, ``lut_args = (log_E_min,
inv_dlogE, n_energy, total_rate, dEds, inv_beta, cdf, alpha)``, ``materials
= (L_Js, L_Zs, L_ks, L_coeffs, L_E_cross, sbethe...)``, ``state = (alive,
clock, pos, dirs, E_keV, E_cut)``, ``segments`` the twelve midpoint row
columns (``seg_dir, seg_mid, seg_len, seg_E, seg_t0, seg_id, seg_lay,
seg_E_end, seg_t_end, seg_flight, seg_substep, seg_event``), ``pe_out =
(seg_count, exit_code)`` and ``straggling = (straggle_on, stragg_dE)``
(``cores.py:184-300``).  ``make_reference_core`` returns a callable with that
signature whose array arguments are device tensors (the reference passes
CuPy arrays on its GPU path); it maps them onto one ``hr_transport_args``
and launches on torch's current stream.  The core mutates ``pos``, ``dirs``,
``E_keV`` and ``clock`` in place, writes row ``s`` of local electron ``i`` to
slot ``i*cap + s`` while ``s < cap`` and reports the true row count, as the
reference's does.  ``elastic_model_code`` 0 (screened Rutherford) and 1
(Mott) both run: the reference's LUT kernel takes the code but never branches
on it (``_jit_lut_kernel.py:74``), because the LUT columns carry the model
(``tables.build_layer_tables(..., elastic_model=...)``).  Unsupported modes
fail closed: the ELSEPA code 2 (its tables enter with the layer tables, not
in the positional ``mott_group`` layout) or a nonzero ``sbethe_on`` raises
``NotImplementedError``; the shell and radiative modes of this backend take a
``hard_events.DeviceHardModes`` through ``hard=`` (their tables live on the
transport grid, not in the reference's ``inelastic_args`` layout, which is
refused).
"""

from __future__ import annotations

import ctypes as C

from .. import _lib

_SEGMENT_COLUMNS = ("seg_dir", "seg_mid", "seg_len", "seg_E", "seg_t0", "seg_id", "seg_lay", "seg_E_end",
                    "seg_t_end", "seg_flight", "seg_substep", "seg_event")
_SEGMENT_DTYPES = ("float64", "float64", "float64", "float64", "float64", "int64", "int16", "float64", "float64",
                   "int64", "int64", "int8")
_STATE = ("alive", "clock", "pos", "dirs", "E_keV", "E_cut")
_STATE_DTYPES = ("uint8", "float64", "float64", "float64", "float64", "float64")


def _ptr(t, name, dtype=None, n=None):
    if t is None:
        raise ValueError(f"reference core: {name} is missing")
    if not hasattr(t, "data_ptr"):
        raise TypeError(f"reference core: {name} must be a device tensor, got {type(t).__name__}")
    if dtype is not None and str(t.dtype).replace("torch.", "") != dtype:
        raise TypeError(f"reference core: {name} must be {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"reference core: {name} must be contiguous")
    if n is not None and t.numel() < n:
        raise ValueError(f"reference core: {name} holds {t.numel()} values, needs {n}")
    return int(t.data_ptr())


def build_args(run, control, geometry, lut_args, materials, state, segments, pe_out, straggling,
               precision="fp64", block_size=0):
    """``hr_transport_args`` of one positional call (no launch)."""
    e_start, e_count, cap, stream_key = run
    max_steps, _max_segments, elastic_code, energy_code, max_dE_frac = control
    if int(elastic_code) not in (0, 1):
        raise NotImplementedError(f"elastic_model_code={int(elastic_code)}: the positional adapter runs the "
                                  "LUT-carried models, screened Rutherford (0) and Mott (1); ELSEPA (2) tables enter "
                                  "with the layer tables (build_layer_tables(..., 'elsepa', elastic_tables=...))")
    if int(energy_code) not in (0, 1):
        raise ValueError("energy_model_code must be 0 (frozen) or 1 (midpoint)")
    n_layers, internal_bounds, z_total, finite, width, height, L_nel, L_top, L_bot = geometry
    log_E_min, inv_dlogE, n_energy, total_rate, dEds, inv_beta, cdf, alpha = lut_args
    L_J, L_Z, L_k, L_coeff, L_E_cross = materials[:5]
    if len(materials) > 5 and materials[5]:
        raise NotImplementedError("sbethe_on: SBETHE stopping tables are not available to the hip LUT core")
    seg_count, exit_code = pe_out
    straggle_on, stragg_dE = straggling
    if len(segments) != 12:
        raise ValueError(f"reference core: segments must hold the 12 midpoint columns, got {len(segments)}")
    e_start, e_count, cap = int(e_start), int(e_count), int(cap)
    n_layers = int(n_layers)
    max_el = int(L_Z.shape[-1]) if L_Z.dim() > 1 else int(L_Z.numel() // max(n_layers, 1))
    a = _lib.TransportArgs()
    a.e_start, a.e_count, a.cap = e_start, e_count, cap
    Ne = e_start + e_count
    a.stream_key = _ptr(stream_key, "stream_key", "uint64", Ne)
    a.max_steps, a.energy_model, a.max_dE_frac = int(max_steps), int(energy_code), float(max_dE_frac)
    a.straggle_on = int(bool(straggle_on))
    a.n_layers = n_layers
    a.internal_bounds = _ptr(internal_bounds, "internal_bounds", "float64")
    a.z_total = float(z_total)
    a.finite_footprint, a.width_ang, a.height_ang = int(bool(finite)), float(width), float(height)
    a.L_nel = _ptr(L_nel, "L_nel", "int32", n_layers)
    a.L_top, a.L_bot = _ptr(L_top, "L_top", "float64", n_layers), _ptr(L_bot, "L_bot", "float64", n_layers)
    a.max_el = max_el
    for name, t in (("L_Z", L_Z), ("L_J", L_J), ("L_k", L_k), ("L_coeff", L_coeff), ("L_E_cross", L_E_cross)):
        setattr(a, name, _ptr(t, name, "float64", n_layers * max_el))
    nE = int(n_energy)
    a.lut_log_E_min, a.lut_inv_dlogE, a.lut_n_energy = float(log_E_min), float(inv_dlogE), nE
    a.lut_total_rate = _ptr(total_rate, "lut_total_rate", "float64", n_layers * nE)
    a.lut_dEds = _ptr(dEds, "lut_dEds", "float64", n_layers * nE)
    a.lut_inv_beta = _ptr(inv_beta, "lut_inv_beta", "float64", nE)
    a.lut_cdf = _ptr(cdf, "lut_cdf", "float64", n_layers * max_el * nE)
    a.lut_alpha = _ptr(alpha, "lut_alpha", "float64", n_layers * max_el * nE)
    for name, dt, t in zip(_STATE, _STATE_DTYPES, state):
        setattr(a, name, _ptr(t, name, dt, Ne * (3 if name in ("pos", "dirs") else 1)))
    a.stragg_dE = _ptr(stragg_dE, "stragg_dE", "float64", Ne) if straggle_on else 0
    n_slots = e_count * cap
    for name, dt, t in zip(_SEGMENT_COLUMNS, _SEGMENT_DTYPES, segments):
        width = 3 if name in ("seg_dir", "seg_mid") else 1
        setattr(a, name, _ptr(t, name, dt, n_slots * width) if n_slots else 0)
    a.seg_count = _ptr(seg_count, "seg_count", "int64", e_count)
    a.exit_code = _ptr(exit_code, "exit_code", "int8", e_count)
    a.block_size = int(block_size)
    a.precision = {"fp64": 0, "mixed": 1}[precision]
    return a


def make_reference_core(hard=None, precision="fp64"):
    """``core(run, control, geometry, lut_args, materials, state, segments, pe_out,
    straggling, inelastic_args=(), *, work=None, stream=None)``.

    ``work`` is the launch's 8-byte claim counter (an int64 device tensor the
    driver owns; the core keeps one of its own when it is omitted).  With
    ``precision="mixed"`` the FP32 LUT copies are made once per LUT tensor."""
    import torch

    own = {}
    f32 = {}

    def core(run, control, geometry, lut_args, materials, state, segments, pe_out, straggling, inelastic_args=(),
             *, work=None, stream=None, block_size=0):
        if inelastic_args:
            raise NotImplementedError("inelastic_args in the reference layout: pass hard=DeviceHardModes to "
                                      "make_reference_core (shell tables on the transport LUT grid)")
        a = build_args(run, control, geometry, lut_args, materials, state, segments, pe_out, straggling,
                       precision, block_size)
        if precision == "mixed":
            names = ("lut_total_rate_f", "lut_dEds_f", "lut_inv_beta_f", "lut_cdf_f", "lut_alpha_f")
            for name, t in zip(names, (lut_args[3], lut_args[4], lut_args[5], lut_args[6], lut_args[7])):
                key = (name, int(t.data_ptr()))
                if key not in f32:
                    f32[key] = t.float()
                setattr(a, name, int(f32[key].data_ptr()))
        if hard is not None and hard.on:
            cols = dict(zip(_SEGMENT_COLUMNS, segments))
            hard.fill(a.hard, cols)
        if work is None:
            dev = segments[0].device
            if dev not in own:
                own[dev] = torch.zeros(1, dtype=torch.int64, device=dev)
            work = own[dev]
        a.work = _ptr(work, "work", "int64", 1)
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        _lib.check(_lib.load().hr_transport_run(C.byref(a), C.c_void_p(stream)), "hr_transport_run")

    return core


def reference_arguments(tables, control, run, state, rows, pe_out, stragg_dE):
    """The positional tuples of one call, from this package's ``DeviceTables``,
    ``TransportControl`` and row dict (used by the tests and by a host that
    drives the reference's batch driver with device tensors)."""
    lt, lut = tables.layers, tables.lut
    dev = tables.total_rate.device
    import torch

    L_nel = tables.L_nel
    geometry = (lt.n_layers, tables.internal_bounds, float(lt.z_total), int(control.finite_footprint),
                float(control.width_ang), float(control.height_ang), L_nel, tables.L_top, tables.L_bot)
    lut_args = (float(lut.log_E_min), float(lut.inv_dlogE), int(lut.n_energy), tables.total_rate, tables.dEds,
                tables.inv_beta, tables.cdf, tables.alpha)
    materials = (tables.L_J, tables.L_Z, tables.L_k, tables.L_coeff, tables.L_E_cross, 0,
                 torch.zeros(1, dtype=torch.int32, device=dev), None, None)
    ctl = (int(control.max_steps), 0, 0, int(control.energy_model), float(control.max_dE_frac))
    segments = tuple(rows[k] for k in _SEGMENT_COLUMNS)
    return run, ctl, geometry, lut_args, materials, state, segments, pe_out, (bool(control.straggle_on), stragg_dE)


__all__ = ["build_args", "make_reference_core", "reference_arguments"]
