"""ctypes binding of libhiprite.so (the C ABI declared in include/hiprite.h).

The library is loaded from ``pyrite_amd/_lib/libhiprite.so`` (built in-tree by
``pyrite_amd.build``).  There is no fallback: when the library is missing or
fails to load, every entry point raises :class:`HipriteUnavailable`.
"""

from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libhiprite.so"

HR_OK = 0
HR_ERR_INVALID = 1
HR_ERR_HIP = 2
HR_ERR_OOM = 3
HR_ERR_UNSUPPORTED = 4

_dp = C.c_void_p  # device pointer


class HipriteUnavailable(RuntimeError):
    """libhiprite.so is not built or cannot be loaded."""


class HipriteError(RuntimeError):
    """A libhiprite entry point returned a non-zero hr_status."""

    def __init__(self, status: int, where: str, text: str):
        super().__init__(f"{where} failed with hr_status={status}: {text}")
        self.status = status


class HipriteOOM(HipriteError, MemoryError):
    """Device allocation failed inside libhiprite."""


class Workspace(C.Structure):
    """Mirror of ``hr_workspace`` (caller-owned device scratch of one call)."""

    _fields_ = [("ptr", _dp), ("bytes", C.c_size_t)]


class DeviceScratch:
    """Caller-owned device workspace of one scorer, grown to what a call needs.

    One instance serves the calls of one stream (calls on different streams
    that may overlap need different instances); growth frees the old buffer
    through torch's stream-ordered allocator."""

    def __init__(self, device):
        self.device = device
        self.buf = None

    def get(self, nbytes: int) -> "Workspace":
        import torch

        nbytes = int(nbytes)
        ws = Workspace()
        if nbytes <= 0:
            ws.ptr, ws.bytes = None, 0
            return ws
        if self.buf is None or self.buf.numel() < nbytes:
            self.buf = None
            self.buf = torch.empty(max(nbytes, 1 << 20), dtype=torch.uint8, device=self.device)
        ws.ptr, ws.bytes = self.buf.data_ptr(), self.buf.numel()
        return ws


def workspace_bytes(name: str, *args) -> int:
    """Bytes an entry point's ``*_workspace_bytes`` query asks for."""
    n = C.c_size_t(0)
    check(getattr(load(), name)(*args, C.byref(n)), name)
    return int(n.value)


class HardEvents(C.Structure):
    """Mirror of ``hr_hard_events`` (the hard-event modes of the transport step)."""

    _fields_ = [
        ("rad_on", C.c_int32), ("shell_on", C.c_int32),
        ("rad_key", _dp), ("lut_rad_soft", _dp), ("lut_rad_rate", _dp),
        ("rad_incident", _dp), ("rad_nominal", _dp), ("rad_top", _dp), ("rad_chi", _dp), ("rad_z", _dp),
        ("rad_ncm3", _dp), ("rad_nt", _dp), ("rad_plasma", _dp),
        ("rad_max_t", C.c_int32), ("rad_n_red", C.c_int32),
        ("rad_cutoff_eV", C.c_double),
        ("lut_sh_rate", _dp), ("lut_sh_omega2", _dp), ("lut_sh_ch_rate", _dp),
        ("sh_U", _dp), ("sh_W", _dp), ("sh_branch", _dp), ("sh_code", _dp), ("sh_nch", _dp),
        ("sh_max_ch", C.c_int32),
        ("sh_cutoff_eV", C.c_double),
        ("seg_hard_W", _dp), ("seg_hard_ch", _dp), ("seg_rad_k", _dp), ("seg_rad_Z", _dp),
    ]


class ConeSplitArgs(C.Structure):
    """Mirror of ``hr_cone_split`` (detector-cone splitting of elastic events)."""

    _fields_ = [
        ("on", C.c_int32), ("pad_", C.c_int32),
        ("nx", C.c_double), ("ny", C.c_double), ("nz", C.c_double),
        ("cos_c", C.c_double), ("sin_c", C.c_double),
        ("target_weight", C.c_double), ("min_energy_keV", C.c_double), ("angle_scale", C.c_double),
        ("split_key", _dp), ("cap", C.c_int64), ("count", _dp), ("killed", _dp),
        ("pos", _dp), ("dir", _dp), ("E_keV", _dp), ("clock", _dp), ("weight", _dp), ("flight", _dp),
    ]


class ElsepaArgs(C.Structure):
    """Mirror of ``hr_elsepa`` (tabulated elastic angles)."""

    _fields_ = [
        ("on", C.c_int32), ("n_mu", C.c_int32), ("guide_cells", C.c_int32), ("pad_", C.c_int32),
        ("start", _dp), ("length", _dp), ("logE", _dp), ("mu", _dp), ("cdf", _dp), ("pdf", _dp), ("guide", _dp),
    ]


class ConeCollectArgs(C.Structure):
    """Mirror of ``hr_cone_collect_args``."""

    _fields_ = [
        ("cone", C.POINTER(ConeSplitArgs)),
        ("e_start", C.c_int64), ("e_count", C.c_int64),
        ("ladder", C.c_int32), ("pad_", C.c_int32),
        ("out_cap", C.c_int64),
        ("parent", _dp), ("flight", _dp),
        ("pos", _dp), ("dir", _dp), ("E_keV", _dp), ("clock", _dp), ("weight", _dp),
        ("stream_key", _dp),
        ("n_out", _dp),
        ("ws", Workspace),
    ]


class TransportArgs(C.Structure):
    """Mirror of ``hr_transport_args`` (include/hiprite.h)."""

    _fields_ = [
        ("e_start", C.c_int64),
        ("e_count", C.c_int64),
        ("cap", C.c_int64),
        ("stream_key", _dp),
        ("max_steps", C.c_int32),
        ("energy_model", C.c_int32),
        ("max_dE_frac", C.c_double),
        ("straggle_on", C.c_int32),
        ("n_layers", C.c_int32),
        ("internal_bounds", _dp),
        ("z_total", C.c_double),
        ("finite_footprint", C.c_int32),
        ("width_ang", C.c_double),
        ("height_ang", C.c_double),
        ("L_nel", _dp),
        ("L_top", _dp),
        ("L_bot", _dp),
        ("max_el", C.c_int32),
        ("L_Z", _dp),
        ("L_J", _dp),
        ("L_k", _dp),
        ("L_coeff", _dp),
        ("L_E_cross", _dp),
        ("lut_log_E_min", C.c_double),
        ("lut_inv_dlogE", C.c_double),
        ("lut_n_energy", C.c_int32),
        ("lut_total_rate", _dp),
        ("lut_dEds", _dp),
        ("lut_inv_beta", _dp),
        ("lut_cdf", _dp),
        ("lut_alpha", _dp),
        ("alive", _dp),
        ("clock", _dp),
        ("pos", _dp),
        ("dirs", _dp),
        ("E_keV", _dp),
        ("E_cut", _dp),
        ("stragg_dE", _dp),
        ("seg_dir", _dp),
        ("seg_mid", _dp),
        ("seg_len", _dp),
        ("seg_E", _dp),
        ("seg_t0", _dp),
        ("seg_id", _dp),
        ("seg_lay", _dp),
        ("seg_E_end", _dp),
        ("seg_t_end", _dp),
        ("seg_flight", _dp),
        ("seg_substep", _dp),
        ("seg_event", _dp),
        ("seg_count", _dp),
        ("exit_code", _dp),
        ("block_size", C.c_int32),
        ("precision", C.c_int32),
        ("lut_total_rate_f", _dp), ("lut_dEds_f", _dp), ("lut_inv_beta_f", _dp), ("lut_cdf_f", _dp),
        ("lut_alpha_f", _dp),
        ("work", _dp),
        ("hard", HardEvents),
        ("cone", ConeSplitArgs),
        ("elsepa", ElsepaArgs),
    ]


class LineArgs(C.Structure):
    """Mirror of ``hr_line_args``."""

    _fields_ = [
        ("n_seg", C.c_int64),
        ("r_mid", _dp), ("v_hat", _dp), ("L", _dp), ("E", _dp),
        ("thickness", C.c_double),
        ("n_dir", C.c_int32),
        ("n_hat", _dp), ("dir_weight", _dp),
        ("n_g", C.c_int32),
        ("g", _dp), ("wm", _dp), ("row", _dp), ("e_s", _dp), ("e_p", _dp),
        ("n_tab", C.c_int64),
        ("E_tab", _dp), ("delta_tab", _dp),
        ("chi_re", _dp), ("chi_im", _dp), ("u_re", _dp), ("u_im", _dp),
        ("n_mu", C.c_int32),
        ("log_mu", _dp),
        ("lo_keep", C.c_double), ("hi_keep", C.c_double),
        ("scale", C.c_double),
        ("row_weight", _dp),
        ("E_r", _dp), ("a", _dp), ("w", _dp),
    ]


class SincsqArgs(C.Structure):
    """Mirror of ``hr_sincsq_args``."""

    _fields_ = [
        ("n_lines", C.c_int64),
        ("E_r", _dp), ("a", _dp), ("w", _dp),
        ("n_E", C.c_int64),
        ("E_grid", _dp),
        ("uniform", C.c_int32),
        ("E0", C.c_double), ("dE", C.c_double),
        ("out", _dp),
        ("ws", Workspace),
    ]


class BremArgs(C.Structure):
    """Mirror of ``hr_brem_args``."""

    _fields_ = [
        ("n_seg", C.c_int64),
        ("r_mid", _dp), ("v_hat", _dp), ("L", _dp), ("E_keV", _dp),
        ("thickness", C.c_double), ("n_hat", C.c_double * 3),
        ("n_el", C.c_int32),
        ("Z", C.c_double * 4), ("n_cm3", C.c_double * 4),
        ("n_E", C.c_int64),
        ("E_grid_eV", _dp), ("mu", _dp),
        ("row_weight", _dp),
        ("out", _dp),
        ("ws", Workspace),
    ]


QUEUE_TALLY = 12


class QueueConfig(C.Structure):
    """Mirror of ``hr_queue_config``."""

    _fields_ = [
        ("t", TransportArgs),
        ("seed", C.c_uint64),
        ("E0_keV", C.c_double), ("E_cut_keV", C.c_double),
        ("beam_dir", C.c_double * 3),
    ]


class QueueLaunch(C.Structure):
    """Mirror of ``hr_queue_launch``."""

    _fields_ = [
        ("configs", _dp), ("counter", _dp), ("n_cfg", C.c_int32),
        ("offsets", _dp), ("total", C.c_int64),
        ("tally", _dp),
        ("accumulate", C.c_int32), ("block_size", C.c_int32), ("per_claim", C.c_int32),
        ("precision", C.c_int32),
    ]


class WindowArgs(C.Structure):
    """Mirror of ``hr_window_args``."""

    _fields_ = [
        ("lines", C.POINTER(LineArgs)),
        ("n_E", C.c_int64),
        ("E0", C.c_double), ("dE", C.c_double),
        ("x_c", C.c_double),
        ("out", _dp),
        ("ws", Workspace),
    ]


class CoherentArgs(C.Structure):
    """Mirror of ``hr_coherent_args``."""

    _fields_ = [
        ("lines", C.POINTER(LineArgs)),
        ("n", C.c_int64), ("rows", C.c_int64),
        ("seg_count", _dp), ("row_offset", _dp),
        ("seg_mid", _dp), ("seg_dir", _dp), ("seg_len", _dp),
        ("seg_E", _dp), ("seg_E_end", _dp), ("seg_t0", _dp),
        ("first", C.c_int64), ("ld", C.c_int64),
        ("rec", _dp),
        ("n_lines", C.c_int64),
    ]


class CoherentFieldArgs(C.Structure):
    """Mirror of ``hr_coherent_field_args``."""

    _fields_ = [
        ("rec", _dp), ("ld", C.c_int64), ("n_lines", C.c_int64), ("n_rows", C.c_int32),
        ("n_E", C.c_int64), ("E_grid", _dp), ("dom", _dp), ("field", _dp),
        ("ws", Workspace),
    ]


class CoherentGroupedArgs(C.Structure):
    """Mirror of ``hr_coherent_grouped_args``."""

    _fields_ = [
        ("field", CoherentFieldArgs), ("group_start", _dp), ("n_groups", C.c_int64), ("grouped", _dp),
    ]


class BeamArgs(C.Structure):
    """Mirror of ``hr_beam_args``."""

    _fields_ = [
        ("start", C.c_int64), ("n", C.c_int64),
        ("mode", C.c_int32),
        ("root", C.c_uint64),
        ("sigma_x", C.c_double), ("sigma_y", C.c_double),
        ("twiss_x", C.c_double * 3), ("twiss_y", C.c_double * 3),
        ("projected", C.c_int32),
        ("e_x", C.c_double * 3), ("e_y", C.c_double * 3), ("beam_dir", C.c_double * 3),
        ("basis", C.c_double * 9),
        ("E0_keV", C.c_double),
        ("pos", _dp), ("dirs", _dp), ("E_keV", _dp), ("clock", _dp),
        ("energy_spread_frac", C.c_double), ("energy_root", C.c_uint64),
        ("axis", C.c_double * 3),
    ]


class VrWindowArgs(C.Structure):
    """Mirror of ``hr_vr_window_args``."""

    _fields_ = [
        ("n", C.c_int64),
        ("exit_code", _dp), ("log_importance", _dp), ("weight", _dp), ("key", _dp),
        ("log_I_ref", C.c_double), ("w_min", C.c_double), ("w_max", C.c_double),
        ("ratio_lo", C.c_double), ("ratio_hi", C.c_double),
        ("n_max", C.c_int32),
        ("count", _dp), ("w_child", _dp),
        ("weight2", _dp), ("w2_child", _dp),
        ("rows_near", C.c_int32), ("rows_far", C.c_int32), ("far_ratio", C.c_double), ("rows_next", _dp),
    ]


class VrSpawnArgs(C.Structure):
    """Mirror of ``hr_vr_spawn_args``."""

    _fields_ = [
        ("n", C.c_int64),
        ("count", _dp), ("offset", _dp), ("w_child", _dp),
        ("pos", _dp), ("dirs", _dp), ("E", _dp), ("clock", _dp),
        ("key", _dp), ("prim", _dp),
        ("out_pos", _dp), ("out_dirs", _dp), ("out_E", _dp), ("out_clock", _dp), ("out_weight", _dp),
        ("out_key", _dp), ("out_prim", _dp),
        ("w2_child", _dp), ("out_weight2", _dp),
        ("rows_next", _dp), ("out_rows", _dp),
    ]


class DoseArgs(C.Structure):
    """Mirror of ``hr_dose_args``."""

    _fields_ = [
        ("n", C.c_int64), ("rows", C.c_int64), ("seg_count", _dp),
        ("seg_mid", _dp), ("seg_dir", _dp), ("seg_len", _dp), ("seg_E", _dp), ("seg_E_end", _dp),
        ("seg_event", _dp), ("weight", _dp), ("scale", C.c_double), ("z0", C.c_double), ("dz", C.c_double),
        ("n_bins", C.c_int32), ("out", _dp),
    ]


class Continuation(C.Structure):
    """Mirror of ``hr_continuation``."""

    _fields_ = [
        ("history", _dp), ("tau", _dp), ("draws", _dp), ("nseg", _dp),
        ("flight", _dp), ("substep", _dp), ("steps", _dp),
    ]


LANE_EMPTY = -1


class VrGatherArgs(C.Structure):
    """Mirror of ``hr_vr_gather_args``."""

    _fields_ = [
        ("n", C.c_int64), ("rows", C.c_int64),
        ("seg_count", _dp), ("row_offset", _dp), ("weight", _dp), ("prim", _dp),
        ("seg_mid", _dp), ("seg_dir", _dp), ("seg_len", _dp), ("seg_E", _dp), ("seg_E_end", _dp),
        ("r_mid", _dp), ("v_hat", _dp), ("L", _dp), ("E", _dp),
        ("row_weight", _dp), ("row_prim", _dp),
        ("weight2", _dp), ("row_weight2", _dp),
        ("n_out", C.c_int64),
    ]


class MlpNet(C.Structure):
    """Mirror of ``hr_mlp``."""

    _fields_ = [("params", _dp), ("w1b", _dp), ("w2b", _dp)]


# name -> (restype, argtypes); every symbol include/hiprite.h declares.
SIGNATURES = {
    "hr_last_error": (C.c_char_p, []),
    "hr_version": (C.c_char_p, []),
    "hr_device_count": (C.c_int32, []),
    "hr_stream_keys": (C.c_int32, [C.c_uint64, C.c_int64, C.c_int64, _dp, _dp]),
    "hr_transport_run": (C.c_int32, [C.POINTER(TransportArgs), _dp]),
    "hr_cone_collect_workspace_bytes": (C.c_int64, [C.c_int64]),
    "hr_cone_collect": (C.c_int32, [C.POINTER(ConeCollectArgs), _dp]),
    "hr_cone_split_keys": (C.c_int32, [_dp, C.c_int64, _dp, _dp]),
    "hr_cone_roulette": (C.c_int32, [C.c_int64, _dp, _dp, C.c_double, C.c_double, _dp, _dp, _dp]),
    "hr_cone_event": (C.c_int32, [C.POINTER(ConeSplitArgs), C.c_int64, _dp, _dp, _dp, _dp, _dp, _dp]),
    "hr_cone_event_tab": (C.c_int32, [C.POINTER(ConeSplitArgs), C.POINTER(ElsepaArgs), C.c_int64, _dp, _dp, _dp, _dp,
                                      _dp, _dp]),
    "hr_radiative_lut": (C.c_int32, [C.POINTER(HardEvents), C.c_int32, C.c_int32, _dp, _dp, C.c_int64, _dp, _dp,
                                     _dp]),
    "hr_radiative_sample": (C.c_int32, [C.POINTER(HardEvents), C.c_int32, C.c_double, C.c_int64, _dp, _dp, _dp,
                                        _dp]),
    "hr_roulette_uniforms": (C.c_int32, [C.c_uint64, C.c_int64, C.c_int64, _dp, _dp]),
    "hr_line_screen": (C.c_int32, [_dp, C.c_int64, C.c_int64, C.c_int64, C.c_double, C.c_double, C.c_uint64,
                                   _dp, _dp, _dp]),
    "hr_depth_dose": (C.c_int32, [C.POINTER(DoseArgs), _dp]),
    "hr_transport_continue": (C.c_int32, [C.POINTER(TransportArgs), C.POINTER(Continuation), _dp]),
    "hr_cxr_lines": (C.c_int32, [C.POINTER(LineArgs), _dp]),
    "hr_sincsq_reduce": (C.c_int32, [C.POINTER(SincsqArgs), _dp]),
    "hr_sincsq_workspace_bytes": (C.c_int32, [C.POINTER(SincsqArgs), C.POINTER(C.c_size_t)]),
    "hr_sincsq_reduce_static": (C.c_int32, [C.POINTER(SincsqArgs), _dp]),
    "hr_brem_reduce": (C.c_int32, [C.POINTER(BremArgs), _dp]),
    "hr_brem_workspace_bytes": (C.c_int32, [C.POINTER(BremArgs), C.POINTER(C.c_size_t)]),
    "hr_mlp_param_count": (C.c_int32, []),
    "hr_mlp_workspace_bytes": (C.c_size_t, [C.c_int64]),
    "hr_mlp_shadow": (C.c_int32, [_dp, _dp, _dp, _dp]),
    "hr_mlp_forward": (C.c_int32, [C.POINTER(MlpNet), _dp, C.c_int64, _dp, _dp]),
    "hr_mlp_grad": (C.c_int32, [C.POINTER(MlpNet), _dp, _dp, C.c_int64, _dp, _dp, _dp, _dp]),
    "hr_mlp_adam": (C.c_int32, [C.POINTER(MlpNet), _dp, _dp, _dp, C.c_float, C.c_int32, _dp]),
    "hr_beam_entry": (C.c_int32, [C.POINTER(BeamArgs), _dp]),
    "hr_queue_run": (C.c_int32, [C.POINTER(QueueLaunch), _dp]),
    "hr_queue_graph_create": (C.c_int32, [C.POINTER(QueueLaunch), C.POINTER(C.c_void_p)]),
    "hr_queue_graph_launch": (C.c_int32, [_dp, _dp]),
    "hr_queue_graph_destroy": (C.c_int32, [_dp]),
    "hr_window_reduce": (C.c_int32, [_dp, _dp, _dp, C.c_int64, C.c_int64, C.c_double, C.c_double, C.c_double,
                                     _dp, Workspace, _dp]),
    "hr_window_reduce_workspace_bytes": (C.c_int32, [C.c_int64, C.c_int64, C.POINTER(C.c_size_t)]),
    "hr_window_spectra": (C.c_int32, [C.POINTER(WindowArgs), _dp]),
    "hr_window_spectra_workspace_bytes": (C.c_int32, [C.POINTER(WindowArgs), C.POINTER(C.c_size_t)]),
    "hr_coherent_lines": (C.c_int32, [C.POINTER(CoherentArgs), _dp]),
    "hr_coherent_field": (C.c_int32, [C.POINTER(CoherentFieldArgs), _dp]),
    "hr_coherent_field_workspace_bytes": (C.c_int32, [C.POINTER(CoherentFieldArgs), C.POINTER(C.c_size_t)]),
    "hr_coherent_finalize": (C.c_int32, [_dp, C.c_int32, C.c_int32, _dp, _dp, C.c_int64, C.c_double, _dp, _dp]),
    "hr_coherent_grouped": (C.c_int32, [C.POINTER(CoherentGroupedArgs), _dp]),
    "hr_coherent_grouped_workspace_bytes": (C.c_int32, [C.POINTER(CoherentGroupedArgs), C.POINTER(C.c_size_t)]),
    "hr_coherent_blend_finalize": (C.c_int32, [_dp, _dp, C.c_int32, C.c_int32, _dp, _dp, C.c_int64, _dp, C.c_double,
                                               C.c_double, _dp, _dp]),
    "hr_vr_round": (C.c_int32, [C.POINTER(TransportArgs), C.c_int32, _dp, _dp]),
    "hr_vr_features": (C.c_int32, [C.c_int64, _dp, _dp, _dp, _dp, C.c_double, C.c_double, C.c_double, _dp, _dp]),
    "hr_vr_window": (C.c_int32, [C.POINTER(VrWindowArgs), _dp]),
    "hr_vr_spawn": (C.c_int32, [C.POINTER(VrSpawnArgs), _dp]),
    "hr_vr_gather": (C.c_int32, [C.POINTER(VrGatherArgs), _dp]),
    "hr_line_roulette": (C.c_int32, [_dp, C.c_int64, C.c_int32, C.c_int32, _dp, C.c_double, C.c_uint64,
                                     C.c_int64, _dp]),
    "hr_line_mass": (C.c_int32, [_dp, _dp, _dp, C.c_int64, C.c_int32, C.c_int32, C.c_double, C.c_double,
                                 C.c_double, _dp, _dp]),
}

_LIB = None


def load(path: Path | None = None):
    """Load libhiprite.so once; raise :class:`HipriteUnavailable` when absent."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = Path(path) if path is not None else LIB_PATH
    if not p.exists():
        raise HipriteUnavailable(
            f"{p} is not built; run `python -m pyrite_amd.build` (hipcc, gfx950)"
        )
    try:
        lib = C.CDLL(str(p))
    except OSError as exc:  # pragma: no cover - depends on the ROCm install
        raise HipriteUnavailable(f"cannot load {p}: {exc}") from exc
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _LIB = lib
    return lib


def check(status: int, where: str) -> None:
    """Raise the typed error for a non-zero hr_status."""
    if status == HR_OK:
        return
    text = (load().hr_last_error() or b"").decode(errors="replace")
    if status == HR_ERR_OOM:
        raise HipriteOOM(status, where, text)
    raise HipriteError(status, where, text)


def call(name: str, *args) -> None:
    """Call a status-returning entry point and raise on failure."""
    check(getattr(load(), name)(*args), name)
