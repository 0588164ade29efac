"""Incoherent PXR/CBS line spectrum and BH+Elwert bremsstrahlung on the hip backend.

Mirrors the reference's scoring entry points for a laterally infinite slab:

* ``mc_cxr_spectrum_solid_angle`` -- ``sum_i w_i mc_cxr_spectrum(n_hat=n_i)``
  in photons/eV/electron (``montecarlo/spectrum/cxr/_spectrum.py:401-480``),
  through the batched incoherent route (``cxr/_batched.py:461-1000``):
  per-(segment, reflection, direction) lines from ``hr_cxr_lines``, then the
  dense sinc^2 reduction ``hr_sincsq_reduce`` (``cxr_jit_kernel.py:561``);
* ``mc_brem_spectrum`` -- track-length BH+Elwert density per sr per electron
  (``montecarlo/spectrum/brem.py:829-1239``) via ``hr_brem_reduce``.

Host code builds the per-case tables once (FP32, the reference's device REAL)
and uploads them; segments stay on the device between transport and scoring.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from .. import _lib
from ..transport.elements import TRANSPORT_ELEMENTS
from . import attenuation
from .crystal import CRYSTALS, coupling_tables, decrement, polarization_pairs, tabulation_grid

# Lines per device pass (bounds the E_r/a/w buffers to ~0.7 GB).
LINE_BLOCK = 1 << 25


def _stream():
    """torch's current HIP stream as a raw handle (all launches go there)."""
    import torch

    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def keep_window(E_grid):
    """Resonance keep window: the axis padded by 20% of its span (``_kernels.py:110-127``)."""
    lo, hi = float(E_grid[0]), float(E_grid[-1])
    pad = 0.2 * (hi - lo)
    return lo - pad, hi + pad


@dataclass
class LineTables:
    """Host (numpy) tables of one case: reflections and couplings on ``E_tab``."""

    g: np.ndarray  # [n_g, 3] sample frame
    wm: np.ndarray  # [n_g]
    row: np.ndarray  # [n_g] int32
    E_tab: np.ndarray  # float32
    delta_tab: np.ndarray
    chi_re: np.ndarray
    chi_im: np.ndarray
    u_re: np.ndarray
    u_im: np.ndarray
    log_mu: np.ndarray  # [n_mu, n_tab]
    lo_keep: float
    hi_keep: float
    composition: list

    def as_dict(self):
        return dict(g=self.g, wm=self.wm, row=self.row, E_tab=self.E_tab, delta_tab=self.delta_tab,
                    chi_re=self.chi_re, chi_im=self.chi_im, u_re=self.u_re, u_im=self.u_im,
                    log_mu=self.log_mu, lo_keep=self.lo_keep, hi_keep=self.hi_keep)


def line_tables(crystal_key, hkl_list, E_grid_eV, composition=None, tab_ratio=1.0 + 2.0e-3) -> LineTables:
    """Reflection rows and FP32 coupling tables for a crystal (``cxr/_batched.py:230-330``).

    ``crystal_key`` names a catalog crystal or is a ``Crystal`` (its own
    ``surface_hkl`` then sets the sample frame)."""
    cr = CRYSTALS[crystal_key] if isinstance(crystal_key, str) else crystal_key
    comp = cr.composition if composition is None else list(composition)
    lo, hi = keep_window(np.asarray(E_grid_eV, dtype=float))
    E_tab = tabulation_grid(max(lo, 10.0), hi, tab_ratio)
    R = cr.orientation()
    g_rows = np.array([R @ cr.g_vector(hkl)[0] for hkl in hkl_list]).reshape(-1, 3)
    chi_re, chi_im, u_re, u_im = coupling_tables(cr, hkl_list, E_tab)
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)  # noqa: E731
    return LineTables(
        g=g_rows,
        wm=np.ones(len(hkl_list)),
        row=np.arange(len(hkl_list), dtype=np.int32),
        E_tab=f32(E_tab),
        delta_tab=f32(decrement(cr, E_tab)),
        chi_re=f32(chi_re), chi_im=f32(chi_im), u_re=f32(u_re), u_im=f32(u_im),
        log_mu=f32(attenuation.elemental_log_mu(comp, E_tab)),
        lo_keep=lo,
        hi_keep=hi,
        composition=comp,
    )


def synthetic_line_tables(n_g, E_grid_eV, rng, thickness_composition=(("Si", 0.04994),),
                          g_scale=(1.0, 4.0), F2_range=(0.1, 10.0), tab_ratio=1.0 + 2.0e-3) -> LineTables:
    """Random reflections with ``|F|^2`` uniform in ``F2_range`` (scoring-only workload).

    Each reflection gets a random reciprocal vector (|g| in ``g_scale``, polar
    angle within 60 deg of +z) and a flat coupling: ``chi = -r_e lambda^2
    |F|/(pi V)`` with a 20 A^3 cell, ``U/m`` from the same ``|F|``.
    """
    lo, hi = keep_window(np.asarray(E_grid_eV, dtype=float))
    E_tab = tabulation_grid(max(lo, 10.0), hi, tab_ratio)
    mag = rng.uniform(*g_scale, size=n_g)
    th = np.arccos(rng.uniform(0.5, 1.0, size=n_g))
    ph = rng.uniform(0.0, 2.0 * np.pi, size=n_g)
    g = mag[:, None] * np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], axis=1)
    F = np.sqrt(rng.uniform(*F2_range, size=n_g))
    lam = 12398.4198 / E_tab
    V = 20.0
    chi = -2.8179403e-5 * lam[None, :] ** 2 / (np.pi * V) * F[:, None]
    u = (4.0 * np.pi * (1.0 / 137.035999) * 1973.269804 / V / 510998.95) * (F / mag**2)[:, None] * np.ones_like(E_tab)[None, :]
    f32 = lambda a: np.ascontiguousarray(a, dtype=np.float32)  # noqa: E731
    comp = list(thickness_composition)
    delta = 2.7e-6 * (1000.0 / E_tab) ** 2
    return LineTables(
        g=g, wm=np.ones(n_g), row=np.arange(n_g, dtype=np.int32), E_tab=f32(E_tab),
        delta_tab=f32(delta), chi_re=f32(chi), chi_im=f32(0.01 * chi), u_re=f32(u), u_im=f32(0.0 * u),
        log_mu=f32(attenuation.elemental_log_mu(comp, E_tab)), lo_keep=lo, hi_keep=hi, composition=comp,
    )


class DeviceLineTables:
    def __init__(self, t: LineTables, device):
        import torch

        self.host = t
        dv = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a), device=device).to(dt)  # noqa: E731
        self.g = dv(t.g, torch.float64)
        self.wm = dv(t.wm, torch.float64)
        self.row = dv(t.row, torch.int32)
        self.E_tab = dv(t.E_tab, torch.float32)
        self.delta_tab = dv(t.delta_tab, torch.float32)
        self.chi_re = dv(t.chi_re, torch.float32)
        self.chi_im = dv(t.chi_im, torch.float32)
        self.u_re = dv(t.u_re, torch.float32)
        self.u_im = dv(t.u_im, torch.float32)
        self.log_mu = dv(t.log_mu, torch.float32)


def segments_fp32(segments, device=None):
    """FP32 device segment records ``(r_mid, v_hat, L, E_repr)`` from a transport result."""
    import torch

    def dev(x):
        if isinstance(x, torch.Tensor):
            return x
        return torch.as_tensor(np.asarray(x), device=device)

    E = segments.get("E_repr_keV")
    if E is None:
        E = segments["E_keV"]
    return {
        "r_mid": dev(segments["r_mid"]).to(torch.float32).contiguous(),
        "v_hat": dev(segments["v_hat"]).to(torch.float32).contiguous(),
        "L": dev(segments["L_ang"]).to(torch.float32).contiguous(),
        "E": dev(E).to(torch.float32).contiguous(),
    }


def _uniform_grid(E_grid):
    E = np.asarray(E_grid, dtype=float)
    if E.size < 2:
        return True, float(E[0]) if E.size else 0.0, 1.0
    dE = (E[-1] - E[0]) / (E.size - 1)
    ok = np.allclose(E, E[0] + dE * np.arange(E.size), rtol=0.0, atol=1e-12 * max(abs(E[-1]), 1.0))
    return bool(ok), float(E[0]), float(dE)


class LineScorer:
    """Reusable device scorer: tables and direction sets uploaded once."""

    def __init__(self, tables: LineTables, E_grid_eV, thickness_ang, n_hats, weights, device):
        import torch

        self.tables = tables
        self.dtab = DeviceLineTables(tables, device)
        self.device = device
        self.E_grid = np.ascontiguousarray(E_grid_eV, dtype=np.float64)
        self.E_grid_d = torch.as_tensor(self.E_grid, device=device)
        self.uniform, self.E0, self.dE = _uniform_grid(self.E_grid)
        self.thickness = float(thickness_ang)
        n_hats = np.asarray(n_hats, dtype=float).reshape(-1, 3)
        self.n_dir = n_hats.shape[0]
        es = np.empty((self.n_dir, len(tables.wm), 3))
        ep = np.empty_like(es)
        for d in range(self.n_dir):
            es[d], ep[d] = polarization_pairs(n_hats[d], tables.g)
        self.n_hats = torch.as_tensor(n_hats, device=device)
        self.weights = torch.as_tensor(np.asarray(weights, dtype=float).reshape(-1), device=device)
        self.e_s = torch.as_tensor(es, device=device)
        self.e_p = torch.as_tensor(ep, device=device)
        self.lib = _lib.load()
        self.scratch = _lib.DeviceScratch(device)  # caller-owned workspace of this scorer's calls
        self.record_events = False
        self.events = []  # ((start, end), cells) per reduction launch when recording (nothing else
        #                   is computed, so recording does not disturb the timed work)

    def line_args(self, seg, scale, j0=0, n=None, row_weight=None):
        """``hr_line_args`` for segments ``[j0, j0+n)`` of ``seg`` (outputs unset)."""
        n_total = int(seg["L"].shape[0])
        n = n_total - j0 if n is None else n
        t = self.dtab
        args = _lib.LineArgs()
        args.n_seg = n
        args.r_mid = seg["r_mid"].data_ptr() + 12 * j0
        args.v_hat = seg["v_hat"].data_ptr() + 12 * j0
        args.L = seg["L"].data_ptr() + 4 * j0
        args.E = seg["E"].data_ptr() + 4 * j0
        args.thickness = self.thickness
        args.n_dir = self.n_dir
        args.n_hat, args.dir_weight = self.n_hats.data_ptr(), self.weights.data_ptr()
        args.n_g = len(self.tables.wm)
        args.g, args.wm, args.row = t.g.data_ptr(), t.wm.data_ptr(), t.row.data_ptr()
        args.e_s, args.e_p = self.e_s.data_ptr(), self.e_p.data_ptr()
        args.n_tab = int(t.E_tab.shape[0])
        args.E_tab, args.delta_tab = t.E_tab.data_ptr(), t.delta_tab.data_ptr()
        args.chi_re, args.chi_im = t.chi_re.data_ptr(), t.chi_im.data_ptr()
        args.u_re, args.u_im = t.u_re.data_ptr(), t.u_im.data_ptr()
        args.n_mu = int(t.log_mu.shape[0])
        args.log_mu = t.log_mu.data_ptr()
        args.lo_keep, args.hi_keep = self.tables.lo_keep, self.tables.hi_keep
        args.scale = float(scale)
        args.row_weight = None if row_weight is None else row_weight.data_ptr() + 4 * j0
        return args

    def lines(self, seg, scale, j0=0, n=None, row_weight=None):
        """(E_r, a, w) device tensors for segments ``[j0, j0+n)`` of ``seg``.

        ``row_weight`` (device FP32 ``[n_seg]``) multiplies each segment's lines:
        the statistical weight of a split or rouletted particle.
        """
        import torch

        n_total = int(seg["L"].shape[0])
        n = n_total - j0 if n is None else n
        m = n * self.n_dir * len(self.tables.wm)
        E_r = torch.empty(m, dtype=torch.float64, device=self.device)
        a = torch.empty(m, dtype=torch.float64, device=self.device)
        w = torch.empty(m, dtype=torch.float32, device=self.device)
        args = self.line_args(seg, scale, j0, n, row_weight)
        args.E_r, args.a, args.w = E_r.data_ptr(), a.data_ptr(), w.data_ptr()
        _lib.check(self.lib.hr_cxr_lines(C.byref(args), _stream()), "hr_cxr_lines")
        return E_r, a, w

    def reduce(self, E_r, a, w, out):
        args = _lib.SincsqArgs()
        args.n_lines = int(E_r.shape[0])
        args.E_r, args.a, args.w = E_r.data_ptr(), a.data_ptr(), w.data_ptr()
        args.n_E = int(self.E_grid.size)
        args.E_grid = self.E_grid_d.data_ptr()
        args.uniform = int(self.uniform)
        args.E0, args.dE = self.E0, self.dE
        args.out = out.data_ptr()
        args.ws = self.scratch.get(_lib.workspace_bytes("hr_sincsq_workspace_bytes", C.byref(args)))
        ev = None
        if self.record_events:
            import torch

            ev = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
            ev[0].record()
        _lib.check(self.lib.hr_sincsq_reduce(C.byref(args), _stream()), "hr_sincsq_reduce")
        if ev is not None:
            ev[1].record()
            self.events.append((ev, args.n_lines * args.n_E))

    def spectrum(self, seg, Ne, out=None, row_weight=None):
        """Solid-angle-integrated per-electron spectrum [photons/eV/electron] (device FP64)."""
        import torch

        if out is None:
            out = torch.zeros(self.E_grid.size, dtype=torch.float64, device=self.device)
        n_total = int(seg["L"].shape[0])
        per_seg = self.n_dir * len(self.tables.wm)
        step = max(1, LINE_BLOCK // per_seg)
        for j0 in range(0, n_total, step):
            n = min(step, n_total - j0)
            E_r, a, w = self.lines(seg, 1.0 / float(Ne), j0, n, row_weight)
            self.reduce(E_r, a, w, out)
        return out

    def window_spectra(self, seg, Ne, out=None, x_c=100.0, row_weight=None):
        """Per-direction (pixel) spectra ``[n_dir, n_E]`` by the fused prologue +
        support-windowed route (``node_support_cutoff = x_c``,
        ``spectrum/cxr/_node_window.py``); the lines are never stored.  Row ``d``
        carries direction ``d``'s weight (its solid angle)."""
        import torch

        if not self.uniform:
            raise ValueError("the windowed route needs a uniform energy grid")
        if out is None:
            out = torch.zeros((self.n_dir, self.E_grid.size), dtype=torch.float64, device=self.device)
        largs = self.line_args(seg, 1.0 / float(Ne), 0, None, row_weight)
        wa = _lib.WindowArgs()
        wa.lines = C.pointer(largs)
        wa.n_E, wa.E0, wa.dE, wa.x_c = int(self.E_grid.size), self.E0, self.dE, float(x_c)
        wa.out = out.data_ptr()
        wa.ws = self.scratch.get(_lib.workspace_bytes("hr_window_spectra_workspace_bytes", C.byref(wa)))
        _lib.check(self.lib.hr_window_spectra(C.byref(wa), _stream()), "hr_window_spectra")
        return out


def mc_cxr_spectrum_solid_angle(segments, E_grid_eV, crystal, hkl_list, *, n_hats, weights,
                                composition=None, electron_limit=None):
    """``sum_i weights_i * mc_cxr_spectrum(n_hat=n_hats_i)`` [photons/eV/electron], host FP64."""
    from ..backend import require_device

    dev = require_device().torch_device
    tables = line_tables(crystal, hkl_list, E_grid_eV, composition)
    scorer = LineScorer(tables, E_grid_eV, segments["thickness_ang"], n_hats, weights, dev)
    seg = segments_fp32(segments, dev)
    Ne = segments["Ne"] if electron_limit is None else electron_limit
    return scorer.spectrum(seg, Ne).cpu().numpy()


class BremScorer:
    """BH+Elwert track-length reducer for one composition and one direction."""

    def __init__(self, composition, E_grid_eV, thickness_ang, n_hat, device):
        import torch

        if not 1 <= len(composition) <= 4:
            raise ValueError("hr_brem_reduce scores 1..4 elements")
        self.composition = list(composition)
        self.E_grid = np.ascontiguousarray(E_grid_eV, dtype=np.float64)
        self.E_grid_d = torch.as_tensor(self.E_grid, device=device)
        mu = attenuation.mu_total_inv_ang(self.composition, self.E_grid)
        self.mu_d = torch.as_tensor(np.ascontiguousarray(mu, dtype=np.float32), device=device)
        self.thickness = float(thickness_ang)
        self.n_hat = np.asarray(n_hat, dtype=float)
        self.device = device
        self.lib = _lib.load()
        self.scratch = _lib.DeviceScratch(device)

    def reduce(self, seg, out, row_weight=None):
        args = _lib.BremArgs()
        args.n_seg = int(seg["L"].shape[0])
        args.r_mid, args.v_hat = seg["r_mid"].data_ptr(), seg["v_hat"].data_ptr()
        args.L, args.E_keV = seg["L"].data_ptr(), seg["E"].data_ptr()
        args.thickness = self.thickness
        args.n_hat = (C.c_double * 3)(*self.n_hat.tolist())
        args.n_el = len(self.composition)
        for i, (el, n) in enumerate(self.composition):
            args.Z[i] = float(TRANSPORT_ELEMENTS[el]["Z"])
            args.n_cm3[i] = float(n) * 1e24
        args.n_E = int(self.E_grid.size)
        args.E_grid_eV = self.E_grid_d.data_ptr()
        args.mu = self.mu_d.data_ptr()
        args.row_weight = None if row_weight is None else row_weight.data_ptr()
        args.out = out.data_ptr()
        args.ws = self.scratch.get(_lib.workspace_bytes("hr_brem_workspace_bytes", C.byref(args)))
        _lib.check(self.lib.hr_brem_reduce(C.byref(args), _stream()), "hr_brem_reduce")

    def spectrum(self, seg, Ne, out=None, row_weight=None):
        """Density [photons/eV/sr/electron] (device FP64): the sum over 4 pi Ne."""
        import torch

        acc = torch.zeros(self.E_grid.size, dtype=torch.float64, device=self.device)
        self.reduce(seg, acc, row_weight)
        acc /= 4.0 * np.pi * float(Ne)
        if out is not None:
            out += acc
            return out
        return acc


def mc_brem_spectrum(segments, E_grid_eV, composition, n_hat, electron_limit=None):
    """BH+Elwert density per eV per sr per incident electron (host FP64)."""
    from ..backend import require_device

    dev = require_device().torch_device
    scorer = BremScorer(composition, E_grid_eV, segments["thickness_ang"], n_hat, dev)
    Ne = segments["Ne"] if electron_limit is None else electron_limit
    return scorer.spectrum(segments_fp32(segments, dev), Ne).cpu().numpy()
