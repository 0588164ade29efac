"""Coherent PXR/CBS line spectrum on the hip backend.

The reference's ``coherent=True`` route of ``mc_cxr_spectrum``
(``montecarlo/spectrum/cxr/_spectrum.py:163-164``: "Sum segment fields
coherently instead of segment intensities"), streamed the way its coherent
stream pipeline is (``coherent_stream_jit_kernel.py:1-22``): every batch of
transport rows becomes per-line complex records (``hr_coherent_lines``), which
are summed into persistent FP64 field planes, one per (direction, reflection)
row (``hr_coherent_field``); the planes are squared only when the spectrum is
read (``hr_coherent_finalize``), so the fields of all batches add before the
square.  Reflections, mosaic orientations and detector directions stay
incoherent with one another.

With ``grouped=True`` the scorer also keeps the inter-electron decoherence
blend's grouped term ``G = sum_e |sum_{j in e} field_j|^2`` per row
(``hr_coherent_grouped``; coherent_grouped_jit_kernel.py:62-230), and
``spectrum(factor=F)`` returns the reference's row estimator
(``mixed_row_power``, coherent_population.py:83-96) from the caller's
inter-electron factor ``F(E)``: ``(1 - F) G + F P`` for a sampled bunch,
``G + (N-1)/(M-1) F (P - G)`` for a physical population of ``N`` electrons
(``pair_scale``).  The Gaussian longitudinal factor ``F_z`` is
``gaussian_longitudinal_factor``; the reference's transverse and empirical
(sampled-offset) factors, ``phase_retention``, layered absorbers and the
coherent ``sinc_cutoff`` window are not built.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from .. import _lib
from .crystal import CRYSTALS, decrement
from .lines import LineScorer, LineTables, line_tables

HBARC = 1973.269804  # eV Angstrom (materials/crystal.py:48)
C_ANG_PER_FS = 2997.92458  # speed of light [Angstrom/fs]
ELEMENTARY_CHARGE_C = 1.602176634e-19
N_FIELDS = 13
# Record bytes per device pass (13 FP64 planes per line and row).
RECORD_BYTES = 1 << 30


def _stream():
    import torch

    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def escape_phase_rate(crystal_key, E_grid_eV):
    """``delta(E) E / hbar c`` [1/A] on the output grid: the in-medium
    propagation phase per unit escape path (``cxr/_setup.py:1030-1045``)."""
    E = np.asarray(E_grid_eV, dtype=np.float64)
    return decrement(CRYSTALS[crystal_key], E) * E / HBARC


def gaussian_longitudinal_factor(E_grid_eV, longitudinal_rms_fs):
    """``F_z(E) = exp(-(omega sigma_z)^2)``, ``omega = E / hbar c``,
    ``sigma_z = c * longitudinal_rms_fs`` (coherent_windows.py:114-119;
    coherent_form_factor.py:22): one at ``E = 0`` or ``sigma_z = 0``."""
    rms = float(longitudinal_rms_fs)
    if not np.isfinite(rms) or rms < 0.0:
        raise ValueError("longitudinal_rms_fs must be finite and nonnegative")
    omega = np.asarray(E_grid_eV, dtype=np.float64) / HBARC
    return np.exp(-((omega * (rms * C_ANG_PER_FS)) ** 2))


def physical_bunch_electrons(bunch_charge_pc=1.0):
    """``N = Q / e`` of a bunch charge in pC (coherent_population.py:37-50)."""
    charge = float(bunch_charge_pc)
    if not np.isfinite(charge) or charge < 0.0:
        raise ValueError("bunch_charge_pc must be finite and nonnegative")
    population = charge * 1e-12 / ELEMENTARY_CHARGE_C
    if not np.isfinite(population) or 0.0 < population < 1.0:
        raise ValueError("a nonzero coherent bunch must contain at least one physical electron")
    return population


def pair_scale(physical_electrons, incident_samples):
    """Cross-pair weight ``(N-1)/(M-1)`` of ``N`` physical electrons sampled by
    ``M`` incident histories (coherent_population.py:53-80): zero for ``N <= 1``."""
    population = float(physical_electrons)
    if not np.isfinite(population) or population < 0 or 0 < population < 1:
        raise ValueError("physical_electrons must be zero or finite and at least one")
    if incident_samples < 1:
        raise ValueError("coherent emission needs a positive incident sample count")
    if population <= 1:
        return 0.0
    if incident_samples < 2:
        raise ValueError("physical coherent cross-electron power requires at least two incident samples")
    return (population - 1.0) / (incident_samples - 1)


class CoherentScorer:
    """Persistent coherent field planes of one case, fed batch by batch."""

    def __init__(self, tables: LineTables, crystal_key, E_grid_eV, thickness_ang, n_hats, weights, device,
                 grouped=False):
        import torch

        self.lines = LineScorer(tables, E_grid_eV, thickness_ang, n_hats, weights, device)
        self.device = device
        self.n_dir = self.lines.n_dir
        self.n_g = len(tables.wm)
        self.n_rows = self.n_dir * self.n_g
        self.n_E = int(np.asarray(E_grid_eV).size)
        self.dom = torch.as_tensor(escape_phase_rate(crystal_key, E_grid_eV), device=device)
        self.field = torch.zeros((self.n_rows, 4, self.n_E), dtype=torch.float64, device=device)
        self.grouped = (torch.zeros((self.n_rows, self.n_E), dtype=torch.float64, device=device)
                        if grouped else None)
        self.n_electrons = 0
        # crystal, tables and detector rays of the incoherent prologue's args; the
        # coherent kernels read the FP64 slot rows instead of its FP32 segments
        self._largs = self.lines.line_args({"L": torch.empty(0, device=device),
                                            "r_mid": torch.empty((0, 3), device=device),
                                            "v_hat": torch.empty((0, 3), device=device),
                                            "E": torch.empty(0, device=device)}, 1.0)
        self._largs.n_seg = 0
        self._largs.r_mid = self._largs.v_hat = self._largs.L = self._largs.E = None
        self._rec = None
        self.lib = _lib.load()
        self.scratch = _lib.DeviceScratch(device)  # field partials of this scorer's calls
        self.record_events = False
        self.events = []  # ((start, end), cells) per field launch when recording

    def reset(self):
        self.field.zero_()
        if self.grouped is not None:
            self.grouped.zero_()
        self.n_electrons = 0

    def _records(self, n_lines):
        import torch

        need = N_FIELDS * self.n_rows * max(int(n_lines), 1)
        if self._rec is None or self._rec.numel() < need:
            self._rec = torch.empty(need, dtype=torch.float64, device=self.device)
        return self._rec

    def records(self, rows, seg_count, n, slots):
        """Record planes ``[13, n_rows, n_lines]`` of slot rows (for inspection)."""
        import torch

        cnt = seg_count[:n].clamp(max=slots)
        off = torch.cumsum(cnt, 0) - cnt
        n_lines = int(cnt.sum().item()) if n else 0
        rec = self._records(n_lines)[: N_FIELDS * self.n_rows * max(n_lines, 1)]
        self._launch_lines(rows, seg_count, off, 0, n, slots, rec, n_lines)
        return rec.view(N_FIELDS, self.n_rows, max(n_lines, 1))[:, :, :n_lines]

    def _launch_lines(self, rows, seg_count, row_offset, i0, m, slots, rec, ld):
        a = _lib.CoherentArgs()
        a.lines = C.pointer(self._largs)
        a.n, a.rows = int(m), int(slots)
        a.seg_count = seg_count.data_ptr() + 8 * i0
        a.row_offset = row_offset.data_ptr()  # indexed from particle i0
        base = i0 * slots
        a.seg_mid = rows["seg_mid"].data_ptr() + 24 * base
        a.seg_dir = rows["seg_dir"].data_ptr() + 24 * base
        for k in ("seg_len", "seg_E", "seg_E_end", "seg_t0"):
            setattr(a, k, rows[k].data_ptr() + 8 * base)
        a.first, a.ld = 0, int(ld)
        a.n_lines = int(ld)  # the pass writes columns [0, ld): one thread per row
        a.rec = rec.data_ptr()
        _lib.check(self.lib.hr_coherent_lines(C.byref(a), _stream()), "hr_coherent_lines")

    def _launch_field(self, rec, n_lines):
        f = _lib.CoherentFieldArgs()
        f.rec, f.ld, f.n_lines, f.n_rows = rec.data_ptr(), int(n_lines), int(n_lines), self.n_rows
        f.n_E, f.E_grid, f.dom = self.n_E, self.lines.E_grid_d.data_ptr(), self.dom.data_ptr()
        f.field = self.field.data_ptr()
        f.ws = self.scratch.get(_lib.workspace_bytes("hr_coherent_field_workspace_bytes", C.byref(f)))
        ev = None
        if self.record_events:
            import torch

            ev = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
            ev[0].record()
        _lib.check(self.lib.hr_coherent_field(C.byref(f), _stream()), "hr_coherent_field")
        if ev is not None:
            ev[1].record()
            self.events.append((ev, n_lines * self.n_rows * self.n_E))

    def _launch_grouped(self, rec, n_lines, group_start):
        g = _lib.CoherentGroupedArgs()
        f = g.field
        f.rec, f.ld, f.n_lines, f.n_rows = rec.data_ptr(), int(n_lines), int(n_lines), self.n_rows
        f.n_E, f.E_grid, f.dom = self.n_E, self.lines.E_grid_d.data_ptr(), self.dom.data_ptr()
        g.group_start, g.n_groups = group_start.data_ptr(), int(group_start.numel()) - 1
        g.grouped = self.grouped.data_ptr()
        f.ws = self.scratch.get(_lib.workspace_bytes("hr_coherent_grouped_workspace_bytes", C.byref(g)))
        _lib.check(self.lib.hr_coherent_grouped(C.byref(g), _stream()), "hr_coherent_grouped")

    def add_rows(self, rows, seg_count, n, slots, live_counts=None, electron=None):
        """Add the fields of ``n`` particles' slot rows (``slots`` rows each,
        ``seg_count`` written) to the planes, in passes of at most
        ``RECORD_BYTES`` of records.  With a ``live_counts`` list the records
        are only built, and each pass's live (line, row) count is appended
        instead of adding fields.  A grouped scorer squares each electron's
        field: particle ``i`` belongs to electron ``electron[i]`` (host array,
        equal ids adjacent; default: each particle its own electron), and
        passes end only between electrons."""
        import torch

        if n == 0:
            return
        cnt = seg_count[:n].clamp(max=slots)
        incl = torch.cumsum(cnt, 0)
        per_line = N_FIELDS * self.n_rows * 8
        max_lines = max(1, RECORD_BYTES // per_line)
        incl_h = incl.cpu().numpy()
        key = None
        if self.grouped is not None and electron is not None:
            key = np.asarray(electron, dtype=np.int64)[:n]
            if key.size != n:
                raise ValueError(f"electron ids: {key.size} for {n} particles")
        i0 = 0
        while i0 < n:
            base = int(incl_h[i0 - 1]) if i0 else 0
            i1 = int(np.searchsorted(incl_h, base + max_lines, side="right"))
            i1 = min(max(i1, i0 + 1), n)
            if key is not None and i1 < n and key[i1] == key[i1 - 1]:
                j = i1
                while j > i0 and key[j - 1] == key[i1]:
                    j -= 1
                if j == i0:
                    raise ValueError(f"electron {int(key[i1])} has more lines than one record pass holds")
                i1 = j
            n_lines = int(incl_h[i1 - 1]) - base
            if n_lines > 0:
                off = incl[i0:i1] - cnt[i0:i1] - base
                rec = self._records(n_lines)
                self._launch_lines(rows, seg_count, off, i0, i1 - i0, slots, rec, n_lines)
                if live_counts is None:
                    self._launch_field(rec, n_lines)
                    if self.grouped is not None:
                        if key is None:
                            starts = off
                        else:
                            k = key[i0:i1]
                            first = np.flatnonzero(np.r_[True, k[1:] != k[:-1]])
                            starts = off[torch.as_tensor(first, device=off.device)]
                        group_start = torch.cat((starts.to(torch.int64),
                                                 torch.tensor([n_lines], dtype=torch.int64, device=off.device)))
                        self._launch_grouped(rec, n_lines, group_start)
                else:
                    coeff = rec[: N_FIELDS * self.n_rows * n_lines].view(N_FIELDS, self.n_rows, n_lines)[4:8]
                    live_counts.append(int((coeff != 0).any(0).sum().item()))
                del off
            i0 = i1

    def add_transport(self, res, n_electrons):
        """Fields of a ``compact=False`` transport result's slot rows."""
        self.add_slots(res.slots, n_electrons)

    def add_slots(self, slots, n_electrons, live_counts=None):
        """Fields of slot rows ``[(e0, m, cap, rows, seg_count)]``; ``n_electrons``
        histories join the count the spectrum is divided by."""
        for _e0, m, cap, rows, seg_count in slots:
            self.add_rows(rows, seg_count, m, cap, live_counts=live_counts)
        self.n_electrons += int(n_electrons)

    def spectrum(self, Ne=None, out=None, factor=None, pair_scale=None):
        """``sum_d w_d sum_g wm_g (|F_s|^2 + |F_p|^2) / Ne`` [photons/eV/electron]
        (device FP64).  With ``factor`` (``F(E)`` on the grid; grouped scorer)
        each row's power is the blend: ``(1 - F) G + F P``, or with a
        ``pair_scale`` (``(N-1)/(M-1)``) ``G + pair_scale F (P - G)``."""
        import torch

        Ne = self.n_electrons if Ne is None else Ne
        if out is None:
            out = torch.zeros(self.n_E, dtype=torch.float64, device=self.device)
        if Ne <= 0:
            return out
        t = self.lines.dtab
        if factor is not None:
            if self.grouped is None:
                raise ValueError("the decoherence blend needs a scorer built with grouped=True")
            F = torch.as_tensor(np.ascontiguousarray(factor, dtype=np.float64), device=self.device)
            if F.numel() != self.n_E:
                raise ValueError(f"factor has {F.numel()} values for {self.n_E} bins")
            pair = -1.0 if pair_scale is None else float(pair_scale)
            if pair_scale is not None and not (np.isfinite(pair) and pair >= 0.0):
                raise ValueError("pair_scale must be finite and nonnegative")
            _lib.check(self.lib.hr_coherent_blend_finalize(
                self.field.data_ptr(), self.grouped.data_ptr(), self.n_dir, self.n_g,
                self.lines.weights.data_ptr(), t.wm.data_ptr(), self.n_E, F.data_ptr(), pair,
                1.0 / float(Ne), out.data_ptr(), _stream()), "hr_coherent_blend_finalize")
            return out
        _lib.check(self.lib.hr_coherent_finalize(self.field.data_ptr(), self.n_dir, self.n_g,
                                                 self.lines.weights.data_ptr(), t.wm.data_ptr(), self.n_E,
                                                 1.0 / float(Ne), out.data_ptr(), _stream()),
                   "hr_coherent_finalize")
        return out


def segment_rows(segments, device):
    """Slot rows (one row per segment) of a reference-style segments dict:
    r_mid, v_hat, L_ang, E_repr_keV (or E_keV) and t_ang (segment start age)."""
    import torch

    def f64(a):
        return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=device)

    E = segments.get("E_repr_keV", segments["E_keV"])
    n = int(np.asarray(segments["L_ang"]).shape[0])
    rows = {"seg_mid": f64(segments["r_mid"]).reshape(-1, 3), "seg_dir": f64(segments["v_hat"]).reshape(-1, 3),
            "seg_len": f64(segments["L_ang"]), "seg_E": f64(E), "seg_E_end": f64(E),
            "seg_t0": f64(segments.get("t_ang", np.zeros(n)))}
    return rows, torch.ones(n, dtype=torch.int64, device=device), n


def mc_cxr_coherent_spectrum(segments, E_grid_eV, crystal, hkl_list, *, n_hats, weights, composition=None,
                             electron_limit=None):
    """Coherent ``sum_i weights_i * mc_cxr_spectrum(n_hat=n_hats_i, coherent=True)``
    [photons/eV/electron] of a segments dict (host FP64 result)."""
    from ..backend import require_device

    dev = require_device().torch_device
    tables = line_tables(crystal, hkl_list, E_grid_eV, composition)
    sc = CoherentScorer(tables, crystal, E_grid_eV, segments["thickness_ang"], n_hats, weights, dev)
    rows, cnt, n = segment_rows(segments, dev)
    sc.add_rows(rows, cnt, n, 1)
    Ne = segments["Ne"] if electron_limit is None else electron_limit
    return sc.spectrum(Ne).cpu().numpy()
