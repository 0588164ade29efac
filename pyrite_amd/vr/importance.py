"""Learned-importance splitting and roulette for one beam/target case.

The importance network (``ImportanceNet``, kernels in ``csrc/mlp.hip``)
predicts, from an electron state ``x = (E/E0, u, z/thickness, tilt)``, the
log of the photon yield the next ``horizon_rows`` transport rows of that
electron will add to the PXR-peak window of the detector tally.  It is
trained on the device on labels produced by the analog pipeline itself:
every analog row's lines are reduced to their expected window content
(``hr_line_mass``) and summed over the row and its successors in the same
history.

Production runs transport in rounds (``hr_vr_round``: at most
``rows_per_round`` rows per particle near the split threshold, ``rows_far``
far below it, then suspend).  Between rounds the
network scores every suspended particle and a weight window
(``hr_vr_window``) sets the target weight ``w_t = clamp(I_ref / I(x),
1/k_max, 1)``: particles heavier than ``w_t`` are split into copies of weight
``w/n``, lighter ones play Russian roulette for weight ``w_t``
(``hr_vr_spawn`` writes the next bank with fresh child streams).  Every row
is scored with its particle's weight, so the tallies keep the analog
expectation whatever the network predicts; the analog pipeline
(``CasePipeline``) stays the reference for bias checks and the figure of
merit ``1 / (rel_var * time)`` on the peak window.

There is no counterpart in the reference: its variance reduction is
likelihood-ratio cone splitting on the CPU core (``src/pyrite/_cone_splitting.py:57``,
``transport/cone_split.py:223``) and line roulette (``src/pyrite/_line_roulette.py:29``).
"""

from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass

import numpy as np

from .. import _lib
from ..spectrum.lines import LINE_BLOCK, segments_fp32
from ..transport.core import alloc_rows, transport_args
from .mlp import ImportanceNet, init_params

EXIT_SUSPENDED = 6


def _stream():
    import torch

    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


@dataclass
class VRConfig:
    rows_per_round: int = 1  # rows between decisions near the split threshold
    rows_far: int = 8  # rows between decisions far below it (w I/I_ref < far_ratio)
    far_ratio: float = 1.0 / 32.0
    horizon_rows: int = 2
    # "second": the label is the square root of the rows' second moment over
    # the bin offset (narrow lines that hit a bin centre deposit their peak
    # weight, so a row's variance grows as w^2/(a dE) while its mean grows as
    # w/(a dE); splitting in proportion to the root of the second moment per
    # unit cost minimises variance x cost).  "mean": their expected
    # peak-window content.
    label: str = "second"
    k_max: float = 16.0  # smallest target weight is 1/k_max
    # weight-window band around the target weight.  Rouletting copies back up
    # to a larger weight multiplies the rare grazing / detector-bound rows they
    # may produce later (the heavy tail of the peak tally), so by default the
    # window only splits: ratio_lo is tiny and copies keep their weight.
    ratio_lo: float = 1.0e-6
    ratio_hi: float = 2.0
    tail_fraction: float = 0.01  # below this share of the batch alive, rounds take tail_rows rows
    tail_rows: int = 32
    n_max: int = 64  # copies per split
    ref_quantile: float = 0.98  # states at this predicted quantile get w_t = 1/k_max
    peak_half_width_eV: float = 400.0
    roulette_keep_mass: float = 0.99  # rows at or above the mass carrying this fraction are scored surely
    roulette_bound: bool = True  # roulette rows on expected mass + peak line weight (a content bound)
    label_eps_frac: float = 1.0e-3  # floor of the regression target, relative to the mean label
    train_rows: int = 2_000_000
    epochs: int = 6
    batch: int = 4096
    lr: float = 2.0e-3
    seed: int = 7


def line_second_moment(w, a_dE):
    """Offset-averaged second moment of the content a line of peak ``w`` and
    width ``a_dE = a dE`` (in bins) deposits on a uniform grid: the larger of
    the narrow-line limit ``w^2 (2 pi/3)/(a dE)`` and the wide-line limit
    ``(w pi/(a dE))^2``.  Works on numpy arrays and torch tensors."""
    return w * w * np.maximum((np.pi / a_dE) ** 2, (2.0 * np.pi / 3.0) / a_dE) if isinstance(a_dE, np.ndarray) \
        else w * w * ((np.pi / a_dE) ** 2).maximum((2.0 * np.pi / 3.0) / a_dE)


class ParticleBank:
    """Structure-of-arrays particle states on the device."""

    def __init__(self, capacity: int, device, E_cut_keV: float):
        import torch

        m = int(capacity)
        self.capacity = m
        self.pos = torch.zeros((m, 3), dtype=torch.float64, device=device)
        self.dirs = torch.zeros((m, 3), dtype=torch.float64, device=device)
        self.E = torch.zeros(m, dtype=torch.float64, device=device)
        self.clock = torch.zeros(m, dtype=torch.float64, device=device)
        self.weight = torch.zeros(m, dtype=torch.float64, device=device)  # line tally weight
        self.weight2 = torch.zeros(m, dtype=torch.float64, device=device)  # brem tally weight
        self.key = torch.zeros(m, dtype=torch.uint64, device=device)
        self.prim = torch.zeros(m, dtype=torch.int64, device=device)
        self.E_cut = torch.full((m,), float(E_cut_keV), dtype=torch.float64, device=device)
        self.stragg = torch.zeros(m, dtype=torch.float64, device=device)
        self.rows = torch.ones(m, dtype=torch.int32, device=device)  # next round length


class LearnedImportance:
    """Training and production of the learned variance reduction for a ``CasePipeline``."""

    def __init__(self, pipe, vcfg: VRConfig | None = None, net: ImportanceNet | None = None):
        self.pipe = pipe
        self.vcfg = vcfg or VRConfig()
        self.lib = _lib.load()
        cfg = pipe.cfg
        if not pipe.lines.uniform:
            raise ValueError("the learned importance needs a uniform energy grid")
        self.dE = float(pipe.lines.dE)
        self.E_peak = self._peak_energy()
        h = self.vcfg.peak_half_width_eV
        self.window = (self.E_peak - h, self.E_peak + h)
        grid = pipe.E_grid
        self.window_bins = (grid >= self.window[0]) & (grid < self.window[1])
        self.net = net or ImportanceNet(init_params(np.random.default_rng(self.vcfg.seed)), pipe.dev)
        self.log_I_ref = None
        self.m_thr_unit = None  # line-roulette threshold for a line scale of 1 (per-electron masses x Ne)
        self.train_history = []
        self._banks = None
        self.overlap = True  # score each round on a second stream beside the next round's transport
        self._score_stream = None
        self._rows_R = 0
        self.rounds = 0

    # -- case geometry ------------------------------------------------------
    def _peak_energy(self):
        """First-order PXR energy of the first reflection toward the cone axis
        for the incident direction: hbar c v.g / (1 - v.n)."""
        p = self.pipe
        g = np.asarray(p.line_tables.g[0], dtype=float)
        beta = np.sqrt(1.0 - 1.0 / (1.0 + p.cfg.E0_keV / 510.99895) ** 2)
        v = beta * np.asarray(p.beam_dir, dtype=float)
        n = np.asarray(p.n_hat, dtype=float)
        return float(1973.269804 * abs(v @ g) / (1.0 - v @ n))

    def peak_tally(self, spec_lines):
        """PXR/CBS line photons per incident electron in the peak window (the
        brem continuum under the peak is a separate tally)."""
        return float(spec_lines[self.window_bins].sum() * self.dE)

    # -- labels -------------------------------------------------------------
    def _features(self, n, pos, dirs, E, seg_len, out):
        cfg = self.pipe.cfg
        _lib.call("hr_vr_features", int(n), pos.data_ptr(), dirs.data_ptr(), E.data_ptr(),
                  None if seg_len is None else seg_len.data_ptr(), float(cfg.E0_keV), float(cfg.thickness_ang),
                  float(cfg.tilt_deg), out.data_ptr(), _stream())

    def row_masses(self, seg, Ne, window=True):
        """Expected peak-window content of every row's lines (device FP64 ``[n_rows]``);
        ``window=False``: their whole-grid content."""
        import torch

        lines = self.pipe.lines
        n_total = int(seg["L"].shape[0])
        n_dir, n_g = lines.n_dir, len(lines.tables.wm)
        mass = torch.empty(n_total, dtype=torch.float64, device=self.pipe.dev)
        step = max(1, LINE_BLOCK // (n_dir * n_g))
        for j0 in range(0, n_total, step):
            n = min(step, n_total - j0)
            E_r, a, w = lines.lines(seg, 1.0 / float(Ne), j0, n)
            self._mass(E_r, a, w, n, mass[j0:], window=window)
        return mass

    def row_second_moments(self, seg, Ne):
        """Second moment over the bin offset of every row's peak-window line
        content (device FP64 ``[n_rows]``, cross terms between lines left out).

        A line of peak ``w`` and width parameter ``a`` on bins of width ``dE``
        deposits ``X = w sum_k sinc^2(a (E_k - E_r))``.  With ``E_r`` uniform
        over a bin, ``E[X] = w pi/(a dE)`` and, for a line narrower than a bin,
        ``E[X^2] = w^2 (2 pi/3)/(a dE)`` (the integral of sinc^4 is 2 pi/3);
        a line wider than a few bins deposits its mean, ``E[X^2] = E[X]^2``.
        The larger of the two is taken."""
        import torch

        lines = self.pipe.lines
        n_total = int(seg["L"].shape[0])
        n_dir, n_g = lines.n_dir, len(lines.tables.wm)
        lo, hi = self.window
        out = torch.empty(n_total, dtype=torch.float64, device=self.pipe.dev)
        step = max(1, LINE_BLOCK // (n_dir * n_g))
        for j0 in range(0, n_total, step):
            n = min(step, n_total - j0)
            E_r, a, w = lines.lines(seg, 1.0 / float(Ne), j0, n)
            E_r = E_r[: n_dir * n * n_g].view(n_dir, n, n_g)
            a = a[: n_dir * n * n_g].view(n_dir, n, n_g)
            w = w[: n_dir * n * n_g].view(n_dir, n, n_g).double()
            live = (w != 0) & (a > 0) & (E_r >= lo) & (E_r < hi)
            m2 = line_second_moment(w, torch.where(live, torch.clamp(a * float(self.dE), min=1.0e-6), torch.ones_like(a)))
            out[j0:j0 + n] = torch.where(live, m2, torch.zeros_like(m2)).sum(dim=(0, 2))
        return out

    def _mass(self, E_r, a, w, n, out, window=True):
        lines = self.pipe.lines
        lo, hi = self.window if window else (-1.0e300, 1.0e300)
        _lib.call("hr_line_mass", E_r.data_ptr(), a.data_ptr(), w.data_ptr(), n, lines.n_dir,
                  len(lines.tables.wm), lo, hi, self.dE, out.data_ptr(), _stream())
        if not window and self.vcfg.roulette_bound:
            # A line's content on the bins is at most w (its peak) plus about its
            # expected mass: rouletting on that bound keeps every survivor's
            # contribution near the threshold, also for narrow lines that hit a bin.
            out[:n] += w.view(lines.n_dir, n, len(lines.tables.wm)).double().sum(dim=(0, 2))

    def score_lines(self, seg, Ne, out, row_weight=None, roulette_key=None):
        """Dense line spectrum of ``seg`` into ``out``; with ``roulette_key`` the
        rows below the mass threshold play line roulette first."""
        import torch

        lines = self.pipe.lines
        n_total = int(seg["L"].shape[0])
        n_dir, n_g = lines.n_dir, len(lines.tables.wm)
        step = max(1, LINE_BLOCK // (n_dir * n_g))
        mass = None
        for j0 in range(0, n_total, step):
            n = min(step, n_total - j0)
            E_r, a, w = lines.lines(seg, 1.0 / float(Ne), j0, n, row_weight)
            if roulette_key is not None:
                if mass is None:
                    mass = torch.empty(min(step, n_total), dtype=torch.float64, device=self.pipe.dev)
                self._mass(E_r, a, w, n, mass, window=False)
                _lib.call("hr_line_roulette", w.data_ptr(), n, n_dir, n_g, mass.data_ptr(),
                          float(self.m_thr_unit) / float(Ne), int(roulette_key), int(self._row0 + j0), _stream())
            lines.reduce(E_r, a, w, out)
        self._row0 += n_total

    def labels_from_rows(self, rows, Ne):
        """``(x [n,6] FP32, y [n] FP64 horizon yield)`` of compacted analog rows."""
        import torch

        dev = self.pipe.dev
        seg = segments_fp32({
            "r_mid": rows["seg_mid"], "v_hat": rows["seg_dir"], "L_ang": rows["seg_len"],
            "E_keV": rows["seg_E"], "E_repr_keV": 0.5 * (rows["seg_E"] + rows["seg_E_end"]),
        }, dev)
        n = int(seg["L"].shape[0])
        mass = self.row_masses(seg, Ne) if self.vcfg.label == "mean" else self.row_second_moments(seg, Ne)
        self._last_full_mass = self.row_masses(seg, Ne, window=False) * float(Ne)
        x = torch.empty((n, 6), dtype=torch.float32, device=dev)
        self._features(n, rows["seg_mid"], rows["seg_dir"], rows["seg_E"], rows["seg_len"], x)
        # horizon sums within each history (rows are electron-major / step-minor)
        ids = rows["seg_id"]
        _, counts = torch.unique_consecutive(ids, return_counts=True)
        starts = torch.cumsum(counts, 0) - counts
        end = torch.repeat_interleave(starts + counts, counts)
        j = torch.arange(n, device=dev)
        hi = torch.minimum(j + int(self.vcfg.horizon_rows), end)
        csum = torch.zeros(n + 1, dtype=torch.float64, device=dev)
        csum[1:] = torch.cumsum(mass, 0)
        y = csum[hi] - csum[j]
        if self.vcfg.label != "mean":
            # the difference of two running sums can fall a rounding below zero
            y = torch.sqrt(torch.clamp(y, min=0.0))
        return x, y

    def training_set(self, n_batches, first_batch=10_000):
        """Labels from ``n_batches`` analog batches (global batch ids from ``first_batch``,
        disjoint from production), subsampled to at most ``train_rows`` rows."""
        import torch

        p = self.pipe
        g = torch.Generator(device=p.dev)
        g.manual_seed(int(self.vcfg.seed))
        xs, ys, ms = [], [], []
        per_batch = max(1, int(self.vcfg.train_rows) // max(1, int(n_batches)))
        for b in range(int(n_batches)):
            res = p.transport(first_batch + b)
            x, y = self.labels_from_rows(res.rows, p.cfg.Ne)
            ms.append(self._last_full_mass)
            if x.shape[0] > per_batch:
                idx = torch.randperm(x.shape[0], generator=g, device=p.dev)[:per_batch]
                x, y = x[idx], y[idx]
            xs.append(x)
            ys.append(y)
        self.calibrate_roulette(torch.cat(ms))
        return torch.cat(xs), torch.cat(ys)

    def calibrate_roulette(self, full_mass):
        """Threshold of the line roulette: rows at or above it carry
        ``roulette_keep_mass`` of the total line mass (per-electron units x Ne)."""
        import torch

        m = torch.sort(full_mass[full_mass > 0], descending=True).values
        if m.numel() == 0:
            self.m_thr_unit = 1.0
            return
        c = torch.cumsum(m, 0)
        k = int(torch.searchsorted(c, float(self.vcfg.roulette_keep_mass) * float(c[-1])))
        self.m_thr_unit = float(m[min(k, m.numel() - 1)])

    def fit(self, x, y):
        """Train the network on ``log(eps + y)`` and set the window reference."""
        import torch

        pos = y[y > 0]
        eps = float(self.vcfg.label_eps_frac) * (float(pos.mean()) if pos.numel() else 1.0)
        t = torch.log(y + eps).to(torch.float32)
        mu, sd = float(t.mean()), float(t.std()) if t.numel() > 1 else 1.0
        # train on the standardised target; fold the affine map into the output layer
        ts = ((t - mu) / max(sd, 1e-6)).contiguous()
        hist = self.net.fit(x.contiguous(), ts, epochs=self.vcfg.epochs, batch=self.vcfg.batch,
                            lr=self.vcfg.lr, seed=self.vcfg.seed)
        self.net.rescale_output(max(sd, 1e-6), mu)
        self.train_history = hist
        pred = self.net.forward(x.contiguous())
        self._pred_sample = pred[: min(pred.numel(), 1 << 24)].double()
        self.calibrate()
        return hist

    def calibrate(self, ref_quantile=None, k_max=None):
        """Place the window: states at the ``ref_quantile`` of the predicted
        importance over the training rows get the smallest target weight 1/k_max."""
        import torch

        if ref_quantile is not None:
            self.vcfg.ref_quantile = float(ref_quantile)
        if k_max is not None:
            self.vcfg.k_max = float(k_max)
        q = torch.quantile(self._pred_sample, float(self.vcfg.ref_quantile))
        self.log_I_ref = float(q) - float(np.log(self.vcfg.k_max))

    # -- production ---------------------------------------------------------
    def _alloc(self, Ne):
        v = self.vcfg
        cap = int(Ne * min(v.k_max, v.n_max) * 1.25) + 4096
        R = max(int(v.rows_per_round), int(v.rows_far))  # slot stride of the round rows
        if self._banks is not None and self._banks[0].capacity >= cap and self._rows_R == R:
            return
        self._rows_R = R
        import torch

        dev = self.pipe.dev
        self._banks = [ParticleBank(cap, dev, self.pipe.cfg.E_cut_keV) for _ in range(2)]
        self._rows = alloc_rows(cap, R, True, dev)
        self._seg_count = torch.zeros(cap, dtype=torch.int64, device=dev)
        self._exit = torch.zeros(cap, dtype=torch.int8, device=dev)
        self._x = torch.empty((cap, 6), dtype=torch.float32, device=dev)
        self._y = torch.empty(cap, dtype=torch.float32, device=dev)
        self._count = torch.empty(cap, dtype=torch.int32, device=dev)
        self._wc = torch.empty(cap, dtype=torch.float64, device=dev)
        m = cap * R
        # two record sets used in turn: round r+1 gathers into one while the
        # scoring stream still reads round r's
        self._sets = [{
            "seg": {"r_mid": torch.empty((m, 3), dtype=torch.float32, device=dev),
                    "v_hat": torch.empty((m, 3), dtype=torch.float32, device=dev),
                    "L": torch.empty(m, dtype=torch.float32, device=dev),
                    "E": torch.empty(m, dtype=torch.float32, device=dev)},
            "w": torch.empty(m, dtype=torch.float32, device=dev),
            "w2": torch.empty(m, dtype=torch.float32, device=dev),
        } for _ in range(2)]
        self._set_free = [None, None]
        self._w2c = torch.empty(cap, dtype=torch.float64, device=dev)
        self._rows_next = torch.empty(cap, dtype=torch.int32, device=dev)

    def _window(self, n, bank, use_vr, n_max):
        a = _lib.VrWindowArgs()
        a.n = int(n)
        a.exit_code = self._exit.data_ptr()
        a.log_importance = self._y.data_ptr()
        a.weight, a.key = bank.weight.data_ptr(), bank.key.data_ptr()
        if use_vr:
            a.log_I_ref = float(self.log_I_ref)
            a.w_min, a.w_max = 1.0 / float(self.vcfg.k_max), 1.0
            a.ratio_lo, a.ratio_hi = float(self.vcfg.ratio_lo), float(self.vcfg.ratio_hi)
        else:
            a.log_I_ref, a.w_min, a.w_max = 0.0, 1.0, 1.0
            a.ratio_lo, a.ratio_hi = 1.0, 1.0
        a.n_max = int(n_max)
        a.count, a.w_child = self._count.data_ptr(), self._wc.data_ptr()
        a.weight2, a.w2_child = bank.weight2.data_ptr(), self._w2c.data_ptr()
        if use_vr:
            a.rows_near, a.rows_far = int(self.vcfg.rows_per_round), int(self.vcfg.rows_far)
            a.far_ratio = float(self.vcfg.far_ratio)
            a.rows_next = self._rows_next.data_ptr()
        _lib.check(self.lib.hr_vr_window(C.byref(a), _stream()), "hr_vr_window")

    def run_batch(self, batch_index, Ne, Ne_total, spec_lines, brem_acc, use_vr=True, stats=None,
                  roulette=False):
        """Histories ``[b*Ne, (b+1)*Ne)`` in rounds; adds weighted tallies
        (lines per electron of ``Ne_total``; brem as the raw sum) to the outputs.

        ``stats`` (a dict) accumulates weighted observables: ``rows`` (sum of
        row weights), ``exit_<code>`` (weight ending with each exit code),
        ``wmass`` (the rows' expected peak-window line content, weighted: a
        smooth observable free of the bin-hit lottery, for bias checks),
        ``particles`` (particle-rounds) and ``rounds``."""
        import torch

        p, v = self.pipe, self.vcfg
        if use_vr and self.log_I_ref is None:
            raise RuntimeError("train the importance network (fit) before a learned-importance run")
        self._alloc(Ne)
        R = self._rows_R  # slot stride
        src, dst = self._banks
        cap = src.capacity
        _lib.call("hr_stream_keys", int(p.cfg.seed), int(batch_index) * Ne, Ne, src.key.data_ptr(), None)
        src.pos[:Ne].zero_()
        src.dirs[:Ne].copy_(p.beam_d.expand(Ne, 3))
        src.E[:Ne].fill_(p.cfg.E0_keV)
        src.clock[:Ne].zero_()
        src.weight[:Ne].fill_(1.0)
        src.weight2[:Ne].fill_(1.0)
        src.rows[:Ne].fill_(int(v.rows_per_round))
        self._row0 = 0
        rkey = _roulette_key(p.cfg.seed, batch_index) if roulette else None
        src.prim[:Ne].copy_(torch.arange(int(batch_index) * Ne, (int(batch_index) + 1) * Ne, device=p.dev))
        n = Ne
        rows = self._rows
        tail_n = max(64, int(v.tail_fraction * Ne))
        # rounds are scored on a second stream while the next round is
        # transported; scoring keeps its order, so tallies are those of the
        # one-stream schedule bit for bit (stats need synchronous reductions)
        overlap = self.overlap and stats is None
        main = torch.cuda.current_stream()
        if overlap:
            if self._score_stream is None:
                self._score_stream = torch.cuda.Stream(device=p.dev)
            sb = self._score_stream
            sb.wait_stream(main)
        k_set = 0
        while n > 0:
            self.rounds += 1
            # the last stragglers take long rounds: fewer launches, decisions rarer where they matter least
            tail = n <= tail_n
            Rr = R if not tail else max(R, min(int(v.tail_rows), (cap * R) // n))
            rows_each = src.rows.data_ptr() if (use_vr and not tail) else None
            if not use_vr and not tail:
                Rr = int(v.rows_per_round)
            a = transport_args(p.tables, p.control, (0, n, Rr, src.key),
                               (None, src.clock, src.pos, src.dirs, src.E, src.E_cut), rows,
                               (self._seg_count, self._exit), (p.control.straggle_on, src.stragg))
            _lib.check(self.lib.hr_vr_round(C.byref(a), Rr, rows_each, _stream()), "hr_vr_round")
            # rows -> weighted FP32 scoring records (record set k_set, once its last reader is done)
            rec = self._sets[k_set]
            if overlap and self._set_free[k_set] is not None:
                main.wait_event(self._set_free[k_set])
            m = self._seg_count[:n].clamp(max=Rr)
            incl = torch.cumsum(m, 0)
            n_rows = int(incl[-1])
            off = incl - m
            g = _lib.VrGatherArgs()
            g.n, g.rows, g.n_out = n, Rr, n_rows
            g.seg_count, g.row_offset = self._seg_count.data_ptr(), off.data_ptr()
            g.weight, g.prim = src.weight.data_ptr(), None
            g.seg_mid, g.seg_dir = rows["seg_mid"].data_ptr(), rows["seg_dir"].data_ptr()
            g.seg_len, g.seg_E, g.seg_E_end = rows["seg_len"].data_ptr(), rows["seg_E"].data_ptr(), \
                rows["seg_E_end"].data_ptr()
            s = rec["seg"]
            g.r_mid, g.v_hat, g.L, g.E = s["r_mid"].data_ptr(), s["v_hat"].data_ptr(), s["L"].data_ptr(), \
                s["E"].data_ptr()
            g.row_weight, g.row_prim = rec["w"].data_ptr(), None
            g.weight2, g.row_weight2 = src.weight2.data_ptr(), rec["w2"].data_ptr()
            _lib.check(self.lib.hr_vr_gather(C.byref(g), _stream()), "hr_vr_gather")
            if stats is not None:
                stats["rows2"] = stats.get("rows2", 0.0) + float((m.to(torch.float64) * src.weight2[:n]).sum())
                w = src.weight[:n]
                stats["rows"] = stats.get("rows", 0.0) + float((m.to(torch.float64) * w).sum())
                codes = self._exit[:n].to(torch.int64)
                ew = torch.zeros(8, dtype=torch.float64, device=p.dev).index_add_(0, codes, w).cpu().numpy()
                for c in range(8):
                    if c != EXIT_SUSPENDED:
                        stats[f"exit_{c}"] = stats.get(f"exit_{c}", 0.0) + float(ew[c])
                stats["particles"] = stats.get("particles", 0) + n
                stats["rounds"] = stats.get("rounds", 0) + 1
            if n_rows:
                seg = {k: t[:n_rows] for k, t in s.items()}
                if stats is not None:
                    wm = self.row_masses(seg, Ne_total) * rec["w"][:n_rows].double()
                    stats["wmass"] = stats.get("wmass", 0.0) + float(wm.sum())
                if overlap:
                    ready = torch.cuda.Event()
                    ready.record(main)
                    with torch.cuda.stream(sb):
                        sb.wait_event(ready)
                        self.score_lines(seg, Ne_total, spec_lines, rec["w"][:n_rows], rkey)
                        p.brem.reduce(seg, brem_acc, row_weight=rec["w2"][:n_rows])
                        done = torch.cuda.Event()
                        done.record(sb)
                    self._set_free[k_set] = done
                    k_set ^= 1
                else:
                    self.score_lines(seg, Ne_total, spec_lines, rec["w"][:n_rows], rkey)
                    p.brem.reduce(seg, brem_acc, row_weight=rec["w2"][:n_rows])
            # split / roulette the suspended particles
            if use_vr:
                self._features(n, src.pos, src.dirs, src.E, None, self._x)
                self.net.forward(self._x[:n], out=self._y[:n])
            self._window(n, src, use_vr, v.n_max)
            incl = torch.cumsum(self._count[:n].to(torch.int64), 0)
            total = int(incl[-1])
            if total > cap:
                survivors = int((self._count[:n] > 0).sum())
                self._window(n, src, use_vr, max(1, min(v.n_max, cap // max(survivors, 1))))
                incl = torch.cumsum(self._count[:n].to(torch.int64), 0)
                total = int(incl[-1])
            if total:
                off = incl - self._count[:n].to(torch.int64)
                sp = _lib.VrSpawnArgs()
                sp.n = n
                sp.count, sp.offset, sp.w_child = self._count.data_ptr(), off.data_ptr(), self._wc.data_ptr()
                sp.pos, sp.dirs, sp.E, sp.clock = src.pos.data_ptr(), src.dirs.data_ptr(), src.E.data_ptr(), \
                    src.clock.data_ptr()
                sp.key, sp.prim = src.key.data_ptr(), src.prim.data_ptr()
                sp.out_pos, sp.out_dirs, sp.out_E = dst.pos.data_ptr(), dst.dirs.data_ptr(), dst.E.data_ptr()
                sp.out_clock, sp.out_weight = dst.clock.data_ptr(), dst.weight.data_ptr()
                sp.out_key, sp.out_prim = dst.key.data_ptr(), dst.prim.data_ptr()
                if use_vr:
                    sp.rows_next, sp.out_rows = self._rows_next.data_ptr(), dst.rows.data_ptr()
                sp.w2_child, sp.out_weight2 = self._w2c.data_ptr(), dst.weight2.data_ptr()
                _lib.check(self.lib.hr_vr_spawn(C.byref(sp), _stream()), "hr_vr_spawn")
            src, dst = dst, src
            n = total
        if overlap:
            main.wait_stream(sb)
        self._banks = [src, dst]

    def measure(self, n_batches, Ne, mode="vr", first_batch=0):
        """Per-batch peak tallies and wall time of ``n_batches`` batches of ``Ne``.

        ``mode``: ``"analog"`` (the plain ``CasePipeline`` transport + scoring),
        ``"roulette"`` (analog plus line roulette by exact row mass),
        ``"rounds"`` (the round machinery with every window at weight 1),
        ``"vr"`` (learned splitting/roulette plus line roulette) or
        ``"vr_split"`` (learned splitting/roulette only).
        Each batch is an independent estimate (tallies per electron of the batch).
        """
        import torch

        p = self.pipe
        dev = p.dev
        n_bins = p.E_grid.size
        vals = []
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for b in range(int(n_batches)):
            lines = torch.zeros(n_bins, dtype=torch.float64, device=dev)
            brem = torch.zeros(n_bins, dtype=torch.float64, device=dev)
            if mode in ("analog", "roulette"):
                res = p.transport(first_batch + b)
                seg = segments_fp32({
                    "r_mid": res.rows["seg_mid"], "v_hat": res.rows["seg_dir"], "L_ang": res.rows["seg_len"],
                    "E_keV": res.rows["seg_E"],
                    "E_repr_keV": 0.5 * (res.rows["seg_E"] + res.rows["seg_E_end"]),
                }, dev)
                if mode == "analog":
                    p.lines.spectrum(seg, p.cfg.Ne, lines)
                else:
                    self._row0 = 0
                    self.score_lines(seg, p.cfg.Ne, lines, None, _roulette_key(p.cfg.seed, first_batch + b))
                p.brem.reduce(seg, brem)
            elif mode in ("rounds", "vr", "vr_split"):
                self.run_batch(first_batch + b, Ne, Ne, lines, brem, use_vr=(mode != "rounds"),
                               roulette=(mode == "vr"))
            else:
                raise ValueError(f"unknown mode {mode!r}")
            m = torch.as_tensor(self.window_bins, device=dev)
            vals.append(lines[m].sum() * self.dE)
        vals = torch.stack(vals).cpu().numpy()
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        return vals, elapsed


def _roulette_key(seed, batch_index):
    """Stream key of a batch's line-roulette draws (disjoint from transport keys)."""
    from ..transport.beam import child_stream_root

    return (child_stream_root(int(seed), 7, 6) + 0x9E3779B97F4A7C15 * (int(batch_index) + 1)) % (1 << 64)


def figure_of_merit(vals, elapsed):
    """``(mean, rel_err, FOM = 1 / (rel_err^2 * t))`` of independent batch estimates."""
    vals = np.asarray(vals, dtype=float)
    mean = float(vals.mean())
    rel = float(vals.std(ddof=1) / np.sqrt(vals.size) / mean) if mean != 0 and vals.size > 1 else float("inf")
    return mean, rel, 1.0 / (rel * rel * elapsed)
