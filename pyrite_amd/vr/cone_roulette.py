"""Learned importance roulette of detector-cone copies.

Cone splitting (``pyrite_amd.cone_split``) makes a copy at every elastic event
whose in-cone outcome is possible; without the reference's angle roulette most
copies carry little of the detector-bound line yield, yet each is transported
as a full history (copies that enter the cone near the slab face run long
along it).  Here the BF16 importance network of ``LearnedImportance`` learns,
from the copies' own futures, the peak-window yield a copy adds per unit
weight: its label is the square root of the summed second moments over the
bin offset of its rows' peak-window line content (the estimator's variance
per copy, cross terms left out), its inputs the copy's state at the event
(``E/E0``, direction, depth, tilt).  With ``per_cost`` the label is divided
by the square root of the copy's row count: a roulette that minimises
variance x cost keeps a copy with probability proportional to
sqrt(E[X^2] / cost), and a copy's cost is its rows.  In production every collected copy of
weight ``w`` survives with ``p = clamp(w I(x) / T, p_min, 1)`` and carries
``w / p`` (``hr_cone_roulette``; the draw is on a key of its own, so the
roulette depends on the copy's state only and every tally keeps its
expectation).  ``T`` is the ``keep_quantile`` of ``w I(x)`` over the training
copies.
"""

from __future__ import annotations



from .. import _lib
from ..spectrum.lines import segments_fp32
from .importance import LearnedImportance, VRConfig, _stream


class LearnedConeRoulette:
    """Train on the copies of analog-roulette-free cone splitting, then thin copies."""

    def __init__(self, pipe, vcfg: VRConfig | None = None, keep_quantile: float = 0.95, p_min: float = 1.0 / 256,
                 per_cost: bool = True):
        if pipe.cone_core is None:
            raise ValueError("the learned cone roulette needs a pipeline with cone_split")
        if not 0.0 <= keep_quantile < 1.0 or not 0.0 < p_min <= 1.0:
            raise ValueError("keep_quantile must lie in [0, 1) and p_min in (0, 1]")
        self.pipe = pipe
        self.li = LearnedImportance(pipe, vcfg or VRConfig(label="second"))
        self.keep_quantile = float(keep_quantile)
        self.p_min = float(p_min)
        self.per_cost = bool(per_cost)
        self.log_T = None
        self.stats = {"seen": 0, "kept": 0}

    def _features(self, cp):
        import torch

        n = int(cp["weight"].numel())
        x = torch.empty((n, 6), dtype=torch.float32, device=self.pipe.dev)
        self.li._features(n, cp["pos"], cp["dir"], cp["E_keV"], None, x)
        return x

    def copy_labels(self, cp):
        """``(x [n,6], y [n])``: each copy's state and the square root of its rows'
        summed peak-window second moments (per unit weight; with ``per_cost``
        per row of the copy)."""
        import torch

        from ..cone_split import transport_copies

        p = self.pipe
        n = int(cp["weight"].numel())
        x = self._features(cp)
        res = transport_copies(cp, p.tables, p.control, p.cfg.E_cut_keV, p.dev, compact=True)
        rows = res.rows
        y2 = torch.zeros(n, dtype=torch.float64, device=p.dev)
        if rows["seg_len"].numel():
            seg = segments_fp32({
                "r_mid": rows["seg_mid"], "v_hat": rows["seg_dir"], "L_ang": rows["seg_len"],
                "E_keV": rows["seg_E"], "E_repr_keV": 0.5 * (rows["seg_E"] + rows["seg_E_end"]),
            }, p.dev)
            m2 = self.li.row_second_moments(seg, p.cfg.Ne)
            # compacted rows are copy-major: per-copy sums from one running sum
            cnt = res.seg_count.to(torch.int64)
            ends = torch.cumsum(cnt, 0)
            cs = torch.zeros(m2.numel() + 1, dtype=torch.float64, device=p.dev)
            cs[1:] = torch.cumsum(m2, 0)
            y2 = cs[ends] - cs[ends - cnt]
            if self.per_cost:
                y2 = y2 / torch.clamp(cnt.double(), min=1.0)
        return x, torch.sqrt(torch.clamp(y2, min=0.0))

    def training_set(self, n_batches, first_batch=30_000):
        """Copies (and labels) of ``n_batches`` primary batches with every copy kept."""
        import torch

        p = self.pipe
        xs, ys, ws = [], [], []
        for b in range(int(n_batches)):
            p.cone_core.reset()
            p.transport(first_batch + b, compact=False, consume=lambda *a: None, core=p.cone_core)
            cp = p.cone_core.copies()
            if cp is None or cp["weight"].numel() == 0:
                continue
            x, y = self.copy_labels(cp)
            xs.append(x)
            ys.append(y)
            ws.append(cp["weight"])
        if not xs:
            raise RuntimeError("no cone copies in the training batches")
        return torch.cat(xs), torch.cat(ys), torch.cat(ws)

    def fit(self, x, y, w):
        """Train the network on ``log(eps + y)``; ``T`` at ``keep_quantile`` of ``w I(x)``."""
        import torch

        self.li.fit(x, y)
        log_imp = self.li.net.forward(x.contiguous()).double()
        s = torch.log(w) + log_imp
        self.log_T = float(torch.quantile(s[: 1 << 24], self.keep_quantile))
        return self.li.train_history

    def apply(self, cp):
        """The surviving copies of ``cp`` (same fields), with their new weights."""
        import torch

        if self.log_T is None:
            raise RuntimeError("fit() the cone roulette before apply()")
        n = int(cp["weight"].numel())
        if n == 0:
            return cp
        x = self._features(cp)
        log_imp = self.li.net.forward(x).float().contiguous()
        w = cp["weight"].clone()
        keep = torch.empty(n, dtype=torch.uint8, device=self.pipe.dev)
        _lib.call("hr_cone_roulette", n, log_imp.data_ptr(), cp["stream_key"].data_ptr(), float(self.log_T),
                  float(self.p_min), w.data_ptr(), keep.data_ptr(), _stream())
        idx = torch.nonzero(keep).squeeze(1)
        out = {k: (v.index_select(0, idx) if k != "killed" else v) for k, v in cp.items() if k != "weight"}
        out["weight"] = w.index_select(0, idx)
        self.stats["seen"] += n
        self.stats["kept"] += int(idx.numel())
        return out


__all__ = ["LearnedConeRoulette"]
