"""Per-primary line roulette on the hip backend.

Mirrors the reference's opt-in ``LineRoulette`` policy
(``src/pyrite/_line_roulette.py:29-80``) and its screen
(``montecarlo/runner/line_roulette.py``): each primary's exact line mass
``m_i = sum w pi / a`` over its surviving lines comes from a prologue pass
with no grid reduction; the leading ``pilot_histories`` primaries are scored
surely and fix the threshold as the ``1 - keep_fraction`` quantile of their
masses; later primaries below it are scored with probability ``floor`` and
weight ``1 / floor``, from a uniform keyed only on (seed, primary id)
(``hr_roulette_uniforms``).  A primary without surviving lines is skipped.
The strata are scored by the production line scorers through the per-row
weight, so the dense and windowed routes share one estimator; brem is not
rouletted.

Like the reference, the screen needs whole primaries in one block: the
batched route provides them, the streamed (continuation) route does not and
refuses the policy.
"""

from __future__ import annotations

import ctypes as C
from collections.abc import Mapping
from dataclasses import dataclass
from typing import Any

import numpy as np

from . import _lib

# Interface string and field names are the reference policy's
# (src/pyrite/_line_roulette.py:29-39), so a saved case selects the same
# estimator; everything below them is this backend's own.
LINE_ROULETTE_MODEL = "exact-line-mass-two-stratum-v1"
_FIELDS = ("floor", "keep_fraction", "pilot_histories", "model")


def _check_fields(floor, keep_fraction, pilot_histories, model):
    """Every reason the four fields do not form a usable policy (empty if none)."""
    bad = []
    for name, x in (("floor", floor), ("keep_fraction", keep_fraction)):
        if type(x) not in (int, float) or not (0.0 < float(x) <= 1.0):
            bad.append(f"{name}={x!r} is not a probability in (0, 1]")
    if type(pilot_histories) is not int or pilot_histories <= 0:
        bad.append(f"pilot_histories={pilot_histories!r} is not a positive int")
    if model != LINE_ROULETTE_MODEL:
        bad.append(f"model={model!r} is not {LINE_ROULETTE_MODEL!r}")
    return bad


@dataclass(frozen=True, slots=True, kw_only=True)
class LineRoulette:
    """Per-primary line roulette: below the pilot threshold a primary is
    scored with probability ``floor`` at weight ``1/floor``."""

    floor: float
    keep_fraction: float = 0.1
    pilot_histories: int = 128
    model: str = LINE_ROULETTE_MODEL

    def __post_init__(self) -> None:
        bad = _check_fields(self.floor, self.keep_fraction, self.pilot_histories, self.model)
        if bad:
            raise ValueError("line roulette policy: " + "; ".join(bad))

    def to_dict(self) -> dict[str, Any]:
        return {f: getattr(self, f) for f in _FIELDS}

    @classmethod
    def from_dict(cls, payload: Mapping[str, Any]) -> LineRoulette:
        """Policy from a case's ``line_roulette`` mapping (the four fields only)."""
        if not isinstance(payload, Mapping):
            raise TypeError(f"line roulette policy must be a mapping, not {type(payload).__name__}")
        extra = set(payload).difference(_FIELDS)
        if extra:
            raise ValueError(f"line roulette policy: fields {sorted(extra)} are not part of it")
        return cls(**dict(payload))

    def validate_case(self, cfg, max_dE_frac: float = 0.0, route: str = "batched") -> None:
        """Refuse a hip-backend case whose line tally is not a sum of
        independent per-primary terms, which the two-stratum estimator needs:
        the coherent route squares a field summed over rows (``cfg.coherent``),
        the streamed route scores a primary's rows over several rounds before
        its whole line mass is known, and sub-stepped flights
        (``max_dE_frac > 0``) split one flight into rows that the screen would
        weigh apart."""
        if getattr(cfg, "coherent", False):
            raise ValueError("line roulette needs incoherent line scoring; this case sets coherent=True")
        if route != "batched":
            raise ValueError(f"line roulette needs whole primaries per block; route {route!r} splits them")
        if max_dE_frac > 0.0:
            raise ValueError(f"line roulette needs one row per flight; max_dE_frac={max_dE_frac} sub-steps flights")


def roulette_uniforms(seed: int, first_id: int, n: int, device):
    """Device FP64 uniforms of primaries ``[first_id, first_id + n)``."""
    import torch

    out = torch.empty(int(n), dtype=torch.float64, device=device)
    _lib.call("hr_roulette_uniforms", int(seed) & 0xFFFFFFFFFFFFFFFF, int(first_id), int(n), out.data_ptr(), None)
    return out


class LineScreen:
    """Masses, threshold and draws of one line population (primaries ``[0, n)``)."""

    def __init__(self, policy: LineRoulette | Mapping, seed: int, n_electrons: int, device):
        import torch

        self.policy = policy if isinstance(policy, LineRoulette) else LineRoulette.from_dict(policy)
        self.seed = int(seed)
        self.n = int(n_electrons)
        self.dev = device
        self.mass = torch.zeros(self.n, dtype=torch.float64, device=device)
        self.probability = torch.ones(self.n, dtype=torch.float64, device=device)
        self.kept = torch.zeros(self.n, dtype=torch.uint8, device=device)
        self.rows = torch.zeros(self.n, dtype=torch.int64, device=device)
        self.threshold: float | None = None

    def _resolve_threshold(self) -> None:
        pilot = self.mass[: self.policy.pilot_histories].cpu().numpy()
        self.threshold = float(np.quantile(pilot, 1.0 - self.policy.keep_fraction))

    def row_weights(self, scorer, seg, prim, Ne):
        """Screen the whole primaries of ``seg`` (row primary ids ``prim``, ascending)
        and return each row's stratum weight ``kept / probability`` (FP32)."""
        import torch

        from .spectrum.lines import LINE_BLOCK

        n_seg = int(seg["L"].shape[0])
        n_dir, n_g = scorer.n_dir, len(scorer.tables.wm)
        step = max(1, LINE_BLOCK // (n_dir * n_g))
        seg_mass = torch.empty(n_seg, dtype=torch.float64, device=self.dev)
        for j0 in range(0, n_seg, step):
            m = min(step, n_seg - j0)
            E_r, a, w = scorer.lines(seg, 1.0 / float(Ne), j0, m)
            _lib.call("hr_line_mass", E_r.data_ptr(), a.data_ptr(), w.data_ptr(), m, n_dir, n_g,
                      -np.inf, np.inf, 1.0, seg_mass[j0:].data_ptr(), None)
        lo, hi = (int(prim[0]), int(prim[-1]) + 1) if n_seg else (0, 0)
        if n_seg:
            if lo < 0 or hi > self.n:
                raise ValueError("line roulette: primary ids outside the screened population")
            block = slice(lo, hi)
            self.mass[block] = 0.0
            self.mass.index_add_(0, prim, seg_mass)
            self.rows[block] = torch.bincount(prim - lo, minlength=hi - lo)
            if hi > self.policy.pilot_histories and self.threshold is None:
                self._resolve_threshold()
            thr = np.inf if self.threshold is None else self.threshold
            _lib.call("hr_line_screen", self.mass[block].data_ptr(), hi - lo, lo, int(self.policy.pilot_histories),
                      float(thr), float(self.policy.floor), self.seed & 0xFFFFFFFFFFFFFFFF,
                      self.probability[block].data_ptr(), self.kept[block].data_ptr(), None)
        wgt = self.kept.to(torch.float64) / self.probability
        return wgt.index_select(0, prim).to(torch.float32) if n_seg else torch.empty(0, dtype=torch.float32,
                                                                                       device=self.dev)

    def record(self) -> dict[str, Any]:
        """Provenance and the conditional variance of the summed line mass
        (``line_roulette.py:178-203``)."""
        m = self.mass.cpu().numpy()
        p = self.probability.cpu().numpy()
        k = self.kept.cpu().numpy().astype(bool)
        rows = self.rows.cpu().numpy()
        culled = p < 1.0
        return {
            **self.policy.to_dict(),
            "threshold": self.threshold,
            "primaries": self.n,
            "primaries_scored": int(k.sum()),
            "primaries_at_floor": int(culled.sum()),
            "primaries_zero_mass": int(np.count_nonzero(m == 0.0)),
            "rows": int(rows.sum()),
            "rows_scored": int(rows[k].sum()),
            "line_mass_analog": float(m.sum()),
            "line_mass_estimate": float(m[k] @ (1.0 / p[k])),
            "added_mass_variance": float((1.0 / p[culled] - 1.0) @ m[culled] ** 2),
        }
