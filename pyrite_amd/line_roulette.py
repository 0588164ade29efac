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
from dataclasses import asdict, dataclass
from typing import Any

import numpy as np

from . import _lib

#: Versioned estimator: pre-screen, strata and draw stream of this policy.
LINE_ROULETTE_MODEL = "exact-line-mass-two-stratum-v1"

#: Case keys whose line scoring is not an additive per-history incoherent sum.
_REFUSED_KEYS = {
    "coherent_emission": "coherent emission squares a sum of amplitudes",
    "temporal_profile": "temporal profiles accumulate unweighted line boxes",
    "channeling": "channeling joins zones across a history",
    "max_dE_frac": "substeps of a flight add coherently, so a row has no separate line mass",
}


def _unit_interval(name, value):
    if isinstance(value, bool) or not isinstance(value, int | float) or not 0 < value <= 1:
        raise ValueError(f"line roulette {name} must lie in (0, 1]")


@dataclass(frozen=True, slots=True, kw_only=True)
class LineRoulette:
    """Keep low-mass histories with probability ``floor`` and weight ``1/floor``."""

    floor: float
    keep_fraction: float = 0.1
    pilot_histories: int = 128
    model: str = LINE_ROULETTE_MODEL

    def __post_init__(self) -> None:
        if self.model != LINE_ROULETTE_MODEL:
            raise ValueError(f"unknown line roulette model {self.model!r}")
        _unit_interval("floor", self.floor)
        _unit_interval("keep_fraction", self.keep_fraction)
        pilot = self.pilot_histories
        if isinstance(pilot, bool) or not isinstance(pilot, int) or pilot < 1:
            raise ValueError("line roulette pilot_histories must be a positive integer")

    def to_dict(self) -> dict[str, Any]:
        return asdict(self)

    @classmethod
    def from_dict(cls, payload: Mapping[str, Any]) -> LineRoulette:
        if not isinstance(payload, Mapping):
            raise TypeError("line_roulette must be a mapping")
        unknown = sorted(set(payload) - set(cls.__dataclass_fields__))
        if unknown:
            raise ValueError(f"unknown line roulette field(s) {unknown}")
        return cls(**payload)

    def validate_case(self, case: Mapping[str, Any]) -> None:
        """Refuse routes whose line tally is not a sum over primary histories."""
        for key, reason in _REFUSED_KEYS.items():
            if case.get(key):
                raise ValueError(f"line roulette does not support {key}: {reason}")
        if (case.get("transition_radiation") or {}).get("model") == "two-beam":
            raise ValueError("line roulette does not support two-beam line scoring")


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
