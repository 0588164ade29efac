"""Numpy restatement of the depth-dose tally (TEST INFRASTRUCTURE ONLY).

Each row's loss E - E_end is spread over the depth bins its flight crosses in
proportion to the depth covered in each (edge bins also take what lies
outside the grid); a row ending in the cutoff event (code 7,
src/pyrite/montecarlo/transport/events.py:43) deposits its residual E_end at
its end depth.  Pins: the exact energy balance sum(dose) + sum(escaped E) =
Ne E0 on the oracle's own histories, and hand-computed single-row cases.
"""

from __future__ import annotations

import numpy as np


def depth_dose(rows, z0, dz, n_bins, scale=1.0, weight=None):
    """``rows``: compacted rows (seg_mid, seg_dir, seg_len, seg_E, seg_E_end, seg_event)."""
    mid = np.asarray(rows["seg_mid"], dtype=float).reshape(-1, 3)
    dz_row = 0.5 * np.asarray(rows["seg_len"], dtype=float) * np.asarray(rows["seg_dir"], dtype=float).reshape(-1, 3)[:, 2]
    za, zb = mid[:, 2] - dz_row, mid[:, 2] + dz_row
    dE = np.asarray(rows["seg_E"], dtype=float) - np.asarray(rows["seg_E_end"], dtype=float)
    w = np.full(dE.size, float(scale)) if weight is None else np.asarray(weight, dtype=float) * scale
    out = np.zeros(int(n_bins))
    last = n_bins - 1

    def bin_of(z):
        return np.clip(np.floor((z - z0) / dz), 0.0, last).astype(np.int64)

    lo, hi = np.minimum(za, zb), np.maximum(za, zb)
    b0, b1 = bin_of(lo), bin_of(hi)
    single = (b0 == b1) | ~(hi > lo)
    np.add.at(out, b0[single], (w * dE)[single])
    multi = np.flatnonzero(~single)
    if multi.size:
        m_lo, m_hi, m_b0, m_b1 = lo[multi], hi[multi], b0[multi], b1[multi]
        m_w = (w * dE)[multi] / (m_hi - m_lo)
        for k in range(int((m_b1 - m_b0).max()) + 1):
            b = m_b0 + k
            on = b <= m_b1
            e_lo = np.where(b == m_b0, m_lo, z0 + b * dz)
            e_hi = np.where(b == m_b1, m_hi, z0 + (b + 1) * dz)
            np.add.at(out, b[on], (m_w * (e_hi - e_lo))[on])
    ev = rows.get("seg_event")
    if ev is not None:
        cut = np.asarray(ev) == 7
        E_end = np.asarray(rows["seg_E_end"], dtype=float)
        np.add.at(out, bin_of(zb[cut]), (w * E_end)[cut])
    return out
