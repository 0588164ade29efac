"""Row-event contract of midpoint transport rows (test helper).

Restates the checks of the reference's ``check_segment_event_contract``
(montecarlo/transport/events.py:156-322) for the event codes the LUT path
emits (ELASTIC 0, SUBSTEP 1, LAYER 2, EXIT_TOP 4, EXIT_BOTTOM 5, EXIT_SIDE 6,
CUTOFF 7; events.py:36-43): no grooves, channeling or hard events here.
"""

import numpy as np

ELASTIC, SUBSTEP, LAYER, GROOVE, EXIT_TOP, EXIT_BOTTOM, EXIT_SIDE, CUTOFF = range(8)
KNOWN = (ELASTIC, SUBSTEP, LAYER, EXIT_TOP, EXIT_BOTTOM, EXIT_SIDE, CUTOFF)
TERMINAL = (EXIT_TOP, EXIT_BOTTOM, EXIT_SIDE, CUTOFF)
CONTINUING = (SUBSTEP,)  # the next row resumes the same flight
KEEPS_DIRECTION = (SUBSTEP, LAYER)


def check(rows, rtol=1e-12, atol_ang=1e-5):
    """Raise AssertionError naming the first violated rule.

    ``rows``: compacted midpoint rows (numpy) with seg_mid, seg_dir, seg_len,
    seg_E, seg_E_end, seg_t0, seg_t_end, seg_id, seg_flight, seg_substep,
    seg_event."""
    kind = np.asarray(rows["seg_event"]).astype(np.int64)
    n = kind.size
    if n == 0:
        return
    bad = ~np.isin(kind, KNOWN)
    assert not bad.any(), f"unknown event codes {sorted(set(kind[bad].tolist()))}"
    eid = np.asarray(rows["seg_id"], dtype=np.int64)
    fid = np.asarray(rows["seg_flight"], dtype=np.int64)
    sid = np.asarray(rows["seg_substep"], dtype=np.int64)
    order = np.lexsort((sid, fid, eid))
    eid, fid, sid, kind = eid[order], fid[order], sid[order], kind[order]
    L = np.asarray(rows["seg_len"], dtype=float)[order]
    v = np.asarray(rows["seg_dir"], dtype=float).reshape(-1, 3)[order]
    mid = np.asarray(rows["seg_mid"], dtype=float).reshape(-1, 3)[order]
    E0 = np.asarray(rows["seg_E"], dtype=float)[order]
    E1 = np.asarray(rows["seg_E_end"], dtype=float)[order]
    t0 = np.asarray(rows["seg_t0"], dtype=float)[order]
    t1 = np.asarray(rows["seg_t_end"], dtype=float)[order]
    start = mid - 0.5 * L[:, None] * v
    end = mid + 0.5 * L[:, None] * v

    def first_violation(mask, rule):
        if mask.any():
            j = int(np.flatnonzero(mask)[0])
            raise AssertionError(f"{rule}: electron {eid[j]} flight {fid[j]} substep {sid[j]} event {kind[j]}")

    opens = np.r_[True, eid[1:] != eid[:-1]]
    closes = np.r_[opens[1:], True]
    first_violation(opens & ((fid != 0) | (sid != 0)), "history does not open at flight 0, substep 0")
    terminal = np.isin(kind, TERMINAL)
    first_violation(terminal & ~closes, "terminal event before the history's last row")
    first_violation(closes & ~terminal, "history's last row is not terminal")

    a = np.flatnonzero(~closes)
    b = a + 1
    ka = kind[a]
    pair = np.zeros(n, dtype=bool)

    def on_pairs(mask):
        out = pair.copy()
        out[a[mask]] = True
        return out

    cont = np.isin(ka, CONTINUING)
    first_violation(on_pairs(cont & ((fid[b] != fid[a]) | (sid[b] != sid[a] + 1))),
                    "substep row does not continue its flight")
    first_violation(on_pairs(~cont & ((fid[b] != fid[a] + 1) | (sid[b] != 0))),
                    "event row does not open the next flight")
    keeps = np.isin(ka, KEEPS_DIRECTION)
    moved = np.any(np.abs(v[b] - v[a]) > 1e-12, axis=1)
    first_violation(on_pairs(keeps & moved), "direction changed across a direction-preserving event")

    def near(x, y, scale):
        return np.abs(x - y) <= rtol * np.maximum(np.abs(scale), 1.0)

    first_violation(on_pairs(~near(E0[b], E1[a], E1[a])), "energy jumps between rows")
    first_violation(on_pairs(~near(t0[b], t1[a], t1[a])), "clock jumps between rows")
    gap = np.linalg.norm(start[b] - end[a], axis=1)
    first_violation(on_pairs(gap > atol_ang), "position jumps between rows")
    # within a row: energy never rises, the clock never runs back
    first_violation(E1 > E0 * (1.0 + rtol), "energy rises along a row")
    first_violation(t1 < t0 * (1.0 - rtol) - atol_ang, "clock runs back along a row")
