"""Oracle transport core: invariances and the row-event contract.

The invariances are the reference's own (test_transport_per_electron.py:
capacity replay at several caps, batch-size invariance, electron-major output)
plus the event contract of transport/events.py:156 check_segment_event_contract.
"""

import numpy as np
import pytest

from tests import event_contract
from tests.transport_cases import CASES, HOPG_5MEV, MULTILAYER, SI_30KEV


def _run(oracle, case, cap=4096, threads=0, e_start=0, e_count=None, max_steps=400_000, **kw):
    lt, lut = case.tables()
    pos, dirs, E, E_cut = case.state()
    keys = oracle.stream_keys(case.seed, case.Ne)
    return oracle.run_transport(
        lt, lut, keys, pos, dirs, E, E_cut, cap=cap, e_start=e_start, e_count=e_count,
        max_steps=max_steps, energy_model=case.energy_model, max_dE_frac=case.max_dE_frac,
        straggle_on=case.straggle, finite=case.finite, threads=threads, **kw)


@pytest.mark.parametrize("case", CASES, ids=lambda c: c.name)
def test_event_contract(oracle, case):
    rows, count, code, state = _run(oracle, case)
    assert count.max() < 4096
    r = oracle.compact(rows, count, 4096)
    n = len(r["seg_len"])
    assert n == count.sum()
    assert np.all(r["seg_len"] >= 0.0)
    # electron-major, step-minor; a history's rows are contiguous
    assert np.all(np.diff(r["seg_id"]) >= 0)
    starts = np.concatenate([[0], np.cumsum(count)[:-1]])
    last = starts + count - 1
    ev_last = r["seg_event"][last]
    # exit code <-> last event (events.py:36-131; exit codes _jit_device.py:63-69)
    # (the frozen schema carries no event column: batching.py:398-400)
    want = {0: 7, 1: 4, 2: 5, 3: 6}
    if case.energy_model == 1:
        for c, ev in zip(code, ev_last):
            assert want[int(c)] == int(ev)
    # energy never increases along a row; the clock only grows
    if case.energy_model == 1:
        assert np.all(r["seg_E_end"] <= r["seg_E"] + 1e-12)
        assert np.all(r["seg_t_end"] >= r["seg_t0"])
    # cutoff histories end exactly at E_cut
    stopped = code == 0
    if stopped.any():
        np.testing.assert_array_equal(state["E_keV"][stopped], case.E_cut)
    # the full row-event contract (events.py:156): flight/substep sequencing,
    # terminal rows last, continuity of energy, clock and position
    if case.energy_model == 1:
        event_contract.check(r)


def test_event_contract_catches_a_broken_history(oracle):
    """Planted in the oracle's own rows: a dropped row, a moved row end and a
    swapped event code each fail the contract."""
    rows, count, code, state = _run(oracle, SI_30KEV)
    r = oracle.compact(rows, count, 4096)
    event_contract.check(r)
    keep = np.ones(len(r["seg_len"]), dtype=bool)
    k = int(np.flatnonzero(r["seg_event"] == 0)[3])
    keep[k] = False  # drop an elastic row from the middle of a history
    with pytest.raises(AssertionError, match="flight|position|energy|clock"):
        event_contract.check({key: v[keep] for key, v in r.items()})
    moved = {key: v.copy() for key, v in r.items()}
    moved["seg_len"][k] *= 1.5
    with pytest.raises(AssertionError, match="position"):
        event_contract.check(moved)
    swapped = {key: v.copy() for key, v in r.items()}
    swapped["seg_event"][k] = 1  # an elastic row claimed as a substep
    with pytest.raises(AssertionError, match="substep|flight"):
        event_contract.check(swapped)


@pytest.mark.parametrize("case", [SI_30KEV, MULTILAYER], ids=lambda c: c.name)
def test_capacity_thread_and_block_invariance(oracle, case):
    ref_rows, ref_count, ref_code, ref_state = _run(oracle, case, cap=4096, threads=1)
    ref = oracle.compact(ref_rows, ref_count, 4096)
    # tiny cap: rows past cap are dropped but the count and the final state are unchanged
    rows, count, code, state = _run(oracle, case, cap=4, threads=8)
    np.testing.assert_array_equal(count, ref_count)
    np.testing.assert_array_equal(code, ref_code)
    for k in ("pos", "dirs", "E_keV", "clock"):
        np.testing.assert_array_equal(state[k], ref_state[k])
    # blocks [0, 150) + [150, Ne) reproduce the single run bit for bit
    parts = []
    for s, n in ((0, 150), (150, case.Ne - 150)):
        rws, cnt, cd, st = _run(oracle, case, cap=4096, e_start=s, e_count=n)
        parts.append(oracle.compact(rws, cnt, 4096))
    for k in ref:
        np.testing.assert_array_equal(np.concatenate([p[k] for p in parts]), ref[k])


def test_straggling_off_path_draws_unchanged(oracle):
    # Straggling samples on its own key domain: the first flight's length and the
    # first collision's angles are identical with straggling on and off.
    import dataclasses

    off = dataclasses.replace(HOPG_5MEV, straggle=False)
    r_on = oracle.compact(*_run(oracle, HOPG_5MEV, max_steps=1)[:2], 4096)
    r_off = oracle.compact(*_run(oracle, off, max_steps=1)[:2], 4096)
    np.testing.assert_array_equal(r_on["seg_len"], r_off["seg_len"])
    np.testing.assert_array_equal(r_on["seg_dir"], r_off["seg_dir"])


def test_hopg_5mev_transmits(oracle):
    rows, count, code, state = _run(oracle, HOPG_5MEV)
    # 100 um of graphite is thin for 5 MeV electrons: nearly all transmit, losing ~40 keV
    assert (code == 2).mean() > 0.95
    dE = HOPG_5MEV.E0 - state["E_keV"][code == 2]
    assert 10.0 < np.median(dE) < 80.0
    assert 15 < count.mean() < 200
