"""Depth-dose tally: the oracle's restatement against hand cases and the exact
energy balance of the oracle's own histories (CPU)."""

import numpy as np
import pytest

from oracle import tally
from tests.transport_cases import MULTILAYER, SI_30KEV_STRAGGLE


def _row(z_start, z_end, dE, event=0, E_end=1.0):
    L = abs(z_end - z_start)
    d = 1.0 if z_end >= z_start else -1.0
    return {"seg_mid": np.array([[0.0, 0.0, 0.5 * (z_start + z_end)]]), "seg_dir": np.array([[0.0, 0.0, d]]),
            "seg_len": np.array([L]), "seg_E": np.array([E_end + dE]), "seg_E_end": np.array([E_end]),
            "seg_event": np.array([event], dtype=np.int8)}


def test_hand_cases():
    # 10 bins of 1 A from z = 0
    out = tally.depth_dose(_row(0.5, 3.5, 6.0), 0.0, 1.0, 10)
    np.testing.assert_allclose(out[:5], [1.0, 2.0, 2.0, 1.0, 0.0])
    out = tally.depth_dose(_row(3.5, 0.5, 6.0), 0.0, 1.0, 10)  # moving up: same profile
    np.testing.assert_allclose(out[:5], [1.0, 2.0, 2.0, 1.0, 0.0])
    out = tally.depth_dose(_row(2.25, 2.75, 3.0, event=7, E_end=5.0), 0.0, 1.0, 10, scale=0.5)
    np.testing.assert_allclose(out[2], 0.5 * (3.0 + 5.0))  # cutoff residual at the row end
    out = tally.depth_dose(_row(-2.0, 1.0, 3.0), 0.0, 1.0, 10)  # the part above the slab goes to bin 0
    np.testing.assert_allclose(out[:2], [3.0, 0.0])


@pytest.mark.parametrize("case", [SI_30KEV_STRAGGLE, MULTILAYER], ids=lambda c: c.name)
def test_energy_balance_of_oracle_histories(oracle, case):
    lt, lut = case.tables()
    pos, dirs, E, E_cut = case.state()
    keys = oracle.stream_keys(case.seed, case.Ne)
    rows, cnt, code, st = oracle.run_transport(lt, lut, keys, pos, dirs, E, E_cut, cap=4096,
                                               energy_model=1, max_dE_frac=case.max_dE_frac,
                                               straggle_on=case.straggle)
    r = oracle.compact(rows, cnt, 4096)
    z_total = float(lt.z_total)
    dose = tally.depth_dose(r, 0.0, z_total / 64, 64)
    escaped = st["E_keV"][code != 0].sum()
    np.testing.assert_allclose(dose.sum() + escaped, case.Ne * case.E0, rtol=1e-12)
    assert np.all(dose >= 0.0) and dose[0] > 0.0
