"""Depth-dose tally on the device (GPU): the kernel against the oracle's
restatement on the same rows, slot-major against compacted rows, and the exact
energy balance sum(dose) + sum(escaped E) = Ne E0 on device histories."""

import numpy as np
import pytest

from oracle import tally
from tests.transport_cases import HOPG_5MEV, MULTILAYER, SI_30KEV_STRAGGLE

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", [SI_30KEV_STRAGGLE, MULTILAYER, HOPG_5MEV], ids=lambda c: c.name)
def test_device_dose_matches_oracle_and_closes_the_energy_balance(device, case):
    import torch

    from pyrite_amd.transport.api import simulate_trajectories
    from pyrite_amd.transport.tally import depth_dose

    out = simulate_trajectories(E0_keV=case.E0, Ne=case.Ne, layers=case.layers, seed=case.seed,
                                straggling=case.straggle, max_dE_frac=case.max_dE_frac,
                                beam_dir=case.beam_dir, keep_segments_on_device=True)
    rows = {"seg_mid": out["r_mid"], "seg_dir": out["v_hat"], "seg_len": out["L_ang"], "seg_E": out["E_start_keV"],
            "seg_E_end": out["E_end_keV"], "seg_event": out["event_kind"]}
    z_total = float(out["thickness_ang"])
    nb = 50
    dose = torch.zeros(nb, dtype=torch.float64, device=device)
    depth_dose(dose, 0.0, z_total / nb, rows, int(out["L_ang"].shape[0]))
    got = dose.cpu().numpy()
    host = {k: v.cpu().numpy() for k, v in rows.items()}
    want = tally.depth_dose(host, 0.0, z_total / nb, nb)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12 * want.max())
    code = out["exit_code"].cpu().numpy()
    escaped = out["final_E_keV"].cpu().numpy()[code != 0].sum()
    np.testing.assert_allclose(got.sum() + escaped, case.Ne * case.E0, rtol=1e-12)


def test_slot_rows_give_the_compacted_dose(device):
    import torch

    from pyrite_amd.pipeline import CaseConfig, CasePipeline
    from pyrite_amd.transport.tally import depth_dose

    pipe = CasePipeline(CaseConfig(Ne=3000, depth_bins=40))
    res = pipe.transport(0, compact=False)
    pipe.score_dose(res, 3000)
    slot = pipe.depth_dose()
    res2 = pipe.transport(0, compact=True)
    dose = torch.zeros(40, dtype=torch.float64, device=device)
    depth_dose(dose, 0.0, pipe.cfg.thickness_ang / 40, res2.rows, res2.n_segments, scale=1.0 / 3000)
    np.testing.assert_allclose(slot, dose.cpu().numpy(), rtol=1e-12, atol=1e-15)
    assert slot.sum() > 0
