"""Campaign description for the work-queue kernel (CPU)."""

import numpy as np

from pyrite_amd.queue import MATERIALS, campaign


def test_campaign_ranges_and_determinism():
    c = campaign(256, seed=2024)
    assert len(c) == 256
    Ne = np.array([q.Ne for q in c])
    th = np.array([q.thickness_ang for q in c])
    E = np.array([q.E0_keV for q in c])
    tilt = np.array([q.tilt_deg for q in c])
    assert Ne.min() >= 1e4 and Ne.max() <= 1e7
    assert th.min() >= 1e5 and th.max() <= 2e7
    assert E.min() >= 1000 and E.max() <= 30000
    assert tilt.min() >= 0 and tilt.max() <= 3
    assert {q.material for q in c} == set(MATERIALS)
    assert [q.seed for q in campaign(256, seed=2024)] == [q.seed for q in c]
    # log-uniform history counts: mean near (1e7 - 1e4) / ln(1e3)
    assert 0.6e6 < Ne.mean() < 2.5e6
