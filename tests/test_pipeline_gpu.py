"""Overlapped batches (transport of batch s+1 while batch s is scored) give
the tallies of back-to-back steps bit for bit (GPU)."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("route", ["dense", "window"])
def test_overlapped_steps_equal_sequential_steps(device, route):
    from pyrite_amd.pipeline import CaseConfig, CasePipeline

    cfg = CaseConfig(Ne=6000, line_route=route, E0_keV=3000.0, thickness_ang=3.0e5)
    a = CasePipeline(cfg)
    for b in range(4):
        a.step(b, 4)
    la, ba = a.spectra()
    p = CasePipeline(cfg)
    p.run_steps(0, 4, overlap=True)
    lb, bb = p.spectra()
    assert p.histories == a.histories and p.segments == a.segments
    assert la.sum() > 0 and ba.sum() > 0
    if route == "dense":
        np.testing.assert_array_equal(lb, la)
        np.testing.assert_array_equal(bb, ba)
    else:  # LDS FP64 atomics in the windowed route: summation order varies
        np.testing.assert_allclose(lb, la, rtol=1e-12, atol=1e-15 * la.max())
        np.testing.assert_array_equal(bb, ba)
