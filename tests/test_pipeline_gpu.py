"""End to end on the device: transport -> lines + brem, against the oracle case (GPU).

The same histories (global ids) run through the HIP pipeline and the FP64 C
oracle.  Spectra are compared bin by bin with |d| <= 3 sqrt(s_a^2 + s_b^2),
sigma from 20 independent history batches (docs/computation/statistical-methods.md
quadrature rule); about 0.3% of bins trip by chance, so the test bounds the
tripped fraction.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_case_spectra_within_3_sigma(device):
    import torch

    from oracle import case as ocase
    from pyrite_amd.pipeline import CaseConfig, CasePipeline

    n_b, per = 20, 150
    cfg = CaseConfig(Ne=per, n_side=3)
    pipe = CasePipeline(cfg)
    g_lines, g_brem, o_lines, o_brem = [], [], [], []
    for b in range(n_b):
        pipe.spec_lines.zero_()
        pipe.spec_brem.zero_()
        pipe.step(b, 1)
        lines, brem = pipe.spectra()
        g_lines.append(lines)
        g_brem.append(brem)
        ol, ob, _ = ocase.run_case(pipe, b * per, per)
        o_lines.append(ol / per)
        o_brem.append(ob / (4 * np.pi * per) * pipe.omega)
    torch.cuda.synchronize()
    for g, o, name in ((g_lines, o_lines, "lines"), (g_brem, o_brem, "brem")):
        g, o = np.array(g), np.array(o)
        mg, mo = g.mean(0), o.mean(0)
        s = np.sqrt(g.var(0, ddof=1) / n_b + o.var(0, ddof=1) / n_b)
        live = s > 0
        trip = np.abs(mg - mo)[live] > 3.0 * s[live] + 1e-9 * np.abs(mo[live])
        assert live.sum() > 500, name
        assert trip.mean() < 0.02, (name, trip.mean())
        # and the totals agree within 3 sigma of the total
        tg, to = g.sum(1), o.sum(1)
        st = np.sqrt(tg.var(ddof=1) / n_b + to.var(ddof=1) / n_b)
        assert abs(tg.mean() - to.mean()) <= 3.0 * st + 1e-9 * abs(to.mean()), name
