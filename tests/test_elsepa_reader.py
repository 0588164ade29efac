"""ELSEPA dcs reader against the reference's committed vendor output
(tests/data/xsgen/elsepa/test-run-output/dcs_1p000e03.dat: Hg, 1 keV), with
the header's integrated cross sections re-derived from the body (CPU)."""

from pathlib import Path

import numpy as np
import pytest

from pyrite_amd.tables.elsepa import angular_cdf, read_dcs, stack_panels

FIXTURE = Path(__file__).parent / "golden" / "elsepa_dcs_1p000e03.dat"


def test_reads_the_vendor_panel():
    p = read_dcs(FIXTURE)
    assert p.energy_eV == pytest.approx(1000.0) and p.Z == 80
    assert p.total_elastic_cm2 == pytest.approx(2.47469e-16)
    assert p.transport1_cm2 == pytest.approx(5.54005e-17)
    assert p.transport2_cm2 == pytest.approx(5.07959e-17)
    assert p.absorption_cm2 == pytest.approx(9.36109e-17)
    assert p.theta_deg.shape == (606,)
    assert p.theta_deg[[0, -1]].tolist() == [0.0, 180.0] and p.mu[[0, -1]].tolist() == [0.0, 1.0]
    assert read_dcs(FIXTURE.read_bytes()).transport2_cm2 == p.transport2_cm2
    # mu is (1 - cos theta)/2 of the theta column
    np.testing.assert_allclose(p.mu, 0.5 * (1.0 - np.cos(np.radians(p.theta_deg))), rtol=2e-5, atol=1e-16)
    # a0^2 and cm^2 columns differ by the Bohr radius squared
    np.testing.assert_allclose(p.dcs_cm2_sr / p.dcs_a0_2_sr, 0.529177210903e-8**2, rtol=1e-4)


def test_body_integrates_to_the_header_cross_sections():
    """sigma = 4 pi int DCS dmu; sigma_1 = 4 pi int 2 mu DCS dmu (1 - cos = 2 mu);
    sigma_2 = 4 pi int 6 mu (1 - mu) DCS dmu (1 - P2(cos) = 3/2 sin^2)."""
    p = read_dcs(FIXTURE)
    quad = lambda f: 4.0 * np.pi * np.trapezoid(f * p.dcs_cm2_sr, p.mu)  # noqa: E731
    assert quad(np.ones_like(p.mu)) == pytest.approx(p.total_elastic_cm2, rel=3e-4)
    assert quad(2.0 * p.mu) == pytest.approx(p.transport1_cm2, rel=3e-4)
    assert quad(6.0 * p.mu * (1.0 - p.mu)) == pytest.approx(p.transport2_cm2, rel=3e-4)


def test_stacked_table_and_cdf():
    p = read_dcs(FIXTURE)
    t = stack_panels([p], energies_eV=[1000.0])
    cdf = t["angular_cdf"][0]
    assert cdf[0] == 0.0 and cdf[-1] == 1.0 and np.all(np.diff(cdf) >= 0.0)
    np.testing.assert_array_equal(cdf, angular_cdf(p))
    with pytest.raises(ValueError, match="requested"):
        stack_panels([p], energies_eV=[1010.0])


def test_malformed_panels_are_refused():
    text = FIXTURE.read_text()
    with pytest.raises(ValueError, match="fewer than two"):
        read_dcs("# Kinetic energy = 1.0E3 eV\n# Total elastic cross section = 1 cm**2\n"
                 "# 1st transport cross section = 1 cm**2\n# 2nd transport cross section = 1 cm**2\n")
    with pytest.raises(ValueError, match="columns"):
        read_dcs(text + "  1.0 2.0 3.0\n")
    with pytest.raises(ValueError, match="header"):
        read_dcs("\n".join(l for l in text.splitlines() if "Total elastic" not in l) + "\n")
