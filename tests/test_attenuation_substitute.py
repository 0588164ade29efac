"""The Ge attenuation substitute (CPU): Se's XCOM table mapped by the
photoelectric law with the energy axis scaled by the K-edge ratio.

There is no Ge table to pin against (the reference reads EPDL, absent; its
XCOM fixtures hold N, C, Al, Si, Se, W, Pb), so the test holds what the
construction fixes: the K edge at 11.10 keV with Se's jump ratio, the
Z^4 / E^3 relation to Se away from edges, and Se itself unchanged."""

import numpy as np

from pyrite_amd.spectrum import attenuation as att


def test_ge_k_edge_sits_at_its_binding_energy():
    below = att.sigma_photo_cm2("Ge", np.array([att.GE_K_EDGE_EV * (1 - 1e-6)]))[0]
    above = att.sigma_photo_cm2("Ge", np.array([att.GE_K_EDGE_EV * (1 + 1e-6)]))[0]
    se_below = att.sigma_photo_cm2("Se", np.array([att.SE_K_EDGE_EV * (1 - 1e-6)]))[0]
    se_above = att.sigma_photo_cm2("Se", np.array([att.SE_K_EDGE_EV * (1 + 1e-6)]))[0]
    assert above / below > 5.0
    np.testing.assert_allclose(above / below, se_above / se_below, rtol=1e-4)
    # no edge left at Se's energy
    a, b = att.sigma_photo_cm2("Ge", np.array([att.SE_K_EDGE_EV * 0.999, att.SE_K_EDGE_EV * 1.001]))
    assert abs(b / a - 1.0) < 0.01


def test_ge_follows_the_photoelectric_law_from_se():
    s = att.SE_K_EDGE_EV / att.GE_K_EDGE_EV
    E = np.array([20e3, 30e3, 50e3, 80e3])
    ge = att.sigma_photo_cm2("Ge", E)
    se = att.sigma_photo_cm2("Se", E * s)
    np.testing.assert_allclose(ge, (32 / 34) ** 4 * s**3 * se, rtol=1e-12)
    # and against Se at the same energy: roughly (32/34)^4, within the edge shift
    r = att.sigma_photo_cm2("Ge", E) / att.sigma_photo_cm2("Se", E)
    assert np.all((r > 0.6) & (r < 0.9))
