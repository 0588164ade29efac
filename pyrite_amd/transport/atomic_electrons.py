"""Angular deflection by atomic electrons folded into the elastic rate.

The reference's ``atomic_electron_deflection="kawrakow"``
(``src/pyrite/montecarlo/transport/atomic_electrons.py``; Kawrakow 1997,
EGSnrc PIRS-701 Sec. 4.7): the nucleus scatters as ``Z^2``, and the part of the
``Z`` atomic electrons' deflection that transport does not resolve explicitly
is folded in by ``Z^2 -> Z (Z + xi)``, i.e. each element's elastic rate times
``1 + xi / Z`` with the elastic angular shape kept.

    xi(T, T_c) = xi_0 (1 - min(g_M / (xi_0 g_R), 1)),   xi_0 = 1,

``g_M(tau, tau_c)``: the ``sin^2 theta`` moment of the primary in Moller
collisions with energy transfer above ``T_c`` (per atomic electron);
``g_R(eta) = (1 + 2 eta) ln(1 + 1/eta) - 2``: the same moment of the
screened-Rutherford nucleus per ``Z^2`` at the Moliere screening ``eta`` of
the layer (EGSnrc Eqs. 4.7.6-4.7.8).  Under continuous stopping nothing is
explicit (``T_c = inf``) and ``xi = 1``.

``g_M`` is evaluated here by Gauss-Legendre quadrature of the Moller cross
section times the primary's ``sin^2 theta`` over ``eps = W/T`` in
``[tau_c/tau, 1/2]``, on ``ln eps`` (the integrand's ``1/eps`` peak becomes
smooth there); tests hold it to the published closed form.
"""

from __future__ import annotations

import numpy as np

from .elements import TRANSPORT_ELEMENTS

MC2_EV = 510998.95
ALPHA2 = (1.0 / 137.035999084) ** 2  # fine-structure constant squared
# EGSnrc Eq. 4.7.6: chi_cc^2 / b_c = 0.1569 MeV^2 / 7821.6 * exp((Z_X - Z_E) / Z_S)
CHI2_OVER_B_MEV2 = 0.1569 / 7821.6
XI0 = 1.0

_GL_X, _GL_W = np.polynomial.legendre.leggauss(96)


def moliere_screening_eta(composition, tau):
    """Moliere screening ``eta = chi_a^2 / 4`` of a layer at ``tau = T / mc^2``
    (only the atom fractions of ``composition = [(element, n), ...]`` enter)."""
    Z = np.array([float(TRANSPORT_ELEMENTS[el]["Z"]) for el, _ in composition])
    frac = np.array([float(n) for _, n in composition])
    frac = frac / frac.sum()
    w = frac * Z * (Z + XI0)
    z_s = w.sum()
    z_e = (w * (-2.0 / 3.0) * np.log(Z)).sum()
    z_x = (w * np.log1p(3.34 * ALPHA2 * Z * Z)).sum()
    m_MeV = MC2_EV * 1e-6
    tau = np.asarray(tau, dtype=np.float64)
    return CHI2_OVER_B_MEV2 * np.exp((z_x - z_e) / z_s) / (4.0 * m_MeV * m_MeV * tau * (tau + 2.0))


def screened_rutherford_g(eta):
    """Half the ``sin^2`` moment of the screened-Rutherford law ``1/(1 - mu + 2 eta)^2``."""
    eta = np.asarray(eta, dtype=np.float64)
    return (1.0 + 2.0 * eta) * np.log1p(1.0 / eta) - 2.0


def _moller_sin2(tau, eps):
    """``(tau + 2) eps / (tau + 2 - tau eps)`` times the Moller bracket: the
    integrand of ``g_M`` in ``eps`` (units where the nucleus's moment is ``2 g_R``)."""
    c = tau / (tau + 1.0)
    d = (2.0 * tau + 1.0) / ((tau + 1.0) * (tau + 1.0))
    om = 1.0 - eps
    bracket = 1.0 / (eps * eps) + 1.0 / (om * om) + c * c - d / (eps * om)
    return (tau + 2.0) * eps / (tau + 2.0 - tau * eps) * bracket


def moller_g(tau, tau_c):
    """``g_M``: zero where no Moller collision transfers more than ``tau_c``
    (``tau <= 2 tau_c``)."""
    tau = np.atleast_1d(np.asarray(tau, dtype=np.float64))
    out = np.zeros_like(tau)
    live = tau > 2.0 * tau_c
    t = tau[live][:, None]
    lo = np.log(tau_c / t)
    hi = np.log(0.5)
    half = 0.5 * (hi - lo)
    s = lo + half * (_GL_X[None, :] + 1.0)  # ln eps at the nodes
    eps = np.exp(s)
    out[live] = (half * (_GL_W[None, :] * _moller_sin2(t, eps) * eps)).sum(axis=1)
    return out


def atomic_electron_xi(energy_eV, composition, cutoff_eV=None):
    """``xi(T, T_c)`` at ``energy_eV`` for one layer; ``cutoff_eV=None``
    (continuous stopping) gives ``xi_0`` everywhere."""
    energy_eV = np.asarray(energy_eV, dtype=np.float64)
    if cutoff_eV is None:
        return np.full(energy_eV.shape, XI0)
    tau = energy_eV / MC2_EV
    ratio = moller_g(tau.ravel(), float(cutoff_eV) / MC2_EV).reshape(tau.shape) / screened_rutherford_g(
        moliere_screening_eta(composition, tau))
    return XI0 * (1.0 - np.minimum(ratio / XI0, 1.0))


def elastic_rate_scale(Z, xi):
    """Per-element factor ``1 + xi / Z`` on the elastic rate."""
    return 1.0 + np.asarray(xi, dtype=np.float64) / np.asarray(Z, dtype=np.float64)
