"""Crystal catalog, reciprocal vectors and the chi_g / U_g coupling tables.

Restates ``materials/crystal.py`` of the reference for the crystals the line
configurations use:

* reciprocal basis ``2 pi (a2 x a3, a3 x a1, a1 x a2)/V``, ``g = hkl @ B``
  (``crystal.py:100-160``);
* Debye-Waller amplitude ``exp(-B (g/4pi)^2)`` (``:302-312``);
* ``chi_g = -r_e lambda^2/(pi V) S(g)``, ``S = sum F_i e^{i 2 pi hkl.R_i} DWF``
  (``:333-420``), emission convention ``conj(chi(-hkl))`` (``:730-770``);
* ``U_g/m = 4 pi e^2 DWF/g^2 sum (Z - f0) e^{i 2 pi hkl.R}/(V m c^2)`` (``:547-604``);
* decrement ``delta = 1 - Re sqrt(1 + chi_0)`` (``:520-545``).

Offline substitutes (the reference reads them through ``xraydb``, absent here):
``f0`` from the published Waasmaier-Kirfel (1995) five-Gaussian fits (C, Si,
Ge) and the International Tables Cromer-Mann fit (W); ``f''`` from the XCOM
photoabsorption cross section by the optical theorem
``sigma_pe = 2 r_e lambda f''``; ``f' = 0``; no bonding-charge correction.
These tables are inputs of the scoring kernels: the kernels and the oracle
read the same arrays, so the parity checks do not depend on them.
"""

from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from ..transport.elements import TRANSPORT_ELEMENTS
from . import attenuation

HC_EV_ANG = 12398.4198
HBARC_EV_ANG = 1973.269804
ALPHA_FS = 1.0 / 137.035999
M_E_EV = 510998.95
R_E_ANG = 2.8179403e-5
E2_EV_ANG = ALPHA_FS * HBARC_EV_ANG

# f0(s) = sum_i a_i exp(-b_i s^2) + c,  s = sin(theta)/lambda = g/(4 pi) [1/A]
F0_COEFFS = {
    # Waasmaier & Kirfel, Acta Cryst. A51, 416 (1995)
    "C": ([2.657506, 1.078079, 1.490909, -4.241070, 0.713791],
          [14.780758, 0.776775, 42.086843, -0.000294, 0.239535], 4.297983),
    "Si": ([5.275329, 3.191038, 1.511514, 1.356849, 2.519114],
           [2.631338, 33.730728, 0.081119, 86.288640, 1.170087], 0.145073),
    "Ge": ([16.540613, 1.567900, 3.727829, 3.345098, 6.785079],
           [2.866618, 0.012198, 13.432163, 58.866046, 0.210974], 0.018726),
    # Cromer-Mann, International Tables for Crystallography C, Table 6.1.1.4
    "W": ([29.0818, 15.43, 14.4327, 5.11982], [1.72029, 9.2259, 0.321703, 57.056], 9.8875),
}


def f0(element, g):
    """Atomic form factor ``f0(g)`` [electrons], ``g = 2 pi / d``."""
    a, b, c = F0_COEFFS[element]
    s2 = (np.asarray(g, dtype=float) / (4.0 * np.pi)) ** 2
    return sum(ai * np.exp(-bi * s2) for ai, bi in zip(a, b)) + c


def f2(element, E_eV):
    """``f''(E)`` from the photoabsorption cross section (optical theorem)."""
    lam_cm = HC_EV_ANG / np.asarray(E_eV, dtype=float) * 1.0e-8
    return attenuation.sigma_photo_cm2(element, E_eV) / (2.0 * R_E_ANG * 1.0e-8 * lam_cm)


@dataclass
class Crystal:
    key: str
    a: float
    b: float
    c: float
    alpha: float
    beta: float
    gamma: float
    basis: list  # [(element, (x, y, z) fractional)]
    B_ang2: float
    surface_hkl: tuple
    lattice_rows: np.ndarray = field(init=False)
    recip: np.ndarray = field(init=False)
    V_cell: float = field(init=False)

    def __post_init__(self):
        al, be, ga = np.deg2rad([self.alpha, self.beta, self.gamma])
        a1 = np.array([self.a, 0.0, 0.0])
        a2 = np.array([self.b * np.cos(ga), self.b * np.sin(ga), 0.0])
        cx = self.c * np.cos(be)
        cy = self.c * (np.cos(al) - np.cos(be) * np.cos(ga)) / np.sin(ga)
        cz = np.sqrt(max(self.c**2 - cx**2 - cy**2, 0.0))
        a3 = np.array([cx, cy, cz])
        self.lattice_rows = np.array([a1, a2, a3])
        self.V_cell = float(np.dot(a1, np.cross(a2, a3)))
        self.recip = 2.0 * np.pi * np.array([np.cross(a2, a3), np.cross(a3, a1), np.cross(a1, a2)]) / self.V_cell

    @property
    def composition(self):
        """``[(element, atoms per A^3)]`` of the cell."""
        counts: dict = {}
        for el, _ in self.basis:
            counts[el] = counts.get(el, 0) + 1
        return [(el, n / self.V_cell) for el, n in counts.items()]

    def g_vector(self, hkl):
        g = np.asarray(hkl, dtype=float) @ self.recip
        return g, float(np.linalg.norm(g))

    def orientation(self):
        """Minimal rotation taking the ``surface_hkl`` normal onto +z (``geometry._orientation_R``)."""
        from .geometry import rotation_between

        g, gm = self.g_vector(self.surface_hkl)
        return rotation_between(g / gm, np.array([0.0, 0.0, 1.0]))


CRYSTALS = {
    "hopg": Crystal("hopg", 2.461, 2.461, 6.711, 90.0, 90.0, 120.0,
                    [("C", (0.0, 0.0, 0.0)), ("C", (1 / 3, 2 / 3, 0.0)),
                     ("C", (0.0, 0.0, 0.5)), ("C", (2 / 3, 1 / 3, 0.5))], 1.30, (0, 0, 1)),
    "silicon": Crystal("silicon", 5.4309, 5.4309, 5.4309, 90.0, 90.0, 90.0,
                       [("Si", p) for p in ((0, 0, 0), (0.25, 0.25, 0.25), (0, 0.5, 0.5), (0.25, 0.75, 0.75),
                                            (0.5, 0, 0.5), (0.75, 0.25, 0.75), (0.5, 0.5, 0), (0.75, 0.75, 0.25))],
                       0.4632, (1, 1, 1)),
    "ge": Crystal("ge", 5.65754, 5.65754, 5.65754, 90.0, 90.0, 90.0,
                  [("Ge", p) for p in ((0, 0, 0), (0, 0.5, 0.5), (0.25, 0.25, 0.25), (0.25, 0.75, 0.75),
                                       (0.5, 0, 0.5), (0.5, 0.5, 0), (0.75, 0.25, 0.75), (0.75, 0.75, 0.25))],
                  0.6, (1, 1, 1)),
    "tungsten": Crystal("tungsten", 3.1648, 3.1648, 3.1648, 90.0, 90.0, 90.0,
                        [("W", (0.0, 0.0, 0.0)), ("W", (0.5, 0.5, 0.5))], 0.16, (1, 1, 0)),
}


def structure_factor_abs(crystal: Crystal, hkl):
    """|sum_j exp(2 pi i h.x_j)| of the basis (unit form factors): the
    extinction rule of the lattice, not a scattering strength."""
    h = np.asarray(hkl, dtype=float)
    ph = [2.0 * np.pi * float(np.dot(h, x)) for _, x in crystal.basis]
    return float(abs(sum(np.exp(1j * p) for p in ph)))


def allowed_reflections(crystal_key, n, h_max=6):
    """The ``n`` shortest reciprocal vectors with a non-vanishing lattice sum,
    ordered by (|g|, hkl) so the choice is deterministic."""
    cr = CRYSTALS[crystal_key]
    rng = range(-h_max, h_max + 1)
    cand = []
    for h in rng:
        for k in rng:
            for l in rng:
                if (h, k, l) == (0, 0, 0):
                    continue
                if structure_factor_abs(cr, (h, k, l)) < 1e-6:
                    continue
                cand.append((round(cr.g_vector((h, k, l))[1], 9), (h, k, l)))
    cand.sort()
    return [hkl for _, hkl in cand[:n]]


def debye_waller(g, B_ang2):
    return np.exp(-B_ang2 * (np.asarray(g) / (4.0 * np.pi)) ** 2)


def coupling_tables(crystal: Crystal, hkl_list, E_eV):
    """Emission couplings ``(chi_re, chi_im, u_re, u_im)`` ``[n_hkl, n_E]`` float64.

    ``conj(chi(-hkl))`` and ``U(hkl)`` (``crystal.py:730-770``); ``U`` is
    dimensionless ``U_g / m c^2``.
    """
    E = np.asarray(E_eV, dtype=float)
    lam = HC_EV_ANG / E
    chi_scale = -R_E_ANG * lam**2 / (np.pi * crystal.V_cell)
    u_scale = 4.0 * np.pi * E2_EV_ANG / crystal.V_cell / M_E_EV
    elements = list(dict.fromkeys(el for el, _ in crystal.basis))
    fpp = {el: f2(el, E) for el in elements}
    out = [np.empty((len(hkl_list), E.size)) for _ in range(4)]
    for i, hkl in enumerate(hkl_list):
        mate = -np.asarray(hkl, dtype=float)
        _, g = crystal.g_vector(mate)
        dwf = debye_waller(g, crystal.B_ang2)
        S = np.zeros(E.size, dtype=complex)
        acc = 0.0 + 0.0j
        for el, R in crystal.basis:
            ph = np.exp(1j * 2.0 * np.pi * np.dot(mate, R))
            S = S + ph * (f0(el, g) + 1j * fpp[el])
            acc += ph * (TRANSPORT_ELEMENTS[el]["Z"] - f0(el, g))
        chi = chi_scale * dwf * S
        U = u_scale * dwf / g**2 * acc
        out[0][i] = chi.real
        out[1][i] = -chi.imag  # conj(chi(-hkl))
        out[2][i] = U.real
        out[3][i] = -U.imag * np.ones(E.size)  # conj(U(-hkl)) = U(hkl)
    return tuple(out)


def decrement(crystal: Crystal, E_eV):
    """``delta = 1 - Re n``, ``n = sqrt(1 + chi_0)`` (``crystal.py:520-545``)."""
    E = np.asarray(E_eV, dtype=float)
    lam = HC_EV_ANG / E
    S0 = np.zeros(E.size, dtype=complex)
    for el, _ in crystal.basis:
        S0 = S0 + TRANSPORT_ELEMENTS[el]["Z"] + 1j * f2(el, E)
    chi0 = -R_E_ANG * lam**2 / (np.pi * crystal.V_cell) * S0
    return 1.0 - np.sqrt(1.0 + chi0).real


def polarization_pairs(k_hat, g_rows):
    """``(e_s, e_p)`` per reflection row (``cxr/_polarization.py:17-40``)."""
    k_hat = np.asarray(k_hat, dtype=float)
    g_rows = np.asarray(g_rows, dtype=float).reshape(-1, 3)
    es = np.empty_like(g_rows)
    ep = np.empty_like(g_rows)
    for i, g in enumerate(g_rows):
        n_plane = np.cross(k_hat, g)
        npl = np.linalg.norm(n_plane)
        if npl < 1e-12:
            n_plane = np.cross(k_hat, np.array([0.0, 1.0, 0.0]))
            npl = np.linalg.norm(n_plane)
        e_s = n_plane / npl
        e_p = np.cross(e_s, k_hat)
        es[i] = e_s
        ep[i] = e_p / np.linalg.norm(e_p)
    return es, ep


def tabulation_grid(lo_eV, hi_eV, ratio=1.0 + 2.0e-3):
    """Geometric tabulation mesh over ``[lo, hi]`` eV (``_tabulation_mesh``-style)."""
    lo = max(float(lo_eV), 1.0)
    n = int(np.ceil(np.log(hi_eV / lo) / np.log(ratio))) + 1
    return lo * ratio ** np.arange(n)
