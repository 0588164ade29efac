"""Photon attenuation from NIST XCOM tables shipped with the package.

The reference reads the EPDL2025 total cross section (a fetched table,
``materials/attenuation.py:194-260``) and validates it against NIST XCOM at
0.5% (C, Si) and 3% (W) (``tests/materials/test_photon_cross_sections.py:33``).
Offline, this package uses the XCOM total-with-coherent attenuation directly,
log-log interpolated between the tabulated nodes with the edge rows kept as
one-sided values.  Ge has no XCOM table here: its cross section is the Se
table mapped by the photoelectric law sigma ~ Z^4 / E^3 with the energy axis
scaled by the ratio of the two K-edge energies (Se 12.660 keV, Ge 11.1031
keV; the table's Se edge node and Bearden & Burr 1967), sigma_Ge(E) =
(32/34)^4 s^3 sigma_Se(s E),
s = E_K(Se)/E_K(Ge), so Ge's K edge sits at 11.10 keV -- a stated substitute,
not a pinned value.

``mu = n sigma`` with ``sigma = (mu/rho) A / N_A`` [cm^2 -> A^2].
"""

from __future__ import annotations

from functools import lru_cache
from pathlib import Path

import numpy as np

from ..transport.elements import TRANSPORT_ELEMENTS

DATA = Path(__file__).resolve().parent.parent / "data" / "xcom"
N_A = 6.02214076e23
# numeric columns of a row: E_MeV, coherent, incoherent, photoelectric, pair_nuclear,
# pair_electron, total_with_coh, total_without_coh
COL_PHOTO, COL_TOTAL = 3, 6
# K edges: Se at the XCOM table's edge node (12.660 keV), Ge 11.1031 keV (Bearden & Burr 1967)
SE_K_EDGE_EV, GE_K_EDGE_EV = 12660.0, 11103.1


@lru_cache(maxsize=None)
def _table(element: str):
    """``(E_eV, mu_rho_total, mu_rho_photo)`` rows in file order [cm^2/g]."""
    scale = 1.0
    e_scale = 1.0
    name = element
    if element == "Ge":
        name = "Se"
        e_scale = GE_K_EDGE_EV / SE_K_EDGE_EV  # node energies: Se edges land on Ge's
        s = 1.0 / e_scale
        scale = (32.0 / 34.0) ** 4 * s**3 * TRANSPORT_ELEMENTS["Se"]["A"] / TRANSPORT_ELEMENTS["Ge"]["A"]
    path = DATA / f"xcom_{name}.txt"
    if not path.exists():
        raise KeyError(f"no XCOM attenuation table for {element}")
    rows = []
    for line in path.read_text().splitlines():
        if not line.strip() or line.startswith("#"):
            continue
        parts = line.split()
        vals = [float(x) for x in parts[1:]]
        rows.append(vals)
    a = np.asarray(rows)
    E = a[:, 0] * 1.0e6 * e_scale
    return E, a[:, COL_TOTAL] * scale, a[:, COL_PHOTO] * scale


def _loglog(E_nodes, v_nodes, E):
    # clamped to the tabulated range (1 keV - 100 GeV): no power-law extrapolation
    E = np.clip(np.asarray(E, dtype=float), E_nodes[0], E_nodes[-1])
    i = np.searchsorted(E_nodes, E, side="right") - 1
    i = np.clip(i, 0, E_nodes.size - 2)
    x0, x1 = E_nodes[i], E_nodes[i + 1]
    y0, y1 = v_nodes[i], v_nodes[i + 1]
    same = x1 <= x0
    t = np.where(same, 0.0, np.log(np.maximum(E, 1e-300) / x0) / np.where(same, 1.0, np.log(x1 / x0)))
    return np.exp(np.log(y0) + t * (np.log(y1) - np.log(y0)))


def sigma_total_ang2(element: str, E_eV) -> np.ndarray:
    """Per-atom total cross section [A^2]."""
    E_n, tot, _ = _table(element)
    A = TRANSPORT_ELEMENTS[element]["A"]
    return _loglog(E_n, tot, E_eV) * A / N_A * 1.0e16


def sigma_photo_cm2(element: str, E_eV) -> np.ndarray:
    """Per-atom photoabsorption cross section [cm^2]."""
    E_n, _, ph = _table(element)
    A = TRANSPORT_ELEMENTS[element]["A"]
    return _loglog(E_n, ph, E_eV) * A / N_A


def mu_total_inv_ang(composition, E_eV) -> np.ndarray:
    """``sum_i n_i sigma_i(E)`` [1/A] for ``[(element, n_per_A3), ...]``."""
    E = np.asarray(E_eV, dtype=float)
    mu = np.zeros_like(E)
    for el, n in composition:
        mu = mu + float(n) * sigma_total_ang2(el, E)
    return mu


def elemental_log_mu(composition, E_eV) -> np.ndarray:
    """``log(n_i sigma_i)`` rows ``[n_el, n_E]`` (the line route sums exp of them)."""
    return np.log(np.stack([float(n) * sigma_total_ang2(el, E_eV) for el, n in composition]))
