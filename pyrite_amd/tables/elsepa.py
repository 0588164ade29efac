"""Reader of ELSEPA elastic differential cross sections (``dcs_*.dat``).

The reference generates its elastic tables by running ELSEPA and reading each
energy panel back (``src/pyrite/xsgen/elsepa/parse.py:48-173``): the header
carries the kinetic energy and the integrated cross sections, the body one row
per polar angle with ``theta [deg], mu = (1 - cos theta)/2, DCS [cm^2/sr],
DCS [a0^2/sr], Sherman function, relative error``.  Panels on one angular
grid stack into the native-grid table the reference stores, with the angular
CDF in ``mu`` by the trapezoid rule on that grid (``dOmega = 4 pi dmu``).

This module reads the format; no transport kernel of this backend consumes
tabulated elastic cross sections yet (the step runs screened Rutherford).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

# header quantities: key phrase -> (field, required)
_HEADER = {
    "Kinetic energy": ("energy_eV", True),
    "Total elastic cross section": ("total_elastic_cm2", True),
    "1st transport cross section": ("transport1_cm2", True),
    "2nd transport cross section": ("transport2_cm2", True),
    "Absorption cross section": ("absorption_cm2", False),
}


@dataclass(frozen=True)
class ElsepaPanel:
    energy_eV: float
    Z: int | None
    theta_deg: np.ndarray
    mu: np.ndarray
    dcs_cm2_sr: np.ndarray
    dcs_a0_2_sr: np.ndarray
    sherman: np.ndarray
    relative_error: np.ndarray
    total_elastic_cm2: float
    transport1_cm2: float
    transport2_cm2: float
    absorption_cm2: float


def _number(token: str) -> float:
    return float(token.replace("D", "E").replace("d", "e"))


def _header_value(line: str) -> float:
    """The first number after ``=`` on a header line (its first unit)."""
    rhs = line.split("=", 1)[1].split()
    return _number(rhs[0])


def read_dcs(source) -> ElsepaPanel:
    """One ``dcs_*.dat`` panel from a path, bytes or text."""
    if isinstance(source, (bytes, bytearray)):
        text = bytes(source).decode("ascii")
    elif isinstance(source, str) and "\n" in source:
        text = source
    else:
        with open(source, encoding="ascii") as f:
            text = f.read()
    found: dict[str, float] = {}
    Z = None
    rows = []
    for raw in text.splitlines():
        line = raw.strip()
        if not line:
            continue
        if line.startswith("#"):
            body = line.lstrip("#").strip()
            for phrase, (field, _) in _HEADER.items():
                if body.startswith(phrase) and "=" in body and field not in found:
                    found[field] = _header_value(body)
            if body.startswith("Element:") and "Z =" in body:
                Z = int(body.split("Z =", 1)[1].split(",")[0])
            continue
        cols = line.split()
        if len(cols) != 6:
            raise ValueError(f"ELSEPA row with {len(cols)} columns (6 expected): {raw!r}")
        try:
            rows.append([_number(c) for c in cols])
        except ValueError as exc:
            raise ValueError(f"ELSEPA row is not numeric: {raw!r}") from exc
    for field, required in _HEADER.values():
        if field not in found:
            if required:
                raise ValueError(f"ELSEPA header has no {field}")
            found[field] = 0.0  # written only when an absorption potential is on
    if len(rows) < 2:
        raise ValueError("ELSEPA panel has fewer than two angles")
    a = np.asarray(rows, dtype=np.float64)
    if not np.isfinite(a).all():
        raise ValueError("ELSEPA panel holds non-finite values")
    theta, mu, dcs, dcs_a0, sherman, err = a.T
    if theta[0] < 0.0 or theta[-1] > 180.0 or np.any(np.diff(theta) <= 0.0):
        raise ValueError("ELSEPA angles must rise strictly within [0, 180] deg")
    if mu[0] < 0.0 or mu[-1] > 1.0 or np.any(np.diff(mu) < 0.0):
        raise ValueError("ELSEPA mu must be nondecreasing within [0, 1]")
    if np.any(dcs <= 0.0) or np.any(dcs_a0 <= 0.0) or np.any(err < 0.0):
        raise ValueError("ELSEPA cross sections must be positive and errors nonnegative")
    return ElsepaPanel(found["energy_eV"], Z, theta, mu, dcs, dcs_a0, sherman, err, found["total_elastic_cm2"],
                       found["transport1_cm2"], found["transport2_cm2"], found["absorption_cm2"])


def angular_cdf(panel: ElsepaPanel) -> np.ndarray:
    """Normalised CDF of the DCS in ``mu`` on the panel's grid (trapezoid rule)."""
    inc = 0.5 * (panel.dcs_cm2_sr[1:] + panel.dcs_cm2_sr[:-1]) * np.diff(panel.mu)
    cdf = np.r_[0.0, np.cumsum(inc)]
    if not (np.isfinite(cdf[-1]) and cdf[-1] > 0.0):
        raise ValueError("ELSEPA DCS has no positive angular integral")
    cdf /= cdf[-1]
    cdf[-1] = 1.0
    return cdf


def stack_panels(panels, energies_eV=None) -> dict:
    """Panels of one angular grid as the native-grid table, ascending in energy."""
    ps = sorted(panels, key=lambda p: p.energy_eV)
    if not ps:
        raise ValueError("no ELSEPA panels")
    ref = ps[0]
    for p in ps[1:]:
        if not (np.array_equal(p.theta_deg, ref.theta_deg) and np.array_equal(p.mu, ref.mu)):
            raise ValueError("ELSEPA panels on different angular grids")
    E = np.array([p.energy_eV for p in ps])
    if energies_eV is not None:
        want = np.sort(np.asarray(energies_eV, dtype=np.float64))
        if want.shape != E.shape or not np.allclose(want, E, rtol=5.1e-6, atol=0.0):
            raise ValueError("ELSEPA panels do not match the requested energies")
        E = want
    return {
        "energy_eV": E,
        "theta_deg": ref.theta_deg,
        "mu": ref.mu,
        "dcs_cm2_sr": np.stack([p.dcs_cm2_sr for p in ps]),
        "dcs_a0_2_sr": np.stack([p.dcs_a0_2_sr for p in ps]),
        "angular_cdf": np.stack([angular_cdf(p) for p in ps]),
        "sherman": np.stack([p.sherman for p in ps]),
        "relative_error": np.stack([p.relative_error for p in ps]),
        "total_elastic_cm2": np.array([p.total_elastic_cm2 for p in ps]),
        "transport1_cm2": np.array([p.transport1_cm2 for p in ps]),
        "transport2_cm2": np.array([p.transport2_cm2 for p in ps]),
        "absorption_cm2": np.array([p.absorption_cm2 for p in ps]),
    }
