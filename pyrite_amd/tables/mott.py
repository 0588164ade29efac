"""NIST SRD 64 Mott transport tables and the Mott-calibrated elastic model.

The reference's ``elastic_model="mott"`` (``src/pyrite/montecarlo/transport/
scattering.py:1-240,494-515``, LUT columns ``lut.py:182-203``) keeps the
screened-Rutherford angular law of the step and changes two LUT columns:

* the elastic rate is Browning's fit to the Mott total cross sections,
  ``sigma = 3e-18 Z^1.7 / (E + 0.005 Z^1.7 sqrt(E) + 7e-4 Z^2 / sqrt(E))``
  [cm^2, E in keV] times the number density;
* the screening parameter ``alpha(E)`` is calibrated per element so that the
  screened-Rutherford first moment reproduces the tabulated Mott transport
  cross section: ``<1 - cos theta>(alpha) = sigma_tr / sigma_Browning`` with
  ``<1 - cos theta>(a) = 2a [(1 + a) ln(1 + 1/a) - 1]``, tabulated as
  ``log10 alpha`` against ``log10 E[eV]`` and interpolated linearly in those
  coordinates onto the LUT nodes.

The device step reads the same LUT columns for either model, so ``"mott"``
needs no kernel change. NIST SRD 64 may not be redistributed, so the tables
are user supplied: ``DisplayCalcTCSTableFor<El>.csv`` exports (rows
``No, Energy [eV], sigma_tr [a0^2]``) in the directory named by
``PYRITE_MOTT_TABLES_DIR`` (the reference also reads it from its config
store, ``mott.tables_dir``). A missing table raises
:class:`MottTableUnavailableError`; transport outside a table's energy range
is refused (:func:`check_mott_coverage`), never extrapolated.
"""

from __future__ import annotations

import os
from functools import lru_cache
from pathlib import Path

import numpy as np

#: Bohr radius squared [cm^2], the unit of the SRD 64 cross sections.
A0_SQ_CM2 = 2.8002852e-17
#: Environment variable naming the directory of the SRD 64 exports.
MOTT_TABLES_ENV = "PYRITE_MOTT_TABLES_DIR"
MOTT_SRD64_URL = "https://srdata.nist.gov/srd64/"

# first-moment targets are kept inside the law's range (0, 1)
_TARGET_MIN, _TARGET_MAX = 1e-12, 0.999


class MottTableUnavailableError(FileNotFoundError):
    """``elastic_model="mott"`` asked for an element whose SRD 64 table is not installed."""


def mott_table_filename(element: str) -> str:
    return f"DisplayCalcTCSTableFor{element}.csv"


def mott_tables_dir(tables_dir=None) -> Path | None:
    """The table directory: ``tables_dir`` if given, else ``PYRITE_MOTT_TABLES_DIR``."""
    value = tables_dir if tables_dir is not None else os.environ.get(MOTT_TABLES_ENV)
    return Path(value).expanduser().resolve() if value else None


def mott_table_path(element: str, tables_dir=None) -> Path:
    root = mott_tables_dir(tables_dir)
    name = mott_table_filename(element)
    if root is not None and (root / name).is_file():
        return root / name
    missing = f"{root / name} does not exist" if root is not None else f"{MOTT_TABLES_ENV} is not set"
    raise MottTableUnavailableError(
        f"elastic_model='mott' needs the NIST SRD 64 transport cross-section table for {element}, but "
        f"{missing}. NIST SRD 64 is not redistributed: export the relativistic transport cross sections "
        f"for {element} from {MOTT_SRD64_URL} as {name} into a directory named by {MOTT_TABLES_ENV}, "
        "or use elastic_model='sr'."
    )


@lru_cache(maxsize=None)
def _read(path: str):
    energies, sigmas = [], []
    with open(path, encoding="utf-8", errors="replace") as f:
        for line in f:
            cols = [c.strip() for c in line.split(",")]
            # data rows are "index, energy, cross section"; headers and notes are skipped
            if len(cols) == 3 and cols[0].isdigit():
                energies.append(float(cols[1]))
                sigmas.append(float(cols[2]))
    if not energies:
        raise ValueError(f"{path}: no SRD 64 transport cross-section rows")
    E = np.asarray(energies)
    if np.any(np.diff(E) <= 0.0):
        raise ValueError(f"{path}: SRD 64 energies must increase")
    return E, np.asarray(sigmas) * A0_SQ_CM2


def read_transport_table(path) -> tuple[np.ndarray, np.ndarray]:
    """``(E_eV, sigma_tr_cm2)`` of one SRD 64 transport cross-section export."""
    E, s = _read(str(Path(path).resolve()))
    return E.copy(), s.copy()


def browning_sigma_cm2(Z, E_keV):
    """Browning's fit to the Mott total elastic cross section [cm^2]
    (J. Appl. Phys. 76, 2016 (1994); ``scattering.py:71-96``)."""
    E = np.asarray(E_keV, dtype=np.float64)
    z17 = float(Z) ** 1.7
    root = np.sqrt(E)
    return 3.0e-18 * z17 / (E + 0.005 * z17 * root + 0.0007 * float(Z) * float(Z) / root)


def first_moment(alpha):
    """``<1 - cos theta>`` of the screened-Rutherford law with screening ``alpha``."""
    a = np.asarray(alpha, dtype=np.float64)
    return 2.0 * a * ((1.0 + a) * np.log1p(1.0 / a) - 1.0)


def alpha_from_first_moment(target):
    """Invert :func:`first_moment` (increasing from 0 to 1 in ``alpha``).

    Newton's method on ``ln alpha`` from the small-alpha asymptote
    ``t ~ 2a (ln(1/a) - 1)``, safeguarded by a bracket that every step
    narrows; targets are clipped to ``[1e-12, 0.999]`` as the reference's
    calibration clips them (``scattering.py:123-141``). The derivative is
    ``d<1-cos>/da = 2 (1 + 2a) ln(1 + 1/a) - 4``."""
    t = np.clip(np.asarray(target, dtype=np.float64), _TARGET_MIN, _TARGET_MAX)
    lo = np.full(t.shape, np.log(1e-14))
    hi = np.full(t.shape, np.log(1e5))
    # initial guess from the small-alpha form, one fixed-point pass
    u = np.log(np.maximum(t / 2.0, 1e-14))
    u = np.log(np.maximum(t / (2.0 * np.maximum(-u - 1.0, 0.5)), 1e-14))
    u = np.clip(u, lo, hi)
    for _ in range(100):
        a = np.exp(u)
        f = first_moment(a) - t
        lo = np.where(f < 0.0, u, lo)
        hi = np.where(f > 0.0, u, hi)
        slope = a * (2.0 * (1.0 + 2.0 * a) * np.log1p(1.0 / a) - 4.0)
        with np.errstate(divide="ignore", invalid="ignore"):
            step = u - f / slope
        bad = ~np.isfinite(step) | (step <= lo) | (step >= hi)
        new = np.where(bad, 0.5 * (lo + hi), step)
        if np.all(np.abs(new - u) <= 1e-15 * np.maximum(1.0, np.abs(u))):
            u = new
            break
        u = new
    return np.exp(u)


@lru_cache(maxsize=None)
def _alpha_table(path: str, Z: float):
    E_eV, sig_tr = _read(path)
    alpha = alpha_from_first_moment(sig_tr / browning_sigma_cm2(Z, E_eV * 1e-3))
    return np.log10(E_eV), np.log10(alpha)


def mott_alpha_table(element: str, Z, tables_dir=None) -> tuple[np.ndarray, np.ndarray]:
    """``(log10 E_eV, log10 alpha)`` nodes of ``element``'s calibrated screening."""
    logE, logA = _alpha_table(str(mott_table_path(element, tables_dir)), float(Z))
    return logE.copy(), logA.copy()


def mott_alpha(logE_eV_nodes, log_alpha_nodes, E_keV):
    """Screening at ``E_keV``: linear in ``log10 E[eV]`` between nodes, ends held."""
    return 10.0 ** np.interp(np.log10(np.asarray(E_keV, dtype=np.float64) * 1e3), logE_eV_nodes, log_alpha_nodes)


def check_mott_coverage(compositions, E_min_keV, E_max_keV, tables_dir=None) -> None:
    """Refuse transport whose energy range leaves any element's table (both
    ends inclusive, in the table's own eV converted to keV;
    ``scattering.py:494-515``)."""
    seen = set()
    for comp in compositions:
        for element, _ in comp:
            if element in seen:
                continue
            seen.add(element)
            E_eV, _ = read_transport_table(mott_table_path(element, tables_dir))
            lower, upper = float(E_eV[0]) / 1e3, float(E_eV[-1]) / 1e3
            if E_min_keV < lower or E_max_keV > upper:
                raise ValueError(
                    f"transport energy range [{E_min_keV:g}, {E_max_keV:g}] keV must be within the Mott "
                    f"SRD 64 table for {element} [{lower:g}, {upper:g}] keV; use elastic_model=\"elsepa\" "
                    "tables (not read by this backend) or elastic_model=\"sr\""
                )
