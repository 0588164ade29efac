"""Reader of ELSEPA elastic differential cross sections (``dcs_*.dat``).

The reference generates its elastic tables by running ELSEPA and reading each
energy panel back (``src/pyrite/xsgen/elsepa/parse.py:48-173``): the header
carries the kinetic energy and the integrated cross sections, the body one row
per polar angle with ``theta [deg], mu = (1 - cos theta)/2, DCS [cm^2/sr],
DCS [a0^2/sr], Sherman function, relative error``.  Panels on one angular
grid stack into the native-grid table the reference stores, with the angular
CDF in ``mu`` by the trapezoid rule on that grid (``dOmega = 4 pi dmu``).

The second half packs native-grid tables for the transport step's
``elastic_model="elsepa"`` (``src/pyrite/montecarlo/transport/scattering.py:
333-470``): per (layer, element) the energy nodes ``ln E[eV]``, the
macroscopic rate ``ln(n sigma)`` (log-log in energy), and per energy row the
piecewise-linear angular density in ``mu`` with its exact CDF and a guide
table that bounds the device's panel search.  The device inverts two
bracketing rows at one uniform and mixes the quantiles linearly in ``ln E``.
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


# ---- packing for the transport step (scattering.py:333-470, 494-510) --------------------------
#: Guide cells per angular row: a power of two, so ``xi * cells`` and the cell
#: bounds ``j / cells`` are exact (scattering.py:356-360).
GUIDE_CELLS = 1024


def angular_pdf(mu, dcs):
    """``(pdf, cdf)`` rows of piecewise-linear densities in ``mu`` from DCS rows.

    ``dOmega = 4 pi dmu`` for an azimuthally symmetric DCS, so ``p(mu)`` is the
    DCS over its trapezoid integral on the native grid; ``cdf`` is the exact
    integral of that piecewise-linear ``pdf`` (0 at the first node, 1 at the last)."""
    mu = np.asarray(mu, dtype=np.float64)
    dcs = np.atleast_2d(np.asarray(dcs, dtype=np.float64))
    if dcs.shape[1] != mu.size:
        raise ValueError("ELSEPA DCS rows must match the angular grid")
    panel = 0.5 * (dcs[:, 1:] + dcs[:, :-1]) * np.diff(mu)[None, :]
    total = panel.sum(axis=1)
    if not np.all(np.isfinite(total)) or np.any(total <= 0.0):
        raise ValueError("ELSEPA DCS rows need a positive finite angular integral")
    pdf = dcs / total[:, None]
    cdf = np.concatenate([np.zeros((dcs.shape[0], 1)), np.cumsum(panel, axis=1) / total[:, None]], axis=1)
    cdf[:, -1] = 1.0
    return pdf, cdf


def guide_table(cdf, cells: int = GUIDE_CELLS) -> np.ndarray:
    """``guide[r, j]``: the last panel ``k`` in ``[0, n_mu - 2]`` with ``cdf[r, k] <= j / cells``.

    For ``xi`` in cell ``j`` the panel a bisection over the whole row picks
    (the last ``k`` with ``cdf[r, k] <= xi``) lies in ``[guide[r, j],
    guide[r, j + 1]]``, so the bisection confined to that window picks the
    same panel, ties and zero-mass panels included; this needs the interior
    CDF non-decreasing, which is checked."""
    cdf = np.atleast_2d(np.asarray(cdf, dtype=np.float64))
    if cells < 1 or (cells & (cells - 1)) != 0:
        raise ValueError("guide cells must be a positive power of two")
    n_mu = cdf.shape[1]
    inner = cdf[:, : n_mu - 1]
    if np.any(np.diff(inner, axis=1) < 0.0):
        raise ValueError("ELSEPA angular CDF rows must be non-decreasing")
    bounds = np.arange(cells + 1, dtype=np.float64) / cells
    out = np.empty((cdf.shape[0], cells + 1), dtype=np.int32)
    for r in range(cdf.shape[0]):
        out[r] = np.clip(np.searchsorted(inner[r], bounds, side="right") - 1, 0, n_mu - 2)
    out[:, 0] = 0
    out[:, cells] = n_mu - 2
    return out


@dataclass
class ElsepaGroup:
    """Packed tables of a layer stack, laid out as the device step reads them.

    ``start[L, i]`` / ``length[L, i]``: the element's energy rows in
    ``logE`` / ``log_rate`` / ``cdf`` / ``pdf`` / ``guide`` (row-major, one
    angular grid ``mu`` shared by every table)."""

    start: np.ndarray  # int64 [n_layers, max_el]
    length: np.ndarray  # int64 [n_layers, max_el]
    logE: np.ndarray  # ln E[eV] per row
    log_rate: np.ndarray  # ln(n_cm3 sigma_el [1/cm]) per row
    mu: np.ndarray  # [n_mu]
    cdf: np.ndarray  # [rows, n_mu]
    pdf: np.ndarray  # [rows, n_mu]
    guide: np.ndarray  # int32 [rows, cells + 1]
    E_range_keV: list  # per layer, per element: (lowest, highest) tabulated energy

    @property
    def guide_cells(self) -> int:
        return int(self.guide.shape[1] - 1)


def pack_tables(elastic_tables, L_ncm3, rate_scales=None, cells: int = GUIDE_CELLS) -> ElsepaGroup:
    """Pack ``elastic_tables[L][i]`` (mappings with ``energy_eV``,
    ``total_elastic_cm2``, ``mu``, ``dcs_cm2_sr``; :func:`stack_panels` makes
    one) for number densities ``L_ncm3[L][i]`` [1/cm^3].  ``rate_scales[L][i]``
    multiplies that element's rate node by node and leaves its angles alone."""
    if len(elastic_tables) != len(L_ncm3):
        raise ValueError("elastic_tables needs one entry per transport layer")
    max_el = max(len(n) for n in L_ncm3)
    start = np.zeros((len(L_ncm3), max_el), dtype=np.int64)
    length = np.zeros((len(L_ncm3), max_el), dtype=np.int64)
    logE, log_rate, pdfs, cdfs, ranges = [], [], [], [], []
    mu_ref = None
    row = 0
    for L, (layer, ncm3) in enumerate(zip(elastic_tables, L_ncm3)):
        if len(layer) != len(ncm3):
            raise ValueError(f"elastic_tables layer {L} needs one table per element")
        ranges.append([])
        for i, (table, n) in enumerate(zip(layer, ncm3)):
            E = np.asarray(table["energy_eV"], dtype=np.float64)
            sigma = np.asarray(table["total_elastic_cm2"], dtype=np.float64)
            mu = np.asarray(table["mu"], dtype=np.float64)
            if mu_ref is None:
                mu_ref = mu
            elif not np.array_equal(mu, mu_ref):
                raise ValueError("every ELSEPA table must share one angular grid")
            if E.size < 2 or np.any(np.diff(E) <= 0.0) or sigma.shape != E.shape or np.any(sigma <= 0.0):
                raise ValueError("ELSEPA tables need at least two rising energies and positive cross sections")
            pdf, cdf = angular_pdf(mu, table["dcs_cm2_sr"])
            if pdf.shape[0] != E.size:
                raise ValueError("ELSEPA tables need one DCS row per energy")
            rate = float(n) * sigma
            if rate_scales is not None:
                rate = rate * np.asarray(rate_scales[L][i], dtype=np.float64)
            start[L, i], length[L, i] = row, E.size
            row += E.size
            logE.append(np.log(E))
            log_rate.append(np.log(rate))
            pdfs.append(pdf)
            cdfs.append(cdf)
            ranges[L].append((float(E[0]) * 1e-3, float(E[-1]) * 1e-3))
    cdf = np.ascontiguousarray(np.vstack(cdfs))
    return ElsepaGroup(start=start, length=length, logE=np.concatenate(logE), log_rate=np.concatenate(log_rate),
                       mu=np.ascontiguousarray(mu_ref), cdf=cdf, pdf=np.ascontiguousarray(np.vstack(pdfs)),
                       guide=guide_table(cdf, cells), E_range_keV=ranges)


def bracket(logE_rows: np.ndarray, x):
    """Lower node and fraction of ``x`` on one ascending row, clamped to its
    ends (``_jit_shell_device.py:58-84``): below the row ``(0, 0)``, above it
    ``(n - 2, 1)``."""
    x = np.asarray(x, dtype=np.float64)
    n = logE_rows.size
    lo = np.clip(np.searchsorted(logE_rows, x, side="right") - 1, 0, n - 2)
    f = (x - logE_rows[lo]) / (logE_rows[lo + 1] - logE_rows[lo])
    f = np.where(x <= logE_rows[0], 0.0, np.where(x >= logE_rows[-1], 1.0, f))
    lo = np.where(x <= logE_rows[0], 0, np.where(x >= logE_rows[-1], n - 2, lo))
    return lo, f


def element_rate(group: ElsepaGroup, L: int, i: int, E_keV) -> np.ndarray:
    """Macroscopic elastic rate [1/cm] of element ``i`` of layer ``L``, log-log between nodes."""
    s, n = int(group.start[L, i]), int(group.length[L, i])
    lo, f = bracket(group.logE[s:s + n], np.log(np.asarray(E_keV, dtype=np.float64) * 1e3))
    r0 = group.log_rate[s + lo]
    return np.exp(r0 + f * (group.log_rate[s + lo + 1] - r0))


def check_coverage(group: ElsepaGroup, E_min_keV: float, E_max_keV: float) -> None:
    """Refuse transport outside any table's energies; never extrapolate
    (``scattering.py:481-491``)."""
    for layer in group.E_range_keV:
        for lower, upper in layer:
            if E_min_keV < lower or E_max_keV > upper:
                raise ValueError(f"transport energy range must be within ELSEPA table [{lower:g}, {upper:g}] keV")
