"""Host-side transport tables: per-layer element rows and the ln-E LUT.

Restates, in NumPy, the reference's host table build:

* per-layer element coefficients -- ``src/pyrite/montecarlo/transport/layer_tables.py:24-154``
  (``coeff = n/0.602214076 * Z``, ``k = 0.731 + 0.0688 log10 Z``,
  SR rate numerator ``5.21e-21 Z^2 4 pi n_cm3``, Joy screening ``3.4e-3 Z^0.67``);
* the Joy-Luo / Berger-Seltzer crossover bisection --
  ``src/pyrite/montecarlo/transport/stopping.py:312-344``;
* the uniform-in-ln(E) transport LUT -- ``src/pyrite/montecarlo/transport/lut.py:28-437``
  (1024 intervals/decade, 256..16384 nodes, pinned endpoints);
* the elastic columns of the two table models the LUT kernel serves
  (``lut.py:182-203``): ``"sr"`` (screened-Rutherford rate, Joy/Bishop
  screening) and ``"mott"`` (Browning rate ``3e-18 Z^1.7 n_cm3 / (E + 0.005
  Z^1.7 sqrt E + 7e-4 Z^2 / sqrt E)``, screening calibrated on the SRD 64
  transport tables, ``pyrite_amd.tables.mott``).

The arrays produced here are the kernel inputs.  They are float64, C-contiguous
and laid out exactly as the device kernels index them.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .elements import TRANSPORT_ELEMENTS

MC2_KEV = 510.99895
BS_PREFACTOR = 1.535e-6
LN2 = 0.6931471805599453
_LN10 = np.log(10.0)
_MIN_DLOG_E = 8.0 * float(np.finfo(np.float32).eps)


# ---- stopping (stopping.py:180-260, 384-411) --------------------------------
def _dEds_joy_luo(Z, A, J, rho, E):
    k = 0.731 + 0.0688 * np.log10(Z)
    return -7.85e-4 * rho * Z / (A * E) * np.log(1.166 * (E + k * J) / J)


def _dEds_berger_seltzer(Z, A, J, rho, E):
    tau = E / MC2_KEV
    gamma = 1.0 + tau
    beta_sq = 1.0 - 1.0 / (gamma * gamma)
    f_minus = 1.0 - beta_sq + (tau * tau / 8.0 - (2.0 * tau + 1.0) * LN2) / (gamma * gamma)
    I_rel = J / MC2_KEV
    log_term = np.log(tau * tau * (tau + 2.0) / (2.0 * I_rel * I_rel))
    return -BS_PREFACTOR * rho * Z / (A * beta_sq) * (log_term + f_minus)


_CROSSOVER_CACHE: dict[str, float] = {}


def crossover_keV(element: str) -> float:
    """Energy where Joy-Luo and Berger-Seltzer stopping agree (stopping.py:312-344).

    200-step bisection on ``[2 J, 300]`` keV; density cancels in the ratio.
    """
    cached = _CROSSOVER_CACHE.get(element)
    if cached is not None:
        return cached
    p = TRANSPORT_ELEMENTS[element]
    Z, A, J = float(p["Z"]), float(p["A"]), float(p["J_keV"])

    def diff(E):
        return _dEds_joy_luo(Z, A, J, 1.0, E) - _dEds_berger_seltzer(Z, A, J, 1.0, E)

    lo, hi = 2.0 * J, 300.0
    if diff(lo) * diff(hi) > 0.0:
        raise ValueError(f"no Joy-Luo/Berger-Seltzer crossover in [{lo}, {hi}] keV for {element}")
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if diff(lo) * diff(mid) <= 0.0:
            hi = mid
        else:
            lo = mid
    out = 0.5 * (lo + hi)
    _CROSSOVER_CACHE[element] = out
    return out


def spliced_dEds(J, k, coeff, E_cross, E):
    """Vectorised spliced stopping [keV/A], negative (stopping.py:347-381)."""
    E = np.asarray(E, dtype=np.float64)
    tau = E / MC2_KEV
    gamma = 1.0 + tau
    beta_sq = 1.0 - 1.0 / (gamma * gamma)
    f_minus = 1.0 - beta_sq + (tau * tau / 8.0 - (2.0 * tau + 1.0) * LN2) / (gamma * gamma)
    jl = np.zeros_like(E)
    bs = np.zeros_like(E)
    for Ji, ki, ci, Ec in zip(J, k, coeff, E_cross):
        use_bs = E >= Ec
        I_rel = Ji / MC2_KEV
        jl += np.where(use_bs, 0.0, ci * np.log(1.166 * (E + ki * Ji) / Ji))
        bs += np.where(use_bs, ci * (np.log(tau * tau * (tau + 2.0) / (2.0 * I_rel * I_rel)) + f_minus), 0.0)
    return -7.85e-4 / E * jl - BS_PREFACTOR / beta_sq * bs


# ---- layers (layer_tables.py:24-154) ----------------------------------------
@dataclass
class LayerTables:
    """Padded ``(n_layers, max_el)`` element rows plus geometry of a layer stack."""

    n_layers: int
    max_el: int
    L_nel: np.ndarray  # int32 [n_layers]
    L_top: np.ndarray  # [n_layers]
    L_bot: np.ndarray  # [n_layers]
    internal_bounds: np.ndarray  # [max(n_layers-1, 1)]
    z_total: float
    L_Z: np.ndarray  # [n_layers, max_el]
    L_J: np.ndarray
    L_k: np.ndarray
    L_coeff: np.ndarray
    L_E_cross: np.ndarray
    L_ncm3: np.ndarray
    L_sr_rate_numer: np.ndarray
    L_sr_joy_numer: np.ndarray
    compositions: list
    elastic_model: str = "sr"
    # Browning rate coefficients (layer_tables.py:104-110): 3e-18 Z^1.7 n_cm3, 0.005 Z^1.7, 7e-4 Z^2
    L_mott_numer: np.ndarray | None = None
    L_mott_denom1: np.ndarray | None = None
    L_mott_denom2: np.ndarray | None = None
    # per layer, per element: (log10 E_eV, log10 alpha) calibration nodes under "mott"
    mott_tables: list | None = None
    # packed tabulated elastic tables under "elsepa" (pyrite_amd.tables.elsepa.ElsepaGroup)
    elsepa: object | None = None


ELASTIC_MODELS = ("sr", "mott", "elsepa")
ATOMIC_ELECTRON_DEFLECTION_MODELS = ("kawrakow",)


def build_layer_tables(layers, elastic_model: str = "sr", mott_tables_dir=None, atomic_electron_deflection=None,
                       hard_cutoff_eV=None, elastic_tables=None) -> LayerTables:
    """``layers = [(z_top, z_bot, [(element, n_per_A3), ...]), ...]`` top first.

    ``elastic_model="mott"`` reads every element's SRD 64 table
    (``mott_tables_dir``, else ``PYRITE_MOTT_TABLES_DIR``) here, before any
    transport, and raises ``MottTableUnavailableError`` for a missing one.
    ``elastic_model="elsepa"`` packs ``elastic_tables[L][i]`` (one ELSEPA
    native-grid table per element, ``pyrite_amd.tables.elsepa``): the LUT's
    rate and element columns then come from their log-log rates and the step
    samples their tabulated angles (``layer_tables.py:29-40``).

    ``atomic_electron_deflection="kawrakow"`` folds the deflection by the Z
    atomic electrons into the elastic rate, ``Z^2 -> Z (Z + xi)``
    (``atomic_electrons.py:1-20``, Kawrakow 1997).  Under continuous stopping
    no atomic-electron collision is explicit, so ``xi = 1`` at every energy and
    both analytic rate numerators scale by ``1 + 1/Z``; the denominators and
    the screening are untouched (``layer_tables.py:121-130``).  With explicit
    hard collisions (``hard_cutoff_eV``, the shell mode) ``xi`` depends on
    energy (``atomic_electrons.atomic_electron_xi``): the ELSEPA tables carry
    it node by node in their rates, and the analytic coefficients cannot, so
    that combination is refused for "sr" and "mott", as the reference does.
    ``None`` leaves every coefficient bit for bit."""
    if elastic_model not in ELASTIC_MODELS:
        raise ValueError(f"elastic_model must be one of {ELASTIC_MODELS}, got {elastic_model!r}")
    if atomic_electron_deflection is not None and atomic_electron_deflection not in ATOMIC_ELECTRON_DEFLECTION_MODELS:
        raise ValueError("atomic_electron_deflection must be None or one of "
                         + ", ".join(repr(m) for m in ATOMIC_ELECTRON_DEFLECTION_MODELS))
    if elastic_model == "elsepa" and elastic_tables is None:
        raise ValueError("elastic_model='elsepa' needs elastic_tables (one ELSEPA table per layer element)")
    if elastic_model != "elsepa" and elastic_tables is not None:
        raise ValueError("elastic_tables is only valid with elastic_model='elsepa'")
    if atomic_electron_deflection is not None and hard_cutoff_eV is not None and elastic_model != "elsepa":
        raise ValueError("atomic_electron_deflection with inelastic_model='shell-soft-hard' needs "
                         "elastic_model='elsepa' (its energy-dependent xi rides on the table nodes); pass "
                         "atomic_electron_deflection=None to opt out")
    n_layers = len(layers)
    max_el = max(len(c) for _, _, c in layers)
    shape = (n_layers, max_el)
    out = {name: np.zeros(shape) for name in ("Z", "J", "k", "coeff", "E_cross", "ncm3", "srn", "sjn", "mn", "md1",
                                               "md2")}
    mott = [[None] * len(c) for _, _, c in layers]
    nel = np.zeros(n_layers, dtype=np.int32)
    for L, (_, _, comp) in enumerate(layers):
        nel[L] = len(comp)
        for i, (el, n_i) in enumerate(comp):
            p = TRANSPORT_ELEMENTS[el]
            Z = float(p["Z"])
            out["Z"][L, i] = Z
            out["J"][L, i] = float(p["J_keV"])
            out["k"][L, i] = 0.731 + 0.0688 * np.log10(Z)
            out["coeff"][L, i] = (n_i / 0.602214076) * Z
            out["E_cross"][L, i] = crossover_keV(el)
            ncm3 = n_i * 1e24
            out["ncm3"][L, i] = ncm3
            out["srn"][L, i] = 5.21e-21 * Z * Z * np.float64(4.0) * np.float64(np.pi) * ncm3
            out["sjn"][L, i] = np.float64(3.4e-3) * Z ** np.float64(0.67)
            z17 = Z ** np.float64(1.7)
            out["mn"][L, i] = np.float64(3.0e-18) * z17 * ncm3
            out["md1"][L, i] = np.float64(0.005) * z17
            out["md2"][L, i] = np.float64(0.0007) * Z * Z
            if atomic_electron_deflection is not None:
                scale = 1.0 + 1.0 / Z  # 1 + xi/Z with xi = 1 (continuous stopping)
                out["srn"][L, i] *= scale
                out["mn"][L, i] *= scale
            if elastic_model == "mott":
                from ..tables.mott import mott_alpha_table

                mott[L][i] = mott_alpha_table(el, Z, mott_tables_dir)
    tops = np.array([float(l[0]) for l in layers])
    bots = np.array([float(l[1]) for l in layers])
    if n_layers > 1:
        internal = bots[:-1].copy()
    else:
        internal = np.zeros(1)
    return LayerTables(
        n_layers=n_layers,
        max_el=max_el,
        L_nel=nel,
        L_top=tops,
        L_bot=bots,
        internal_bounds=internal,
        z_total=float(bots[-1]),
        L_Z=out["Z"],
        L_J=out["J"],
        L_k=out["k"],
        L_coeff=out["coeff"],
        L_E_cross=out["E_cross"],
        L_ncm3=out["ncm3"],
        L_sr_rate_numer=out["srn"],
        L_sr_joy_numer=out["sjn"],
        compositions=[list(c) for _, _, c in layers],
        elastic_model=elastic_model,
        L_mott_numer=out["mn"],
        L_mott_denom1=out["md1"],
        L_mott_denom2=out["md2"],
        mott_tables=mott if elastic_model == "mott" else None,
        elsepa=_pack_elsepa(elastic_tables, layers, out["ncm3"], out["Z"], nel, atomic_electron_deflection,
                            hard_cutoff_eV) if elastic_model == "elsepa" else None,
    )


def _pack_elsepa(elastic_tables, layers, ncm3, Z, nel, atomic_electron_deflection, hard_cutoff_eV):
    """ELSEPA group of the stack; the atomic-electron factor 1 + xi(E)/Z per
    table node, xi from the layer's composition and the shell cutoff
    (``layer_tables.py:131-146``)."""
    from ..tables.elsepa import pack_tables
    from .atomic_electrons import atomic_electron_xi, elastic_rate_scale

    rows = [ncm3[L, :m] for L, m in enumerate(nel.tolist())]
    scales = None
    if atomic_electron_deflection is not None:
        scales = [[elastic_rate_scale(Z[L, i], atomic_electron_xi(np.asarray(t["energy_eV"], dtype=np.float64),
                                                                    layers[L][2], hard_cutoff_eV))
                   for i, t in enumerate(layer)] for L, layer in enumerate(elastic_tables)]
    return pack_tables(elastic_tables, rows, scales)


# ---- LUT (lut.py:28-437) ----------------------------------------------------
@dataclass
class TransportLUT:
    log_E_min: float
    inv_dlogE: float
    n_energy: int
    E_grid: np.ndarray
    total_rate: np.ndarray  # [n_layers, n_E]
    dEds: np.ndarray  # [n_layers, n_E]
    inv_beta: np.ndarray  # [n_E]
    cdf: np.ndarray  # [n_layers, max_el, n_E]
    alpha: np.ndarray  # [n_layers, max_el, n_E]
    max_rel_interp_error: float = 0.0


def _sr_rates(srn, sjn, E):
    """Screened-Rutherford rate [1/cm] (scattering.py:208-218)."""
    a = sjn / E
    rel = (E + 511.0) / (E + 1024.0)
    return srn / (E * E) / (a * (1.0 + a)) * (rel * rel)


def _browning_rates(numer, d1, d2, E):
    """Browning rate [1/cm] (scattering.py:88-96, lut.py:184-189)."""
    root = np.sqrt(E)
    return numer / (E + d1 * root + d2 / root)


def _evaluate_nodes(tables: LayerTables, E_grid):
    n_layers, max_el = tables.n_layers, tables.max_el
    n = E_grid.size
    total = np.empty((n_layers, n))
    dEds = np.empty((n_layers, n))
    cdf = np.ones((n_layers, max_el, n))
    alpha = np.zeros((n_layers, max_el, n))
    g = 1.0 + E_grid / 510.99895
    beta = np.sqrt(1.0 - 1.0 / (g * g))
    inv_beta = 1.0 / beta
    for L in range(n_layers):
        m = int(tables.L_nel[L])
        dEds[L] = spliced_dEds(tables.L_J[L, :m], tables.L_k[L, :m], tables.L_coeff[L, :m], tables.L_E_cross[L, :m], E_grid)
        rates = np.empty((m, n))
        for i in range(m):
            if tables.elastic_model == "elsepa":
                from ..tables.elsepa import element_rate

                rates[i] = element_rate(tables.elsepa, L, i, E_grid)  # alpha unused: tabulated angles
            elif tables.elastic_model == "mott":
                from ..tables.mott import mott_alpha

                rates[i] = _browning_rates(tables.L_mott_numer[L, i], tables.L_mott_denom1[L, i],
                                           tables.L_mott_denom2[L, i], E_grid)
                alpha[L, i] = mott_alpha(*tables.mott_tables[L][i], E_grid)
            else:
                rates[i] = _sr_rates(tables.L_sr_rate_numer[L, i], tables.L_sr_joy_numer[L, i], E_grid)
                alpha[L, i] = tables.L_sr_joy_numer[L, i] / E_grid
        layer_total = rates.sum(axis=0)
        if np.any(~np.isfinite(layer_total)) or np.any(layer_total <= 0.0):
            raise ValueError("transport LUT produced a non-positive elastic rate")
        total[L] = layer_total
        c = np.cumsum(rates, axis=0) / layer_total[None, :]
        c[-1, :] = 1.0
        cdf[L, :m, :] = c
    return total, dEds, inv_beta, cdf, alpha


class TransportLUTToleranceWarning(UserWarning):
    """The built transport LUT does not meet its requested interpolation tolerance."""


def build_transport_lut(E_min_keV, E_max_keV, tables: LayerTables, intervals_per_decade=1024,
                        min_points=256, max_points=16384, audit=None, rel_tol=1e-4) -> TransportLUT:
    """Uniform-ln(E) LUT over ``[E_min, E_max]`` (lut.py:275-437).

    With ``rel_tol > 0`` (the default 1e-4) the linear interpolant's relative
    error is measured at the quarter, half and three-quarter points of every
    interval in the total rate, stopping, 1/beta and the live elements'
    screening (the CDF is a probability and is not held to it) and a
    ``TransportLUTToleranceWarning`` says when it is unmet (``lut.py:380-419``);
    ``max_rel_interp_error`` reports what was achieved.  ``audit=False`` skips
    the measurement."""
    E_min_keV = float(E_min_keV)
    E_max_keV = float(E_max_keV)
    if not (np.isfinite(E_min_keV) and np.isfinite(E_max_keV)) or E_min_keV <= 0.0:
        raise ValueError("transport LUT energy bounds must be finite and positive")
    if E_max_keV <= E_min_keV:
        E_max_keV = float(np.nextafter(E_min_keV, np.inf))
    log_E_min = float(np.log(E_min_keV))
    log_E_max = float(np.log(E_max_keV))
    log_span = log_E_max - log_E_min
    decades = log_span / _LN10
    requested = int(np.ceil(decades * float(intervals_per_decade))) + 1
    n_energy = min(max_points, max(min_points, requested))
    if log_span / (n_energy - 1) < _MIN_DLOG_E:
        n_energy = max(2, int(log_span / _MIN_DLOG_E) + 1)
    inv_dlogE = np.float64((n_energy - 1) / log_span)
    logE = np.linspace(log_E_min, log_E_max, n_energy, dtype=np.float64)
    E_grid = np.exp(logE)
    E_grid[0] = E_min_keV
    E_grid[-1] = E_max_keV
    total, dEds, inv_beta, cdf, alpha = _evaluate_nodes(tables, E_grid)
    err = 0.0
    worst_keV = E_min_keV
    if audit is None:
        audit = float(rel_tol) > 0.0
    if audit:
        for w in (0.25, 0.5, 0.75):
            Es = np.exp((1.0 - w) * logE[:-1] + w * logE[1:])
            st, sd, sb, _, sa = _evaluate_nodes(tables, Es)
            # the screening of each layer's live elements too (lut.py:251-270); the CDF is a
            # probability and is not held to a relative bound
            screens = [] if tables.elastic_model == "elsepa" else \
                [(alpha[L, :m], sa[L, :m]) for L, m in enumerate(tables.L_nel.tolist())]
            for node, samp in ((total, st), (dEds, sd), (inv_beta[None, :], sb[None, :]), *screens):
                pred = (1.0 - w) * node[..., :-1] + w * node[..., 1:]
                with np.errstate(divide="ignore", invalid="ignore"):
                    e = np.abs(pred - samp) / np.abs(samp)
                e = np.where(np.isfinite(e), e, 0.0)
                while e.ndim > 1:
                    e = e.max(axis=0)
                if e.size and float(e.max()) > err:
                    err = float(e.max())
                    worst_keV = float(Es[int(np.argmax(e))])
        if float(rel_tol) > 0.0 and err > float(rel_tol):
            import warnings

            capped = n_energy < requested
            warnings.warn(
                f"transport LUT tolerance unmet: measured max relative interpolation error {err:.3e} > rel_tol "
                f"{float(rel_tol):.3e} at {worst_keV:.6g} keV with {n_energy} points over {E_min_keV:.6g}-"
                f"{E_max_keV:.6g} keV" + (f"; max_points={max_points} capped the requested {requested} points"
                                          if capped else "; raise intervals_per_decade, or accept the O(h) error "
                                          "a model join leaves in the interval that straddles it"),
                TransportLUTToleranceWarning, stacklevel=2)
    return TransportLUT(
        log_E_min=log_E_min,
        inv_dlogE=float(inv_dlogE),
        n_energy=int(n_energy),
        E_grid=E_grid,
        total_rate=np.ascontiguousarray(total),
        dEds=np.ascontiguousarray(dEds),
        inv_beta=np.ascontiguousarray(inv_beta),
        cdf=np.ascontiguousarray(cdf),
        alpha=np.ascontiguousarray(alpha),
        max_rel_interp_error=err,
    )
