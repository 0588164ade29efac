"""Tables of the hard-event transport modes (coupled radiative, shell soft/hard).

The reference runs both modes from fetched data: BremsLib scaled SDCS tables
for the radiative partition (``montecarlo/transport/hard_radiative.py:605``)
and SBETHE stopping with PENELOPE oscillator shells for the shell partition
(``montecarlo/transport/shell_transport.py``).  Neither data set ships with the
reference or exists offline, so this module builds analytic substitutes in
the same layouts.  The step algorithm that reads them (``csrc/radiative.hpp``,
``csrc/shell.hpp``, ``csrc/transport_step.hpp``) is the reference's; the
numbers in the tables are this module's models and do not reproduce BremsLib
or PENELOPE values.

* Radiative: ``chi(Z, T, kappa) = k dsigma/dk / Z^2`` [mb] (the layout's
  convention: ``dsigma/dk = chi Z^2 / k``, brem_bremslib.py:129-133) on a fixed
  grid of incident energies (1 keV - 1 GeV, 10 per decade) and reduced photon
  energies, taken as the larger of the Butcher-Messel screened relativistic
  Bethe-Heitler cross section (Thomas-Fermi screening functions, no Coulomb
  correction) and the nonrelativistic Bethe-Heitler cross section with the
  Elwert factor and a Thomas-Fermi momentum cut.  Layer plasma energies
  ``hbar w_p = hbar c sqrt(4 pi r_e n_e)`` switch on the Ter-Mikaelian
  suppression the reference applies in every coupled run
  (``hard_radiative.py:765-772``).
* Shell: per element, bound shells (X-ray binding energies ``U`` and
  occupations ``f`` of the common elements below) plus a conduction group
  (``U = 0``) with the remaining electrons; resonance energies
  ``W = sqrt((a U)^2 + 2/3 (f/Z) Omega_p^2)`` (conduction: ``Omega_p
  sqrt(f/Z)``) with ``a`` fixed by ``sum f ln W = Z ln I``; three channels per
  shell (distant longitudinal, distant transverse, close Moller) with the
  PENELOPE-2008 distant cross sections and the loss laws the step samples.
  The soft stopping closes on the spliced collision stopping of the plain
  step: ``S_soft = S_spliced - S_hard``; the soft variance is the model's.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .elements import TRANSPORT_ELEMENTS
from .tables import LayerTables, spliced_dEds

MC2_EV = 510998.95
ALPHA = 1.0 / 137.035999
R_E_ANG = 2.8179403262e-5
R_E2_MB = (R_E_ANG * 1.0e-8) ** 2 / 1.0e-27  # r_e^2 in mb
HBARC_EV_ANG = 1973.269804
# Salts of the mode key domains: radiative keys are stream keys of seed ^ salt
# (_jit_radiative.py:17-22); hard-collision keys splitmix64(stream_key ^ salt)
# (hard_inelastic.py:40,77-80).
RADIATIVE_STREAM_SALT = 0xA54FF53A5F1D36F1
HARD_STREAM_SALT = 0x3C6EF372FE94F82B

# ---- radiative ---------------------------------------------------------------
RAD_INCIDENT_KEV = np.logspace(0.0, 6.0, 61)
RAD_NOMINAL = np.array([0.0, 1e-4, 3e-4, 1e-3, 3e-3, 0.01, 0.02, 0.04, 0.06, 0.08, 0.1, 0.15, 0.2, 0.25, 0.3,
                        0.4, 0.5, 0.6, 0.7, 0.8, 0.85, 0.9, 0.93, 0.95, 0.97, 0.99, 0.995, 0.999, 1.0])


def _beta2(T_eV):
    g = 1.0 + T_eV / MC2_EV
    return 1.0 - 1.0 / (g * g)


def _chi_butcher_messel(Z, T_eV, kappa):
    """Screened relativistic Bethe-Heitler ``k dsigma/dk / Z^2`` [mb], clipped at 0."""
    E0 = T_eV + MC2_EV
    k = kappa * T_eV
    Ep = E0 - k
    delta = 136.0 * MC2_EV * Z ** (-1.0 / 3.0) * k / (E0 * Ep)
    lnz = np.log(Z) / 3.0
    low = delta <= 1.0
    d = np.where(low, delta, 1.0)
    big = 21.12 - 4.184 * np.log(np.where(low, 1.0, delta) + 0.952)
    F1 = np.where(low, 20.867 - 3.242 * d + 0.625 * d * d, big)
    F2 = np.where(low, 20.209 - 1.930 * d - 0.086 * d * d, big)
    r = Ep / E0
    bracket = (1.0 + r * r) * (F1 / 4.0 - lnz) - (2.0 / 3.0) * r * (F2 / 4.0 - lnz)
    return np.maximum(4.0 * ALPHA * R_E2_MB * bracket, 0.0)


def _chi_nonrelativistic(Z, T_eV, kappa):
    """Nonrelativistic Bethe-Heitler ``k dsigma/dk / Z^2`` with the Elwert factor and a Thomas-Fermi cut [mb]."""
    kap = np.minimum(kappa, 1.0 - 1.0e-9)
    Tf = T_eV * (1.0 - kap)
    t_i, t_f = T_eV / MC2_EV, Tf / MC2_EV
    p_i = np.sqrt(t_i * (t_i + 2.0))
    p_f = np.sqrt(t_f * (t_f + 2.0))
    b_i, b_f = p_i / (1.0 + t_i), p_f / (1.0 + t_f)
    dp = p_i - p_f
    q_tf = ALPHA * Z ** (1.0 / 3.0) / 0.885
    L = np.maximum(np.log((p_i + p_f) / np.sqrt(dp * dp + q_tf * q_tf)), 0.0)
    a = 2.0 * np.pi * Z * ALPHA
    elwert = b_i * (-np.expm1(-a / b_i)) / (b_f * (-np.expm1(-a / b_f)))
    return (16.0 / 3.0) * ALPHA * R_E2_MB * L * elwert / (p_i * p_i)


def scaled_sdcs(Z, T_keV, kappa):
    """Substitute scaled SDCS ``chi(Z, T, kappa)`` [mb] on broadcast ``T_keV`` x ``kappa``."""
    T = np.asarray(T_keV, dtype=np.float64) * 1.0e3
    kap = np.asarray(kappa, dtype=np.float64)
    return np.maximum(_chi_butcher_messel(float(Z), T, kap), _chi_nonrelativistic(float(Z), T, kap))


def plasma_energy_eV(composition):
    """``hbar w_p`` [eV] of a layer from its atom densities [1/A^3]."""
    n_e = sum(float(n) * TRANSPORT_ELEMENTS[el]["Z"] for el, n in composition)
    return HBARC_EV_ANG * np.sqrt(4.0 * np.pi * R_E_ANG * n_e)


@dataclass
class RadiativeTables:
    """``hr_hard_events`` radiative arrays, element slot ``s = layer*max_el + i``."""

    incident: np.ndarray  # [slots, max_t] keV
    nominal: np.ndarray  # [slots, n_red]
    top: np.ndarray  # [slots, max_t]
    chi: np.ndarray  # [slots, max_t, n_red] mb
    z: np.ndarray  # [slots]
    ncm3: np.ndarray  # [slots]
    nt: np.ndarray  # int32 [slots]
    plasma: np.ndarray  # [n_layers] eV
    max_t: int
    n_red: int
    cutoff_eV: float


def build_radiative_tables(layers: LayerTables, cutoff_eV: float, dielectric: bool = True) -> RadiativeTables:
    """Radiative tables of a layer stack with photon cutoff ``cutoff_eV``."""
    cutoff_eV = float(cutoff_eV)
    if not np.isfinite(cutoff_eV) or cutoff_eV <= 0.0:
        raise ValueError("radiative cutoff must be positive and finite")
    if layers.max_el > 8:
        raise ValueError("the radiative mode takes at most 8 elements per layer")
    n_slots = layers.n_layers * layers.max_el
    max_t, n_red = RAD_INCIDENT_KEV.size, RAD_NOMINAL.size
    incident = np.tile(RAD_INCIDENT_KEV, (n_slots, 1))
    nominal = np.tile(RAD_NOMINAL, (n_slots, 1))
    top = np.ones((n_slots, max_t))
    chi = np.zeros((n_slots, max_t, n_red))
    z = np.zeros(n_slots)
    ncm3 = np.zeros(n_slots)
    nt = np.full(n_slots, max_t, dtype=np.int32)
    plasma = np.zeros(layers.n_layers)
    for L in range(layers.n_layers):
        comp = layers.compositions[L]
        plasma[L] = plasma_energy_eV(comp) if dielectric else 0.0
        for i in range(int(layers.L_nel[L])):
            s = L * layers.max_el + i
            Z = float(layers.L_Z[L, i])
            z[s] = Z
            ncm3[s] = float(layers.L_ncm3[L, i])
            chi[s] = scaled_sdcs(Z, RAD_INCIDENT_KEV[:, None], RAD_NOMINAL[None, :])
    return RadiativeTables(incident=incident, nominal=nominal, top=top, chi=np.ascontiguousarray(chi), z=z,
                           ncm3=ncm3, nt=nt, plasma=plasma, max_t=max_t, n_red=n_red, cutoff_eV=cutoff_eV)


# ---- shell ------------------------------------------------------------------
# Bound shells (binding energy eV, occupation) of the elements the shell mode
# accepts; the remaining electrons form the conduction group (U = 0).
SHELLS = {
    "C": [(284.2, 2)],
    "N": [(409.9, 2)],
    "O": [(543.1, 2)],
    "Al": [(1559.6, 2), (117.8, 2), (72.9, 6)],
    "Si": [(1839.0, 2), (149.7, 2), (99.8, 6)],
    "Ge": [(11103.1, 2), (1414.6, 2), (1248.1, 2), (1217.0, 4), (180.1, 2), (124.9, 2), (120.8, 4), (29.8, 4),
           (29.2, 6)],
    "W": [(69525.0, 2), (12099.8, 2), (11544.0, 2), (10206.8, 4), (2819.6, 2), (2574.9, 2), (2281.0, 4),
          (1872.0, 4), (1809.0, 6), (594.1, 2), (490.4, 2), (423.6, 4), (255.9, 4), (243.5, 6), (33.6, 6),
          (31.4, 8), (75.6, 2), (45.3, 2), (36.8, 4)],
}
K_R_ANG2_EV = 2.0 * np.pi * R_E_ANG * R_E_ANG * MC2_EV  # 2 pi r_e^2 mc^2 [A^2 eV]; / beta^2 per electron


def _qmin(E, W):
    p0 = np.sqrt(E * (E + 2.0 * MC2_EV))
    rem = E - W
    p1 = np.sqrt(np.maximum(rem, 0.0) * (np.maximum(rem, 0.0) + 2.0 * MC2_EV))
    dp = W * (2.0 * (E + MC2_EV) - W) / (p0 + p1)
    return dp * dp / (np.sqrt(dp * dp + MC2_EV * MC2_EV) + MC2_EV)


def _loss_bounds(E, U, W, branch, Wc):
    """(lower, upper, peak_end) of a channel's hard interval, arrays over E."""
    w_max = 0.5 * (E + U)
    if branch == 2:
        base = U if U > 0.0 else W
        return np.maximum(base, Wc) + 0.0 * E, w_max, 0.0 * E
    if U == 0.0:
        return W + 0.0 * E, W + 0.0 * E, 0.0 * E
    pe = np.minimum(3.0 * W - 2.0 * U, E)
    return np.maximum(U, Wc) + 0.0 * E, np.minimum(pe, w_max), pe


def _moller_J(E, prime, w):
    """Antiderivative of F(-)(E', W)/W^2 (the close-collision loss law)."""
    a = (E / (E + MC2_EV)) ** 2
    rest = prime - w
    return -1.0 / w + 1.0 / rest + (1.0 - a) / prime * np.log(rest / w) + a * w / (prime * prime)


def _moller_F(E, prime, w):
    a = (E / (E + MC2_EV)) ** 2
    r = w / (prime - w)
    return 1.0 + r * r - (1.0 - a) * r + a * (w / prime) ** 2


_GL_X, _GL_W = np.polynomial.legendre.leggauss(48)


def _log_quad(f, lo, hi):
    """int_lo^hi f(W) dW by Gauss-Legendre in ln W, elementwise over arrays lo < hi (0 elsewhere)."""
    ok = hi > lo
    lo_s = np.where(ok, lo, 1.0)
    hi_s = np.where(ok, hi, 2.0)
    a, b = np.log(lo_s), np.log(hi_s)
    x = 0.5 * (b - a)[..., None] * _GL_X + 0.5 * (a + b)[..., None]
    w = np.exp(x)
    val = (f(w) * w * _GL_W).sum(-1) * 0.5 * (b - a)
    return np.where(ok, val, 0.0)


def _resonances(Z, shells, f_cb, omega_p, I_eV):
    """Resonance energies of the bound shells and the conduction group."""
    U = np.array([u for u, _ in shells], dtype=np.float64)
    f = np.array([n for _, n in shells], dtype=np.float64)
    W_cb = omega_p * np.sqrt(f_cb / Z) if f_cb > 0 else 0.0

    def resid(a):
        W = np.sqrt((a * U) ** 2 + (2.0 / 3.0) * (f / Z) * omega_p ** 2)
        s = float((f * np.log(W)).sum())
        if f_cb > 0:
            s += f_cb * np.log(W_cb)
        return s - Z * np.log(I_eV)

    lo, hi = 1.0e-3, 50.0
    if resid(lo) > 0.0:
        a = lo
    elif resid(hi) < 0.0:
        a = hi
    else:
        for _ in range(200):
            mid = 0.5 * (lo + hi)
            if resid(mid) < 0.0:
                lo = mid
            else:
                hi = mid
        a = 0.5 * (lo + hi)
    W = np.sqrt((a * U) ** 2 + (2.0 / 3.0) * (f / Z) * omega_p ** 2)
    return W, W_cb, a


@dataclass
class ShellChannel:
    element: str
    shell: int  # index into SHELLS[element]; len(...) for the conduction group
    branch: int  # 0 distant longitudinal, 1 distant transverse, 2 close
    U: float
    W: float
    f: float


@dataclass
class ShellTables:
    """``hr_hard_events`` shell arrays on the transport LUT grid."""

    rate: np.ndarray  # [n_layers, n_E] 1/A
    omega2: np.ndarray  # [n_layers, n_E] keV^2/A
    ch_rate: np.ndarray  # [n_layers, max_ch, n_E] 1/A
    U: np.ndarray  # [n_layers, max_ch] eV
    W: np.ndarray
    branch: np.ndarray  # int32
    code: np.ndarray  # int16
    nch: np.ndarray  # int32 [n_layers]
    max_ch: int
    cutoff_eV: float
    soft_dEds: np.ndarray  # [n_layers, n_E] keV/A (negative): the LUT dEds column of the mode
    hard_dEds: np.ndarray  # [n_layers, n_E] keV/A (positive): mean hard loss per length
    channels: list  # per layer, [ShellChannel]


def _channel_moments(E, ch: ShellChannel, Z, Wc):
    """Per-atom (hard rate [A^2], hard first [A^2 eV], soft first, soft second) over E [eV]."""
    b2 = _beta2(E)
    K = K_R_ANG2_EV / b2 * ch.f
    U, W = ch.U, ch.W
    zero = np.zeros_like(E)
    if ch.branch == 2:
        prime = E + U
        lo_h, up, _ = _loss_bounds(E, U, W, 2, Wc)
        base = U if U > 0.0 else W
        hard = np.where(up > lo_h, K * (_moller_J(E, prime, np.maximum(up, lo_h)) - _moller_J(E, prime, lo_h)),
                        0.0)
        h1 = K * _log_quad(lambda w: _moller_F(E[..., None], prime[..., None], w) / w, lo_h, up)
        s_hi = np.minimum(Wc, up)
        base_a = base + zero
        s1 = K * _log_quad(lambda w: _moller_F(E[..., None], prime[..., None], w) / w, base_a, s_hi)
        s2 = K * _log_quad(lambda w: _moller_F(E[..., None], prime[..., None], w), base_a, s_hi)
        return np.maximum(hard, 0.0), h1, s1, s2
    live = E > W
    if ch.branch == 0:
        q = _qmin(E, np.minimum(W, E))
        arg = (W / np.where(q > 0, q, 1.0)) * (q + 2.0 * MC2_EV) / (W + 2.0 * MC2_EV)
        tot = np.where(live & (q > 0), K / W * np.log(np.maximum(arg, 1.0)), 0.0)
    else:
        tot = np.where(live, K / W * np.maximum(-np.log1p(-b2) - b2, 0.0), 0.0)
    if U == 0.0:
        if W > Wc:
            return tot, tot * W, zero, zero
        return zero, zero, tot * W, tot * W * W
    lo_h, up, pe = _loss_bounds(E, U, W, ch.branch, Wc)
    Jt = lambda w: pe * np.log(w / U) - (w - U)  # noqa: E731
    norm = np.where(up > U, Jt(np.maximum(up, U)), 0.0)
    ok = norm > 0.0
    nrm = np.where(ok, norm, 1.0)
    hard = np.where(ok & (up > lo_h), tot * (Jt(np.maximum(up, lo_h)) - Jt(lo_h)) / nrm, 0.0)
    mom1 = lambda l, u: pe * (u - l) - 0.5 * (u * u - l * l)  # noqa: E731
    mom2 = lambda l, u: 0.5 * pe * (u * u - l * l) - (u ** 3 - l ** 3) / 3.0  # noqa: E731
    h1 = np.where(ok & (up > lo_h), tot * mom1(lo_h, np.maximum(up, lo_h)) / nrm, 0.0)
    s_hi = np.maximum(np.minimum(Wc, up), U)
    s1 = np.where(ok, tot * mom1(U, s_hi) / nrm, 0.0)
    s2 = np.where(ok, tot * mom2(U, s_hi) / nrm, 0.0)
    return hard, h1, s1, s2


def build_shell_tables(layers: LayerTables, E_grid_keV, cutoff_eV: float = 1000.0) -> ShellTables:
    """Shell-mode tables on the transport LUT's energy grid ``E_grid_keV``."""
    cutoff_eV = float(cutoff_eV)
    if not np.isfinite(cutoff_eV) or cutoff_eV <= 0.0:
        raise ValueError("shell cutoff must be positive and finite")
    E_keV = np.asarray(E_grid_keV, dtype=np.float64)
    E = E_keV * 1.0e3
    n_layers, n_E = layers.n_layers, E.size
    per_layer, rates, omega2, soft, hard_s = [], [], [], [], []
    for L in range(n_layers):
        comp = layers.compositions[L]
        omega_p = plasma_energy_eV(comp)
        chans, crates = [], []
        h1_tot = np.zeros(n_E)
        s2_tot = np.zeros(n_E)
        for el, n_i in comp:
            if el not in SHELLS:
                raise ValueError(f"the shell mode has no shell data for {el!r} (has {sorted(SHELLS)})")
            p = TRANSPORT_ELEMENTS[el]
            Z, I_eV = float(p["Z"]), float(p["J_keV"]) * 1.0e3
            shells = SHELLS[el]
            f_cb = Z - sum(f for _, f in shells)
            W, W_cb, _ = _resonances(Z, shells, f_cb, omega_p, I_eV)
            N = float(n_i)  # atoms / A^3
            groups = [(u, w, f) for (u, f), w in zip(shells, W)]
            if f_cb > 0:
                groups.append((0.0, W_cb, f_cb))
            for k, (u, w, f) in enumerate(groups):
                for branch in (0, 1, 2):
                    ch = ShellChannel(el, k, branch, float(u), float(w), float(f))
                    hr_, h1, _s1, s2 = _channel_moments(E, ch, Z, cutoff_eV)
                    chans.append(ch)
                    crates.append(N * hr_)
                    h1_tot += N * h1
                    s2_tot += N * s2
        dEds = spliced_dEds(layers.L_J[L, : layers.L_nel[L]], layers.L_k[L, : layers.L_nel[L]],
                            layers.L_coeff[L, : layers.L_nel[L]], layers.L_E_cross[L, : layers.L_nel[L]], E_keV)
        s_hard = h1_tot * 1.0e-3  # keV / A
        soft_dEds = dEds + s_hard
        if np.any(soft_dEds >= 0.0):
            raise ValueError("shell tables: the hard part exceeds the spliced stopping at some energy; "
                             "raise the cutoff")
        per_layer.append(chans)
        rates.append(np.asarray(crates))
        omega2.append(s2_tot * 1.0e-6)
        soft.append(soft_dEds)
        hard_s.append(s_hard)
    max_ch = max(len(c) for c in per_layer)
    ch_rate = np.zeros((n_layers, max_ch, n_E))
    U = np.zeros((n_layers, max_ch))
    Wr = np.zeros((n_layers, max_ch))
    branch = np.zeros((n_layers, max_ch), dtype=np.int32)
    code = np.full((n_layers, max_ch), -1, dtype=np.int16)
    nch = np.zeros(n_layers, dtype=np.int32)
    for L, chans in enumerate(per_layer):
        nch[L] = len(chans)
        ch_rate[L, : len(chans)] = rates[L]
        for c, ch in enumerate(chans):
            U[L, c], Wr[L, c], branch[L, c], code[L, c] = ch.U, ch.W, ch.branch, c
    rate = ch_rate.sum(axis=1)
    return ShellTables(rate=rate, omega2=np.asarray(omega2), ch_rate=ch_rate, U=U, W=Wr, branch=branch, code=code,
                       nch=nch, max_ch=max_ch, cutoff_eV=cutoff_eV, soft_dEds=np.asarray(soft),
                       hard_dEds=np.asarray(hard_s), channels=per_layer)


# ---- device side ----------------------------------------------------------------
# Row columns each mode adds (name, torch dtype name).
RAD_ROW_COLUMNS = (("seg_rad_k", "float64"), ("seg_rad_Z", "int16"))
SHELL_ROW_COLUMNS = (("seg_hard_W", "float64"), ("seg_hard_ch", "int16"))


class DeviceHardModes:
    """Device copies of the mode tables, the radiative LUT columns and the
    ``hr_hard_events`` record a transport call passes.

    ``rad`` and ``shell`` are host tables (``None``: mode off).  The radiative
    LUT columns are evaluated on the device over the transport LUT's grid
    (``hr_radiative_lut``).  With the shell mode the caller's transport LUT
    must carry ``shell.soft_dEds`` as its stopping column (``lut_for``)."""

    def __init__(self, layers: LayerTables, lut, device, rad: RadiativeTables | None = None,
                 shell: ShellTables | None = None):
        import ctypes as C

        import torch

        from .. import _lib

        self.device = device
        self.rad, self.shell = rad, shell
        self.rad_keys = None
        f64 = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=device)  # noqa: E731
        i32 = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.int32), device=device)  # noqa: E731
        h = _lib.HardEvents()
        self._hold = {}
        if rad is not None:
            d = {k: f64(getattr(rad, k)) for k in ("incident", "nominal", "top", "chi", "z", "ncm3", "plasma")}
            d["nt"] = i32(rad.nt)
            self._hold.update({"rad_" + k: v for k, v in d.items()})
            h.rad_on = 1
            h.rad_incident, h.rad_nominal, h.rad_top = (d[k].data_ptr() for k in ("incident", "nominal", "top"))
            h.rad_chi, h.rad_z, h.rad_ncm3 = d["chi"].data_ptr(), d["z"].data_ptr(), d["ncm3"].data_ptr()
            h.rad_nt, h.rad_plasma = d["nt"].data_ptr(), d["plasma"].data_ptr()
            h.rad_max_t, h.rad_n_red, h.rad_cutoff_eV = int(rad.max_t), int(rad.n_red), float(rad.cutoff_eV)
            E = f64(lut.E_grid)
            nel = i32(layers.L_nel)
            soft = torch.empty((layers.n_layers, lut.n_energy), dtype=torch.float64, device=device)
            rate = torch.empty_like(soft)
            stream = C.c_void_p(torch.cuda.current_stream(device).cuda_stream)
            _lib.check(_lib.load().hr_radiative_lut(C.byref(h), int(layers.n_layers), int(layers.max_el),
                                                     nel.data_ptr(), E.data_ptr(), int(lut.n_energy),
                                                     soft.data_ptr(), rate.data_ptr(), stream), "hr_radiative_lut")
            self._hold.update(rad_E=E, rad_nel=nel, lut_rad_soft=soft, lut_rad_rate=rate)
            h.lut_rad_soft, h.lut_rad_rate = soft.data_ptr(), rate.data_ptr()
        if shell is not None:
            d = {k: f64(getattr(shell, k)) for k in ("rate", "omega2", "ch_rate", "U", "W")}
            d["branch"] = i32(shell.branch)
            d["code"] = torch.as_tensor(np.ascontiguousarray(shell.code, dtype=np.int16), device=device)
            d["nch"] = i32(shell.nch)
            self._hold.update({"sh_" + k: v for k, v in d.items()})
            h.shell_on = 1
            h.lut_sh_rate, h.lut_sh_omega2, h.lut_sh_ch_rate = (d[k].data_ptr() for k in ("rate", "omega2",
                                                                                          "ch_rate"))
            h.sh_U, h.sh_W, h.sh_branch = d["U"].data_ptr(), d["W"].data_ptr(), d["branch"].data_ptr()
            h.sh_code, h.sh_nch = d["code"].data_ptr(), d["nch"].data_ptr()
            h.sh_max_ch, h.sh_cutoff_eV = int(shell.max_ch), float(shell.cutoff_eV)
        self.record = h

    @property
    def on(self) -> bool:
        return self.rad is not None or self.shell is not None

    @property
    def row_columns(self):
        return (RAD_ROW_COLUMNS if self.rad is not None else ()) + (SHELL_ROW_COLUMNS if self.shell is not None
                                                                    else ())

    def radiative_columns(self):
        """Device LUT columns (soft keV/A, rate 1/A) ``[n_layers, n_E]`` (radiative mode)."""
        return self._hold["lut_rad_soft"], self._hold["lut_rad_rate"]

    def set_keys(self, seed: int, start: int, n: int):
        """Radiative stream keys of electrons ``[start, start + n)`` (stream keys of seed ^ salt)."""
        import torch

        from .. import _lib

        if self.rad is None:
            return
        if self.rad_keys is None or self.rad_keys.numel() != int(n):
            self.rad_keys = torch.empty(int(n), dtype=torch.uint64, device=self.device)
        stream = torch.cuda.current_stream(self.device).cuda_stream
        _lib.call("hr_stream_keys", (int(seed) ^ RADIATIVE_STREAM_SALT) & (2**64 - 1), int(start), int(n),
                  self.rad_keys.data_ptr(), stream)

    def fill(self, hard, rows):
        """Copy the mode record into ``hard`` (an ``HardEvents`` field) with this call's keys and rows."""
        import ctypes as C

        C.memmove(C.addressof(hard), C.addressof(self.record), C.sizeof(self.record))
        if self.rad is not None:
            if self.rad_keys is None:
                raise RuntimeError("radiative mode: set_keys() before transport")
            hard.rad_key = self.rad_keys.data_ptr()
        for name, _ in self.row_columns:
            t = rows.get(name)
            setattr(hard, name, 0 if t is None else t.data_ptr())


def lut_for(lut, shell: ShellTables | None):
    """The transport LUT a mode runs on: the shell mode's stopping column is the soft stopping."""
    if shell is None:
        return lut
    from dataclasses import replace

    return replace(lut, dEds=np.ascontiguousarray(shell.soft_dEds))
