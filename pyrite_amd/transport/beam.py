"""Incident beam on the entrance face: pencil, Gaussian spot or Twiss beam.

Host side of ``hr_beam_entry``.  Mirrors the reference's ungrooved analytic
entry (``montecarlo/transport/beam_entry.py:23-126`` ``initial_beam_positions``,
``:347-359`` ``primary_directions``; ``montecarlo/transverse.py:117-224``
Courant-Snyder resolution and sampling; ``montecarlo/geometry.py:330-413``
entry projection, ``:575-609`` beam-frame basis).  The host resolves the
scalar parameters (child-stream root, Twiss moments, rotations); the device
draws every electron's state, so a 1e8-electron beam never crosses PCIe.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from .. import _lib
from ..spectrum.geometry import rotation_between

MC2_KEV = 510.99895
FWHM_TO_SIGMA = 1.0 / (2.0 * np.sqrt(2.0 * np.log(2.0)))
#: |z| clamp of the energy-spread normal (``beam_entry.py:244-248``)
ENERGY_SPREAD_CLAMP_SIGMA = 6.0


def child_stream_root(seed: int, n_children: int, index: int) -> int:
    """Root of the ``SeedSequence(seed).spawn(n)[index]`` namespace (``kinematics.py:255-264``)."""
    child = np.random.SeedSequence(seed).spawn(n_children)[index]
    return int(child.generate_state(1, np.uint64)[0])


def upper_beam_energy_keV(E0_keV, energy_spread_frac=None) -> float:
    """Largest incident energy the spread can draw, ``E0 (1 + 6 |f|)``
    (``beam_entry.py:251-261``): the top of the transport tables' range."""
    return float(E0_keV) * (1.0 + ENERGY_SPREAD_CLAMP_SIGMA * abs(float(energy_spread_frac or 0.0)))


def beam_entry_basis(tilt_polar_rad, tilt_azim_rad=0.0):
    """Sample-frame lab spot axes and nominal beam axis ``(e_x, e_y, b)`` (``geometry.py:396-413``)."""
    if not tilt_polar_rad:
        eye = np.eye(3)
        return eye[:, 0], eye[:, 1], eye[:, 2]
    st, ct = np.sin(tilt_polar_rad), np.cos(tilt_polar_rad)
    normal_lab = np.array([st * np.cos(tilt_azim_rad), st * np.sin(tilt_azim_rad), ct])
    Rt = rotation_between(np.array([0.0, 0.0, 1.0]), normal_lab).T
    return Rt[:, 0], Rt[:, 1], Rt[:, 2]


def beam_frame_basis(beam_dir):
    """Shortest-arc rotation from +z to ``beam_dir`` (``geometry.py:575-609``)."""
    d = np.asarray(beam_dir, dtype=float)
    d = d / np.linalg.norm(d)
    sin_theta = float(np.hypot(d[0], d[1]))
    if sin_theta < 1e-15:
        return np.eye(3)
    theta = float(np.arctan2(sin_theta, d[2]))
    dx, dy = -d[1] * theta / sin_theta, d[0] * theta / sin_theta
    ang = float(np.hypot(dx, dy))
    kx, ky = dx / ang, dy / ang
    K = np.array([[0.0, 0.0, ky], [0.0, 0.0, -kx], [-ky, kx, 0.0]])
    return np.eye(3) + np.sin(ang) * K + (1.0 - np.cos(ang)) * (K @ K)


@dataclass(frozen=True)
class TwissPlane:
    """One plane's Courant-Snyder sampling factors (``transverse.py:99-114,169-178``)."""

    sigma_pos_mm: float  # sqrt(eps beta)
    slope_scale_rad: float  # sqrt(eps / beta)
    alpha: float

    @staticmethod
    def from_normalized(emittance_n_mm_mrad, beta_twiss_m, alpha, E0_keV):
        gamma_rel = 1.0 + float(E0_keV) / MC2_KEV
        beta_gamma = float(np.sqrt(gamma_rel**2 - 1.0))
        eps = emittance_n_mm_mrad * 1.0e-3 / beta_gamma
        beta = beta_twiss_m * 1.0e3
        return TwissPlane(float(np.sqrt(eps * beta)), float(np.sqrt(eps / beta)), float(alpha))


@dataclass(frozen=True)
class BeamSpec:
    """Incident beam of a case.

    ``kind``: ``"pencil"``; ``"spot"`` (Gaussian spot of ``fwhm_mm`` /
    ``fwhm_y_mm``, collimated); ``"twiss"`` (planes ``twiss_x``, ``twiss_y``).
    ``energy_spread_frac``: RMS relative energy spread, ``E = E0 (1 + f z)``
    with ``z`` a standard normal clamped to ``|z| <= 6`` drawn in its own child
    namespace (``beam_entry.py:264-297``); 0 is the monoenergetic beam.
    """

    kind: str = "pencil"
    fwhm_mm: float = 0.0
    fwhm_y_mm: float | None = None
    twiss_x: TwissPlane | None = None
    twiss_y: TwissPlane | None = None
    energy_spread_frac: float = 0.0

    @staticmethod
    def gaussian_waist(sigma_mm, divergence_rad, E0_keV):
        """Round Twiss beam at its waist with rms size ``sigma_mm`` and rms
        divergence ``divergence_rad``: eps = sigma sigma', beta = sigma / sigma'."""
        eps_geom = float(sigma_mm) * float(divergence_rad)  # mm rad
        beta_m = float(sigma_mm) / float(divergence_rad) * 1.0e-3
        gamma_rel = 1.0 + float(E0_keV) / MC2_KEV
        eps_n = eps_geom * float(np.sqrt(gamma_rel**2 - 1.0)) * 1.0e3
        plane = TwissPlane.from_normalized(eps_n, beta_m, 0.0, E0_keV)
        return BeamSpec(kind="twiss", twiss_x=plane, twiss_y=plane)


def beam_args(spec: BeamSpec, seed, start, n, E0_keV, tilt_polar_rad, tilt_azim_rad=0.0, beam_dir=None):
    """``hr_beam_args`` (outputs unset) for electrons ``[start, start+n)``.

    The lab spot is carried to the entrance face along the tilt's beam axis;
    the directions follow ``beam_dir`` (default: that axis), as the
    reference's flat entry takes a slab-frame ``beam_dir`` beside the tilt
    (``beam_entry.py:347-359``, ``geometry.py:330-394``)."""
    a = _lib.BeamArgs()
    a.start, a.n = int(start), int(n)
    e_x, e_y, b = beam_entry_basis(tilt_polar_rad, tilt_azim_rad)
    if spec.kind == "pencil":
        a.mode = 0
    elif spec.kind == "spot":
        a.mode = 1
        a.root = child_stream_root(seed, 2, 1)
        fwhm_y = spec.fwhm_mm if spec.fwhm_y_mm is None else spec.fwhm_y_mm
        a.sigma_x = float(spec.fwhm_mm or 0.0) * 1.0e7 * FWHM_TO_SIGMA
        a.sigma_y = float(fwhm_y or 0.0) * 1.0e7 * FWHM_TO_SIGMA
    elif spec.kind == "twiss":
        a.mode = 2
        a.root = child_stream_root(seed, 5, 4)
        for dst, pl in ((a.twiss_x, spec.twiss_x), (a.twiss_y, spec.twiss_y)):
            dst[0], dst[1], dst[2] = pl.sigma_pos_mm, pl.slope_scale_rad, pl.alpha
    else:
        raise ValueError(f"unknown beam kind {spec.kind!r}")
    a.projected = int(bool(tilt_polar_rad) and a.mode != 0)
    d = b if beam_dir is None else np.asarray(beam_dir, dtype=np.float64)  # a unit vector from the caller
    for dst, src in ((a.e_x, e_x), (a.e_y, e_y), (a.beam_dir, d), (a.axis, b)):
        dst[0], dst[1], dst[2] = (float(v) for v in src)
    basis = beam_frame_basis(d)
    for k in range(9):
        a.basis[k] = float(basis.reshape(-1)[k])
    a.E0_keV = float(E0_keV)
    f = float(spec.energy_spread_frac or 0.0)
    if not np.isfinite(f) or abs(f) * ENERGY_SPREAD_CLAMP_SIGMA >= 1.0:
        raise ValueError(f"energy_spread_frac={f} could draw a non-finite or non-positive electron energy; the "
                         "Gaussian spread model needs spread << 1")
    a.energy_spread_frac = f
    a.energy_root = child_stream_root(seed, 6, 5) if f else 0
    return a


def fill_entry(spec: BeamSpec, seed, start, n, E0_keV, tilt_polar_rad, pos, dirs, E_keV, clock,
               tilt_azim_rad=0.0, stream=None, beam_dir=None):
    """Write entry states of electrons ``[start, start+n)`` into device tensors."""
    a = beam_args(spec, seed, start, n, E0_keV, tilt_polar_rad, tilt_azim_rad, beam_dir)
    a.pos, a.dirs, a.E_keV, a.clock = pos.data_ptr(), dirs.data_ptr(), E_keV.data_ptr(), clock.data_ptr()
    if stream is None:
        import torch

        stream = torch.cuda.current_stream().cuda_stream
    _lib.check(_lib.load().hr_beam_entry(C.byref(a), C.c_void_p(stream)), "hr_beam_entry")


def plane_from_resolved(plane) -> TwissPlane:
    """A ``TwissPlane`` from the reference's resolved plane mapping
    (``transverse.py:74-87``: geometric emittance [mm rad], beta [mm/rad], alpha)."""
    eps = float(plane["geometric_emittance_mm_rad"])
    beta = float(plane["beta_twiss_mm_per_rad"])
    return TwissPlane(float(np.sqrt(eps * beta)), float(np.sqrt(eps / beta)), float(plane["alpha_twiss"]))


def flat_beam_spec(beam_fwhm_mm=None, beam_fwhm_y_mm=None, transverse_distribution=None,
                   energy_spread_frac=None) -> BeamSpec:
    """The ``BeamSpec`` of the reference's flat beam keywords
    (``beam_entry.py:23-80``): a resolved Courant-Snyder mapping, else a
    Gaussian spot of FWHM ``beam_fwhm_mm`` (``beam_fwhm_y_mm`` for y), else
    the pencil."""
    f = float(energy_spread_frac or 0.0)
    if transverse_distribution is not None:
        if beam_fwhm_mm or beam_fwhm_y_mm:
            raise ValueError("transverse_distribution is incompatible with the spot FWHM fields")
        td = transverse_distribution
        return BeamSpec(kind="twiss", twiss_x=plane_from_resolved(td["x"]), twiss_y=plane_from_resolved(td["y"]),
                        energy_spread_frac=f)
    if beam_fwhm_mm or beam_fwhm_y_mm:
        return BeamSpec(kind="spot", fwhm_mm=float(beam_fwhm_mm or 0.0),
                        fwhm_y_mm=None if beam_fwhm_y_mm is None else float(beam_fwhm_y_mm), energy_spread_frac=f)
    return BeamSpec(energy_spread_frac=f)
