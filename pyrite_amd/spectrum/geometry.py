"""Observation geometry in the sample frame.

* ``rotation_between`` -- minimal (Rodrigues) rotation, ``materials/crystal.py:894-916``;
* ``tilted_geometry`` -- beam and detector directions for a tilted slab,
  ``montecarlo/geometry.py:263-305``;
* ``detector_directions`` -- a flat square chip tiled into pixel rays with
  inverse-square/obliquity solid angles rescaled to the detector's total
  ``Omega``, ``montecarlo/geometry.py:450-495`` with the ray construction of
  ``_planar_geometry.py:468-515`` (local +x toward larger polar angle, +y
  toward larger azimuth, pixel offsets ``(i - (n-1)/2) * pitch``).
"""

from __future__ import annotations

import numpy as np


def rotation_between(u_hat, t_hat):
    """Minimal rotation ``R`` with ``R @ u_hat = t_hat`` (unit vectors)."""
    u_hat = np.asarray(u_hat, dtype=float)
    t_hat = np.asarray(t_hat, dtype=float)
    c = float(np.dot(u_hat, t_hat))
    axis = np.cross(u_hat, t_hat)
    s = float(np.linalg.norm(axis))
    if s < 1e-12:
        if c > 0:
            return np.eye(3)
        tmp = np.array([1.0, 0.0, 0.0]) if abs(u_hat[0]) < 0.9 else np.array([0.0, 1.0, 0.0])
        axis = np.cross(u_hat, tmp)
        axis /= np.linalg.norm(axis)
        K = np.array([[0.0, -axis[2], axis[1]], [axis[2], 0.0, -axis[0]], [-axis[1], axis[0], 0.0]])
        return np.eye(3) + 2.0 * K @ K
    axis = axis / s
    K = np.array([[0.0, -axis[2], axis[1]], [axis[2], 0.0, -axis[0]], [-axis[1], axis[0], 0.0]])
    return np.eye(3) + s * K + (1.0 - c) * (K @ K)


def _sample_to_lab(tilt_polar_rad, tilt_azim_rad):
    st, ct = np.sin(tilt_polar_rad), np.cos(tilt_polar_rad)
    normal_lab = np.array([st * np.cos(tilt_azim_rad), st * np.sin(tilt_azim_rad), ct])
    return rotation_between(np.array([0.0, 0.0, 1.0]), normal_lab)


def tilted_geometry(theta_obs_rad, tilt_polar_rad, tilt_azim_rad=0.0):
    """``(beam_dir, n_hat)`` in the sample frame for a tilted sample."""
    R = _sample_to_lab(tilt_polar_rad, tilt_azim_rad)
    beam_dir = R.T @ np.array([0.0, 0.0, 1.0])
    n_hat = R.T @ np.array([np.sin(theta_obs_rad), 0.0, np.cos(theta_obs_rad)])
    return beam_dir, n_hat


def detector_directions(theta_obs_rad, tilt_polar_rad=0.0, tilt_azim_rad=0.0, *, n_side=1,
                        chip_mm=14.0, dist_mm=30.0, domega_sr):
    """``(n_hats (N,3), weights (N,))`` tiling a square chip; weights sum to ``domega_sr``."""
    if n_side < 1:
        raise ValueError("n_side must be >= 1")
    R = _sample_to_lab(tilt_polar_rad, tilt_azim_rad)
    th = float(theta_obs_rad)
    center = dist_mm * np.array([np.sin(th), 0.0, np.cos(th)])
    x_axis = np.array([np.cos(th), 0.0, -np.sin(th)])
    y_axis = np.array([0.0, 1.0, 0.0])
    normal = center / np.linalg.norm(center)
    step = chip_mm / n_side
    off = (np.arange(n_side) - (n_side - 1) / 2.0) * step
    centers = center[None, None, :] + off[:, None, None] * y_axis[None, None, :] + off[None, :, None] * x_axis[None, None, :]
    dist = np.linalg.norm(centers, axis=-1)
    dirs = centers / dist[..., None]
    obliquity = dirs @ normal
    solid = step * step * obliquity / dist**2
    n_hats = dirs.reshape(-1, 3) @ R
    w = solid.reshape(-1).copy()
    w *= domega_sr / w.sum()
    return n_hats, w


def cone_detector(theta_obs_rad, tilt_polar_rad, half_angle_rad, n_side, tilt_azim_rad=0.0):
    """A detector cone of ``half_angle_rad`` as a square chip of equal solid angle.

    The chip side subtends ``sqrt(pi) * half_angle`` at unit distance, so its
    area equals the cone's disc; weights are rescaled to ``Omega = pi half^2``
    (small-angle cone).
    """
    omega = np.pi * half_angle_rad**2
    dist = 1000.0
    chip = np.sqrt(np.pi) * half_angle_rad * dist
    return detector_directions(theta_obs_rad, tilt_polar_rad, tilt_azim_rad, n_side=n_side,
                               chip_mm=chip, dist_mm=dist, domega_sr=omega)
