"""Numpy restatement of the reference's analytic beam entry (TEST INFRASTRUCTURE ONLY).

Follows src/pyrite/montecarlo/transport/kinematics.py:267-312 (counter_uniforms,
counter_normals), transport/beam_entry.py:23-126,347-359 (initial_beam_positions,
primary_directions), transverse.py:169-224 (sample_transverse) and
geometry.py:374-394 (_project_entry_rows), in the reference's operation order.
The normal quantile is scipy.special.ndtri, the function the reference calls.
Pins: the SplitMix64 known-answer keys (stream construction), sample moments
against the Courant-Snyder closed forms, and the projection's p0_z = 0 and
zero-tilt identity.
"""

from __future__ import annotations

import numpy as np
from scipy.special import ndtri

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)


def _sm64(x):
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        x = (x ^ (x >> np.uint64(30))) * M1
        x = (x ^ (x >> np.uint64(27))) * M2
    return x ^ (x >> np.uint64(31))


def counter_uniforms(root, n, n_draws, start=0):
    e = np.arange(start, start + n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        key = _sm64(np.uint64(root) + GOLDEN * (e + np.uint64(1)))
        out = np.empty((n, n_draws))
        for c in range(n_draws):
            z = _sm64(key + GOLDEN * np.uint64(c + 1))
            out[:, c] = ((z >> np.uint64(12)).astype(np.float64) + 0.5) * 2.0**-52
    return out


def counter_normals(root, n, n_draws, start=0):
    return ndtri(counter_uniforms(root, n, n_draws, start))


def entry(args, n, start=0):
    """``(pos, dirs)`` for the fields of a ``pyrite_amd._lib.BeamArgs``."""
    mode = int(args.mode)
    pos = np.zeros((n, 3))
    xp = yp = None
    if mode == 1:
        z = counter_normals(args.root, n, 2, start)
        off = np.stack((z[:, 0] * args.sigma_x, z[:, 1] * args.sigma_y), axis=1)
    elif mode == 2:
        z = counter_normals(args.root, n, 4, start)
        tx, ty = list(args.twiss_x), list(args.twiss_y)
        x_mm = tx[0] * z[:, 0]
        xp = tx[1] * (z[:, 1] - tx[2] * z[:, 0])
        y_mm = ty[0] * z[:, 2]
        yp = ty[1] * (z[:, 3] - ty[2] * z[:, 2])
        off = np.stack((x_mm * 1.0e7, y_mm * 1.0e7), axis=1)
    else:
        off = np.zeros((n, 2))
    ex, ey, b = np.array(list(args.e_x)), np.array(list(args.e_y)), np.array(list(args.beam_dir))
    if args.projected:  # along the lab beam axis of the tilt
        ax = np.array(list(args.axis))
        u, v = off[:, 0], off[:, 1]
        s = -(u * ex[2] + v * ey[2]) / ax[2]
        pos[:, 0] = (u * ex[0] + v * ey[0]) + s * ax[0]
        pos[:, 1] = (u * ex[1] + v * ey[1]) + s * ax[1]
    else:
        pos[:, :2] = off
    if mode == 2:
        basis = np.array(list(args.basis)).reshape(3, 3)
        dirs = xp[:, None] * basis[:, 0] + yp[:, None] * basis[:, 1] + b
        dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    else:
        dirs = np.tile(b, (n, 1))
    return pos, dirs


def energies(args, n, start=0):
    """Incident energies for the fields of a ``BeamArgs`` (beam_entry.py:264-297):
    E0 (1 + f z), z draw 0 under the energy root clamped to |z| <= 6."""
    E = np.full(n, float(args.E0_keV))
    f = float(args.energy_spread_frac)
    if f:
        z = np.clip(counter_normals(args.energy_root, n, 1, start)[:, 0], -6.0, 6.0)
        E = E * (1.0 + f * z)
    return E
