"""Detector direction tiling and observation geometry (CPU).

The reference's own checks of ``detector_directions``
(tests/detectors/test_solid_angle.py: n_side = 1 is the central direction
with the whole Omega and follows the tilt, weights sum to Omega and are
positive, unit directions, the weighted mean points at the centre, the
centre cell is heaviest, a wider chip spreads the polar angles, n_side >= 1),
plus closed forms: the cell weights scale as the cube of the cosine to the
chip normal (inverse square times obliquity, so a dropped obliquity fails),
the raw cell solid angles converge to the exact solid angle of a square chip,
``cone_detector`` is a chip of the cone's solid angle, and
``rotation_between`` / ``tilted_geometry`` on worked cases.
"""

import numpy as np
import pytest

from pyrite_amd.spectrum.geometry import cone_detector, detector_directions, rotation_between, tilted_geometry

THETA = np.deg2rad(119.0)
OMEGA = 0.066


def test_nside1_is_single_central_direction():
    n_hats, weights = detector_directions(THETA, n_side=1, domega_sr=OMEGA)
    assert n_hats.shape == (1, 3) and weights.shape == (1,)
    _, n_hat = tilted_geometry(THETA, 0.0)
    assert np.allclose(n_hats[0], n_hat)
    assert weights[0] == pytest.approx(OMEGA)


def test_nside1_tracks_tilt():
    tp, ta = np.deg2rad(20.0), np.deg2rad(35.0)
    n_hats, _ = detector_directions(THETA, tp, ta, n_side=1, domega_sr=OMEGA)
    _, n_hat = tilted_geometry(THETA, tp, ta)
    assert np.allclose(n_hats[0], n_hat)


@pytest.mark.parametrize("n_side", [1, 2, 3, 5, 8])
def test_weights_sum_to_total_omega(n_side):
    _, weights = detector_directions(THETA, n_side=n_side, domega_sr=OMEGA)
    assert weights.shape == (n_side * n_side,)
    assert weights.sum() == pytest.approx(OMEGA)
    assert np.all(weights > 0)


@pytest.mark.parametrize("n_side", [1, 3, 6])
def test_directions_are_unit_vectors(n_side):
    n_hats, _ = detector_directions(THETA, n_side=n_side, domega_sr=OMEGA)
    assert np.allclose(np.linalg.norm(n_hats, axis=1), 1.0)


def test_grid_is_centred_on_central_direction():
    n_hats, weights = detector_directions(THETA, n_side=6, domega_sr=OMEGA)
    mean = (weights[:, None] * n_hats).sum(axis=0)
    mean /= np.linalg.norm(mean)
    assert np.allclose(mean, [np.sin(THETA), 0.0, np.cos(THETA)], atol=2e-3)


def test_centre_cell_has_largest_weight():
    n_side = 5
    _, weights = detector_directions(THETA, n_side=n_side, domega_sr=OMEGA)
    assert np.argmax(weights) == (n_side // 2) * n_side + (n_side // 2)


def test_wider_chip_spreads_polar_angles():
    def polar_spread(chip_mm):
        n_hats, _ = detector_directions(THETA, n_side=5, chip_mm=chip_mm, dist_mm=30.0, domega_sr=OMEGA)
        polar = np.arccos(np.clip(n_hats[:, 2], -1.0, 1.0))
        return polar.max() - polar.min()

    assert polar_spread(30.0) > polar_spread(5.0)


def test_nside_must_be_positive():
    with pytest.raises(ValueError, match=r"n_side must be >= 1"):
        detector_directions(THETA, n_side=0, domega_sr=OMEGA)


def test_cell_weights_scale_as_cosine_cubed():
    """Cell solid angle = pitch^2 cos / r^2 = pitch^2 cos^3 / D^2 for a cell at
    distance r from the source, D the chip distance, cos = D/r: corner and edge
    cells of a 3x3 chip against the centre cell."""
    chip, dist = 14.0, 30.0
    n_hats, w = detector_directions(THETA, n_side=3, chip_mm=chip, dist_mm=dist, domega_sr=1.0)
    s = chip / 3.0
    centre = w[4]
    edge = (dist / np.hypot(dist, s)) ** 3
    corner = (dist / np.sqrt(dist**2 + 2 * s * s)) ** 3
    np.testing.assert_allclose(w[[1, 3, 5, 7]] / centre, edge, rtol=1e-12)
    np.testing.assert_allclose(w[[0, 2, 6, 8]] / centre, corner, rtol=1e-12)
    # the rays point at the cell centres: angle to the chip axis atan(s/D) and atan(s sqrt2/D)
    axis = n_hats[4]
    np.testing.assert_allclose(np.degrees(np.arccos(n_hats[[1, 3, 5, 7]] @ axis)),
                               np.degrees(np.arctan(s / dist)), rtol=1e-9)
    np.testing.assert_allclose(np.degrees(np.arccos(n_hats[[0, 2, 6, 8]] @ axis)),
                               np.degrees(np.arctan(s * np.sqrt(2) / dist)), rtol=1e-9)


def test_raw_cell_solid_angles_converge_to_the_square_chip():
    """The unnormalised cell solid angles (the weights times their raw sum) add up
    to the exact solid angle of a square of side c at distance D,
    4 asin(a^2 / (a^2 + D^2)) with a = c/2, as the tiling is refined."""
    chip, dist = 14.0, 30.0
    a = chip / 2.0
    exact = 4.0 * np.arcsin(a * a / (a * a + dist * dist))
    errs = []
    for n in (4, 16, 64):
        step = chip / n
        off = (np.arange(n) - (n - 1) / 2.0) * step
        r2 = dist**2 + off[:, None] ** 2 + off[None, :] ** 2
        raw = (step * step * dist / r2**1.5).sum()
        _, w = detector_directions(0.3, n_side=n, chip_mm=chip, dist_mm=dist, domega_sr=raw)
        # the tiling's relative weights are the raw cell solid angles
        np.testing.assert_allclose(np.sort(w), np.sort((step * step * dist / r2**1.5).ravel()), rtol=1e-10)
        errs.append(abs(raw - exact) / exact)
    assert errs[2] < 1e-4 and errs[0] > errs[1] > errs[2]


def test_cone_detector_is_a_chip_of_the_cone_solid_angle():
    half = 0.01
    n_hats, w = cone_detector(np.pi / 2, np.deg2rad(0.5), half, 3)
    assert w.sum() == pytest.approx(np.pi * half**2)
    _, n_hat = tilted_geometry(np.pi / 2, np.deg2rad(0.5))
    ang = np.arccos(np.clip(n_hats @ n_hat, -1.0, 1.0))
    side = np.sqrt(np.pi) * half  # chip side at unit distance
    assert ang[4] < 1e-12
    np.testing.assert_allclose(ang[[1, 3, 5, 7]], np.arctan(side / 3.0), rtol=1e-6)


def test_rotation_between_and_tilted_geometry():
    rng = np.random.default_rng(5)
    for _ in range(20):
        u, t = rng.normal(size=3), rng.normal(size=3)
        u, t = u / np.linalg.norm(u), t / np.linalg.norm(t)
        R = rotation_between(u, t)
        np.testing.assert_allclose(R @ u, t, atol=1e-12)
        np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-12)
        assert np.linalg.det(R) == pytest.approx(1.0)
    R = rotation_between(np.array([0.0, 0.0, 1.0]), np.array([0.0, 0.0, -1.0]))
    np.testing.assert_allclose(R @ [0, 0, 1], [0, 0, -1], atol=1e-12)
    assert np.linalg.det(R) == pytest.approx(1.0)
    beam, n = tilted_geometry(np.pi / 2, 0.0)
    np.testing.assert_allclose(beam, [0, 0, 1])
    np.testing.assert_allclose(n, [1, 0, 0], atol=1e-15)
    # a tilt of the sample by t about y turns both lab directions by -t in the sample frame
    t = np.deg2rad(0.5)
    beam, n = tilted_geometry(np.pi / 2, t)
    np.testing.assert_allclose(beam, [-np.sin(t), 0, np.cos(t)], atol=1e-15)
    np.testing.assert_allclose(n, [np.cos(t), 0, np.sin(t)], atol=1e-15)
    assert beam @ n == pytest.approx(0.0, abs=1e-15)
