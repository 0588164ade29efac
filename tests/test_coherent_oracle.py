"""The coherent line route of the oracle pinned by golden vectors and mathematics (CPU).

What the reference fixes for this route (cxr/_batched.py:561-690,
cxr/_formation.py:66-115, cxr/_recoil.py:118-147,
cxr/_batched_finalize.py:201-235):
  * the field/square ordering: the worked two-energy block of the coherent
    reduction (jit-spectrum-kernel-walkthrough.md:167-211) gives
    I = (5 + 20/pi^2)/4 = 1.756606 in both bins;
  * the formation factor is exp(-tau_c/2) sinh(w)/w, w = i v - q, i.e. the
    average of exp(-tau(s)/2 + i v s) over s in [-1, 1] with tau linear
    (checked against quadrature), and Parseval then gives its squared
    integral as pi times the mean transmission;
  * the coherent amplitude c = sqrt(alpha omega / (4 pi^2 hbar c)) t_L A and
    the incoherent weight w = |A|^2 ... T_abs wm share A, so |c|^2 T_abs wm = w
    line by line, and in the vacuum limit (delta = 0) the exact centre and
    width equal the incoherent line's E_r and a;
  * "the pieces of a straight flight sum exactly to its field"
    (cxr/_setup.py:910-918, _recoil.py:130-134): splitting a straight,
    constant-energy segment into collinear pieces with consistent ages leaves
    the field unchanged, with refraction, absorption and recoil on.  That
    fixes the phase slope, lattice phase, recoil shifts, width, escape terms
    and attenuation coefficients against one another; two planted mistakes
    show the check has teeth.
"""

import numpy as np
import pytest

from oracle import oracle as O

HBARC = 1973.269804


@pytest.fixture(scope="module")
def hopg():
    from pyrite_amd.pipeline import CaseConfig, CaseTables
    from pyrite_amd.spectrum.crystal import CRYSTALS, decrement, polarization_pairs

    cfg = CaseConfig(Ne=10)
    t = CaseTables(cfg)
    tab = t.line_tables.as_dict()
    n_hat = t.n_hats[4]
    es, ep = polarization_pairs(n_hat, tab["g"])
    E = np.linspace(1000.0, 50000.0, 12000)
    dom = decrement(CRYSTALS["hopg"], E) * E / HBARC
    return dict(cfg=cfg, t=t, tab=tab, n_hat=n_hat, es=es, ep=ep, E=E, dom=dom)


def _straight(rng, beam_dir, n_split, Ek=None, L=None):
    v = beam_dir + 0.1 * rng.normal(size=3)
    v /= np.linalg.norm(v)
    Ek = rng.uniform(500.0, 5000.0) if Ek is None else Ek
    L = rng.uniform(100.0, 5000.0) if L is None else L
    start = np.array([0.0, 0.0, 2.0e5]) + 100.0 * rng.normal(size=3)
    t0 = rng.uniform(0.0, 1.0e5)
    gam = 1.0 + Ek / 510.99895
    beta = np.sqrt(1.0 - 1.0 / gam**2)

    def seg(ns):
        Ls = L / ns
        return {"r_mid": np.array([start + (i + 0.5) * Ls * v for i in range(ns)]),
                "v_hat": np.tile(v, (ns, 1)), "L": np.full(ns, Ls), "E": np.full(ns, Ek),
                "t0": t0 + np.arange(ns) * Ls / beta}

    return [seg(ns) for ns in n_split]


def test_worked_two_energy_block_gives_the_published_intensity():
    # jit-spectrum-kernel-walkthrough.md:173-211: vacuum phase, w_m = 1/4
    rec = np.zeros((2, 13))
    rec[:, 1] = np.pi / 2  # a
    rec[:, 4], rec[:, 6] = 1.0, 2.0  # C_s, C_p
    rec[:, 10] = 2.0  # apb = a + b with no absorption: the formation factor is sin v / v
    rec[1, 0] = 1.0  # E_r
    rec[1, 3] = -np.pi / 2  # g
    F = O.coherent_field(rec, [0.0, 1.0], [0.0, 0.0])
    I = 0.25 * (F**2).sum(0)
    np.testing.assert_allclose(I, 0.25 * (5.0 + 20.0 / np.pi**2), rtol=1e-14)
    assert abs(I[0] - 1.756606) < 5e-7
    # the square after the sum: the in-phase pair doubles the field, the
    # incoherent sum of squares only doubles the intensity
    rec[1, 3] = 0.0
    rec[1, 0] = 0.0
    F2 = O.coherent_field(rec, [0.0], [0.0])
    np.testing.assert_allclose((F2**2).sum(), 4.0 * 5.0, rtol=1e-14)


@pytest.mark.parametrize("q", [0.0, 1e-4, 0.3, 0.49, 0.51, 2.0, 7.0])
def test_formation_factor_is_the_damped_piece_average(q):
    tau_c = 1.3
    a, b = np.exp(-0.5 * (tau_c - 2.0 * q)), np.exp(-0.5 * (tau_c + 2.0 * q))
    s, wq = np.polynomial.legendre.leggauss(200)
    for v in (-40.0, -3.0, -0.2, 0.0, 1e-4, 0.7, 5.0, 60.0):
        if q < 0.5:  # the small-q coefficients (formation_coefficients)
            apb, bma = 2.0 * np.exp(-0.5 * tau_c) * np.cosh(q), -2.0 * np.exp(-0.5 * tau_c) * np.sinh(q)
        else:
            apb, bma = a + b, b - a
        F = O.formation_factor(v, apb, bma, q)
        quad = 0.5 * np.sum(wq * np.exp(-0.5 * tau_c + (1j * v - q) * s))
        assert abs(F - quad) <= 2e-12 * max(1.0, abs(quad)), (v, F, quad)


def test_formation_factor_series_branch_is_continuous():
    for q, v in ((3e-4, 5e-4), (0.0, 9.99e-4), (6e-4, -7e-4)):
        apb, bma = 2.0 * np.cosh(q), -2.0 * np.sinh(q)
        inside = O.formation_factor(v, apb, bma, q)
        outside = O.formation_factor(v * 1.02, apb, bma, q)
        assert abs(inside - outside) < 1e-5
        s, wq = np.polynomial.legendre.leggauss(60)
        quad = 0.5 * np.sum(wq * np.exp((1j * v - q) * s))
        assert abs(inside - quad) < 1e-12


def test_parseval_squared_formation_integral_is_pi_times_mean_transmission():
    tau0, tau1 = 0.4, 2.2
    q = 0.25 * (tau1 - tau0)
    apb, bma = np.exp(-0.5 * tau0) + np.exp(-0.5 * tau1), np.exp(-0.5 * tau1) - np.exp(-0.5 * tau0)
    V = 400.0
    v = np.linspace(-V, V, 400_001)
    F2 = np.array([abs(O.formation_factor(x, apb, bma, q)) ** 2 for x in v[::50]])
    integral = np.trapezoid(F2, v[::50])
    # |F|^2 -> (apb^2 sin^2 + bma^2 cos^2)/(4 v^2) far out: the tails beyond +-V add (apb^2+bma^2)/(4V)
    integral += (apb**2 + bma**2) / (4.0 * V)
    T_mean = (np.exp(-tau0) - np.exp(-tau1)) / (tau1 - tau0)
    assert abs(integral - np.pi * T_mean) < 2e-4 * np.pi * T_mean


def test_coherent_amplitudes_and_width_match_the_incoherent_line(hopg):
    h = hopg
    rng = np.random.default_rng(11)
    n = 64
    v = h["t"].beam_dir + 0.2 * rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    seg = {"r_mid": np.c_[rng.normal(size=(n, 2)) * 50.0, rng.uniform(1e3, 9.9e5, n)], "v_hat": v,
           "L": rng.uniform(50.0, 3000.0, n), "E": rng.uniform(100.0, 5000.0, n), "t0": rng.uniform(0, 1e6, n)}
    seg32 = {k: np.asarray(seg[k], dtype=np.float32).astype(np.float64) for k in ("r_mid", "v_hat", "L", "E")}
    seg32["t0"] = seg["t0"]
    for vacuum in (False, True):
        tab = dict(h["tab"])
        if vacuum:
            tab["delta_tab"] = np.zeros_like(tab["delta_tab"])
        rec, good = O.coherent_lines(seg32, h["cfg"].thickness_ang, h["n_hat"], tab, h["es"], h["ep"])
        E_r, aw, w, g_inc = O.cxr_lines(seg32, h["cfg"].thickness_ang, h["n_hat"], tab, h["es"], h["ep"])
        good, rec = good[:, 0], rec[:, 0, :]
        assert good.sum() > n // 2
        np.testing.assert_array_equal(good, g_inc)
        r = rec[good]
        c2 = r[:, 4] ** 2 + r[:, 5] ** 2 + r[:, 6] ** 2 + r[:, 7] ** 2
        # attenuation coefficients: (b - a)/(a + b) = -tanh(q), a b = exp(-mu L_esc)
        np.testing.assert_allclose(r[:, 11] / r[:, 10], -np.tanh(r[:, 12]), rtol=1e-12, atol=1e-15)
        mu = 2.0 * r[:, 12] / r[:, 9]
        np.testing.assert_allclose(r[:, 10] ** 2 - r[:, 11] ** 2, 4.0 * np.exp(-mu * r[:, 8]), rtol=1e-10)
        p0, p1 = (r[:, 8] - r[:, 9]) * mu, (r[:, 8] + r[:, 9]) * mu
        T_abs = np.exp(-np.minimum(p0, p1)) * -np.expm1(-np.abs(p1 - p0)) / np.abs(p1 - p0)
        np.testing.assert_allclose(c2 * T_abs * h["tab"]["wm"][0], w[good], rtol=1e-10)
        if vacuum:
            np.testing.assert_allclose(r[:, 0], E_r[good], rtol=1e-12)
            np.testing.assert_allclose(r[:, 1], aw[good], rtol=1e-12)
        else:  # the in-medium root sits O(delta |grad L|) away (grazing exit: |grad L| ~ 100)
            assert np.all(np.abs(r[:, 0] / E_r[good] - 1.0) < 2e-2)


def test_straight_flight_pieces_sum_to_its_field(hopg):
    h = hopg
    rng = np.random.default_rng(5)
    done = 0
    for _ in range(30):
        segs = _straight(rng, h["t"].beam_dir, (1, 2, 7))
        fields = []
        for seg in segs:
            rec, good = O.coherent_lines(seg, h["cfg"].thickness_ang, h["n_hat"], h["tab"], h["es"], h["ep"])
            fields.append((rec[:, 0, :], good.all()))
        if not all(g for _, g in fields):
            continue
        F = [O.coherent_field(r, h["E"], h["dom"]) for r, _ in fields]
        scale = np.abs(F[0]).max()
        assert scale > 0
        for Fi in F[1:]:
            assert np.abs(Fi - F[0]).max() < 1e-9 * scale
        # planted mistakes: drop the in-piece refraction term, or take the
        # escape distance at the piece start instead of its mean
        r7 = fields[2][0].copy()
        r7[:, 9] = 0.0
        assert np.abs(O.coherent_field(r7, h["E"], h["dom"]) - F[0]).max() > 1e-3 * scale
        r7 = fields[2][0].copy()
        r7[:, 8] -= fields[2][0][:, 9]
        assert np.abs(O.coherent_field(r7, h["E"], h["dom"]) - F[0]).max() > 1e-3 * scale
        done += 1
    assert done >= 6


def test_identical_pieces_add_their_fields():
    rng = np.random.default_rng(2)
    rec = np.zeros((1, 13))
    rec[0, :4] = [3000.0, 0.02, 14.0, 3.0]
    rec[0, 4:8] = rng.normal(size=4)
    rec[0, 8:] = [1e4, 200.0, 1.7, -0.1, 0.05]
    E = np.linspace(2900.0, 3100.0, 64)
    dom = 1e-6 * np.ones_like(E)
    one = O.coherent_field(rec, E, dom)
    five = O.coherent_field(np.repeat(rec, 5, axis=0), E, dom)
    np.testing.assert_allclose(five, 5.0 * one, rtol=1e-13, atol=1e-13 * np.abs(one).max())
    rec[0, 4:8] = 0.0  # a rejected line adds nothing, even with a NaN centre
    rec[0, 0] = np.nan
    assert not O.coherent_field(rec, E, dom).any()
