"""Detector-cone splitting: the FP64 oracle's estimator and the host-side records (CPU).

Mirrors the reference's own checks (tests/montecarlo/test_cone_split.py and
test_cone_split_weights.py): single-event weights preserve the cap integrals
of the analog law, the energy gate leaves transport analog bit for bit,
primaries follow their analog streams until cone entry, copy records are
prefix-stable in the electron count, and split tallies (primary plus
weighted copies per history) match analog within statistical error while the
in-cone tally's variance drops.  The analog law here is screened Rutherford,
so the single-event check is also run at finite screening, against the
law's band integral by quadrature.
"""

import math

import numpy as np
import pytest
from scipy import integrate

from pyrite_amd.cone_split import (
    CONE_SPLIT_MODEL,
    ConeSplit,
    ConeSplitting,
    history_sums,
    weighted_history_statistics,
)
from tests.transport_cases import Case

THETA_C = 0.4


def _cone(n_hat=(0.0, 0.0, 1.0), theta=THETA_C, target=0.05, angle_scale=0.0, **kw):
    return dict(n_hat=n_hat, half_angle_rad=theta, target_weight=target, angle_scale=angle_scale, **kw)


def _event_scores(oracle, direction, alpha, theta, target, angle_scale, count, E_keV=5.0):
    keys = [oracle.cone_event_key(918273, i) for i in range(count)]
    made, out = oracle.cone_event(np.tile(direction, (count, 1)), alpha, E_keV, keys,
                                  _cone(theta=theta, target=target, angle_scale=angle_scale))
    o, w = out[:, :3], out[:, 3]
    assert np.all(w[made] > 0.0)
    assert np.all(np.abs(np.linalg.norm(o[made], axis=1) - 1.0) < 1e-12)
    assert np.all(o[made, 2] >= np.cos(theta) - 1e-12)  # inside the cap about n = z
    w = np.where(made, w, 0.0)
    return np.stack([w, w * o[:, 2], w * (o[:, 2] >= np.cos(theta / 2.0))], axis=1)


@pytest.mark.parametrize("direction", [(0.0, 0.0, 1.0), (1.0, 0.0, 0.0), (0.0, 0.0, -1.0)])
@pytest.mark.parametrize("target", [0.001, 0.2])
@pytest.mark.parametrize("angle_scale", [0.0, 0.05])
def test_single_event_weights_preserve_isotropic_cap_integrals(oracle, direction, target, angle_scale):
    """Screening alpha -> infinity makes the law isotropic: the weights then
    integrate 1, cos(theta_out) and a smaller cap under dOmega/(4 pi)
    (the reference's test_cone_split_weights.py numbers)."""
    scores = _event_scores(oracle, np.asarray(direction, dtype=float), 1e15, THETA_C, target, angle_scale, 8192)
    expected = np.array([(1.0 - np.cos(THETA_C)) / 2.0, np.sin(THETA_C) ** 2 / 4.0,
                         (1.0 - np.cos(THETA_C / 2.0)) / 2.0])
    se = scores.std(axis=0, ddof=1) / np.sqrt(scores.shape[0])
    assert np.all(np.abs(scores.mean(axis=0) - expected) <= 6.0 * se + 1e-14)


def _sr_cap_integrals(alpha, cos_dn, theta_c):
    """Integrals of 1 and of Omega.n over the cap under the screened-Rutherford
    law about an incident direction at angle acos(cos_dn) to n, by quadrature
    over the polar angle of the deflection (the azimuth arc in closed form)."""
    sin_dn = math.sqrt(max(0.0, 1.0 - cos_dn * cos_dn))
    cos_c = math.cos(theta_c)
    th_dn = math.acos(cos_dn)
    lo, hi = max(0.0, th_dn - theta_c), min(math.pi, th_dn + theta_c)

    def pdf_theta(t):  # dF/dtheta, F(mu) = (1 + alpha) mu / (alpha + mu), mu = (1 - cos t)/2
        mu = 0.5 * (1.0 - math.cos(t))
        return alpha * (1.0 + alpha) / (alpha + mu) ** 2 * 0.5 * math.sin(t)

    def arc(t):
        s, c = math.sin(t), math.cos(t)
        if s * sin_dn <= 1e-300:
            return math.pi if c * cos_dn >= cos_c else 0.0
        x = (cos_c - c * cos_dn) / (s * sin_dn)
        return math.pi if x <= -1 else 0.0 if x >= 1 else math.acos(x)

    def mean_cos(t):  # mean of Omega.n over the arc |phi| <= arc(t)
        h = arc(t)
        if h <= 0.0:
            return 0.0
        s, c = math.sin(t), math.cos(t)
        return c * cos_dn + s * sin_dn * math.sin(h) / h

    p = integrate.quad(lambda t: pdf_theta(t) * arc(t) / math.pi, lo, hi, limit=400, epsabs=1e-14)[0]
    q = integrate.quad(lambda t: pdf_theta(t) * arc(t) / math.pi * mean_cos(t), lo, hi, limit=400,
                       epsabs=1e-14)[0]
    return p, q


@pytest.mark.parametrize("alpha", [0.02, 0.5])
@pytest.mark.parametrize("angle", [0.1, 1.0, 2.5])
@pytest.mark.parametrize("target", [1e-4, 0.2])
def test_single_event_weights_match_the_screened_rutherford_cap(oracle, alpha, angle, target):
    """At finite screening the mean copy weight is the law's probability of
    the cap and its cos-weighted mean the cap's first moment."""
    d = np.array([np.sin(angle), 0.0, np.cos(angle)])
    scores = _event_scores(oracle, d, alpha, THETA_C, target, 0.0, 20000)
    p, q = _sr_cap_integrals(alpha, float(np.cos(angle)), THETA_C)
    se = scores.std(axis=0, ddof=1) / np.sqrt(scores.shape[0])
    assert abs(scores[:, 0].mean() - p) <= 6.0 * se[0] + 1e-15, (scores[:, 0].mean(), p, se[0])
    assert abs(scores[:, 1].mean() - q) <= 6.0 * se[1] + 1e-15, (scores[:, 1].mean(), q, se[1])


def test_cap_quadrature_reduces_to_the_isotropic_cap():
    """The quadrature helper itself: alpha -> infinity is the isotropic cap."""
    for cos_dn in (1.0, 0.3, -1.0):
        p, q = _sr_cap_integrals(1e12, cos_dn, THETA_C)
        assert abs(p - (1.0 - math.cos(THETA_C)) / 2.0) < 1e-9
        assert abs(q - math.sin(THETA_C) ** 2 / 4.0) < 1e-9


def test_weight_ladder_keeps_expectation(oracle):
    assert oracle.cone_ladder_weight(0.25) == 0.25
    assert oracle.cone_ladder_weight(3.0) == 4.0
    assert oracle.cone_ladder_weight(0.3) == 0.5
    assert oracle.cone_ladder_weight(np.nextafter(1.0, 2.0)) == 2.0
    assert oracle.cone_ladder_weight(np.nextafter(1.0, 0.0)) == 1.0
    # E[W 1{u < w/W}] = w with u counter 4 of the event key
    w = 0.3
    u = np.array([oracle.stream_uniform(oracle.cone_event_key(77, f), 4) for f in range(40000)])
    kept = u < w / 0.5
    assert abs(kept.mean() * 0.5 - w) < 4 * 0.5 * np.sqrt(0.6 * 0.4 / u.size)


def test_history_sums_and_weighted_statistics():
    values = history_sums(np.array([1.0, 0.0, 2.0]), np.array([4.0, 2.0]), np.array([1, 1]),
                          np.array([0.25, 0.5]), 3)
    np.testing.assert_allclose(values, [1.0, 2.0, 2.0])
    stats = weighted_history_statistics(values)
    assert stats["mean"] == pytest.approx(5.0 / 3.0)
    assert stats["max_share"] == pytest.approx(0.4)
    assert stats["effective_sample_size"] == pytest.approx(25.0 / 9.0)
    assert stats["relative_se"] == pytest.approx(np.std(values, ddof=1) / (np.sqrt(3) * 5.0 / 3.0))
    with pytest.raises(ValueError, match="one value per history"):
        history_sums(np.zeros(2), [], [], [], 3)


def test_cone_split_validates_its_request():
    with pytest.raises(ValueError, match=r"n_hat must be a finite nonzero three-vector"):
        ConeSplit((0.0, 0.0, 0.0), THETA_C, 0.05)
    with pytest.raises(ValueError, match=r"half_angle_rad must lie in"):
        ConeSplit((1.0, 0.0, 0.0), 2.0, 0.05)
    with pytest.raises(ValueError, match=r"target_weight must be positive and finite"):
        ConeSplit((1.0, 0.0, 0.0), THETA_C, 0.0)
    with pytest.raises(ValueError, match=r"angle_scale must be positive"):
        ConeSplit((1.0, 0.0, 0.0), THETA_C, 0.05, angle_scale=0.0)
    c = ConeSplit((2.0, 0.0, 0.0), THETA_C, 0.05)
    assert c.n_hat == (1.0, 0.0, 0.0)


def test_policy_record():
    p = ConeSplitting()
    assert p.to_dict() == {"cone_factor": 8.0, "min_energy_fraction": 0.4, "angle_scale": 1.0,
                           "target_weight": 1e-3, "max_copies_per_electron": 1024, "model": CONE_SPLIT_MODEL}
    assert ConeSplitting.from_dict(p.to_dict()) == p
    with pytest.raises(ValueError, match="unknown cone split field"):
        ConeSplitting.from_dict({"cone_factr": 8.0})
    with pytest.raises(ValueError, match="unknown cone split model"):
        ConeSplitting(model="v0")
    with pytest.raises(ValueError, match="min_energy_fraction"):
        ConeSplitting(min_energy_fraction=1.0)
    gamma = 1.0 + 5000.0 / 510.99895
    assert p.half_angle_rad(5000.0) == pytest.approx(8.0 / gamma)
    with pytest.raises(ValueError, match="not below pi/2"):
        p.half_angle_rad(100.0)
    r = p.request((0.0, 1.0, 0.0), 5000.0)
    assert r.min_energy_keV == pytest.approx(2000.0) and r.weight_ladder and r.angle_scale == 1.0

    class Cfg:
        E0_keV = 5000.0
        coherent = False
        line_roulette = None
        depth_bins = 0

    p.validate_case(Cfg)
    for key, val in (("coherent", True), ("line_roulette", {"floor": 0.1}), ("depth_bins", 8)):
        bad = type("Bad", (Cfg,), {key: val})
        with pytest.raises(ValueError, match="does not support"):
            p.validate_case(bad)


# ---- transport with splitting (the oracle core) --------------------------

HBN_50KEV = Case("hbn50", 50.0, 300, [(0.0, 5.0e4, [("B", 0.0553), ("N", 0.0553)])], seed=3, E_cut=1.0,
                 finite=(5.0e5, 5.0e5))
N_HAT = (1.0, 0.0, 0.0)
# the same target on ELSEPA-format tables: N on the vendor DCS shape, B on the screened-Rutherford law's
HBN_50KEV_ELSEPA = Case("hbn50_elsepa", 50.0, 300, [(0.0, 5.0e4, [("B", 0.0553), ("N", 0.0553)])], seed=3,
                        E_cut=1.0, finite=(5.0e5, 5.0e5), elastic="elsepa", elsepa_energies=(0.8, 60.0, 40),
                        elsepa_vendor=("N",))


def _run(oracle, case, seed, ne, cone=None, keys=None, state=None, cap=4096):
    lt, lut = case.tables()
    if state is None:
        bd = np.asarray(case.beam_dir, dtype=float)
        pos, dirs = np.zeros((ne, 3)), np.tile(bd / np.linalg.norm(bd), (ne, 1))
        E, E_cut = np.full(ne, case.E0), np.full(ne, case.E_cut)
        clock = None
    else:
        pos, dirs, E, clock = state
        E_cut = np.full(len(E), case.E_cut)
    if keys is None:
        keys = oracle.stream_keys(seed, ne)
    return oracle.run_transport(lt, lut, keys, pos, dirs, E, E_cut, cap=cap, clock=clock,
                                energy_model=1, straggle_on=False, finite=case.finite, cone=cone)


def _tallies(rows, count, n, cap=4096):
    """Per-history in-cone and total path length (the reference's test tallies)."""
    ids = np.repeat(np.arange(n), np.minimum(count, cap))
    keep = (np.arange(cap)[None, :] < count[:, None]).reshape(-1)
    v = rows["seg_dir"][keep]
    L = rows["seg_len"][keep]
    inside = v @ np.asarray(N_HAT) >= np.cos(THETA_C)
    return (np.bincount(ids, weights=L * inside, minlength=n), np.bincount(ids, weights=L, minlength=n))


def test_energy_gate_above_the_beam_is_analog(oracle):
    rows_a, cnt_a, code_a, _ = _run(oracle, HBN_50KEV, 3, 300)
    rows_g, cnt_g, code_g, st = _run(oracle, HBN_50KEV, 3, 300, cone=_cone(N_HAT, target=0.05, min_energy_keV=60.0))
    assert st["cone"]["count"].sum() == 0 and not st["cone"]["killed"].any()
    np.testing.assert_array_equal(cnt_a, cnt_g)
    np.testing.assert_array_equal(code_a, code_g)
    for k in ("seg_mid", "seg_dir", "seg_len", "seg_E"):
        np.testing.assert_array_equal(rows_a[k], rows_g[k])


def test_primaries_follow_the_analog_streams_until_cone_entry(oracle):
    rows_a, cnt_a, code_a, _ = _run(oracle, HBN_50KEV, 3, 300)
    rows_s, cnt_s, code_s, st = _run(oracle, HBN_50KEV, 3, 300, cone=_cone(N_HAT, target=0.05))
    rec = st["cone"]
    assert rec["killed"].any() and rec["count"].sum() > 0
    cap = 4096
    for e in range(300):
        a = rows_a["seg_mid"][e * cap:e * cap + cnt_a[e]]
        b = rows_s["seg_mid"][e * cap:e * cap + cnt_s[e]]
        if rec["killed"][e]:
            assert code_s[e] == 6 and b.shape[0] <= a.shape[0]
            np.testing.assert_array_equal(a[:b.shape[0]], b)
        else:
            np.testing.assert_array_equal(a, b)
            assert code_s[e] == code_a[e]
    cp = oracle.cone_collect(rec)
    assert np.all(cp["dir"] @ np.asarray(N_HAT) >= np.cos(THETA_C) - 1e-12)
    assert np.all(cp["weight"] > 0.0)
    # ladder weights are powers of two
    m, _ = np.frexp(cp["weight"])
    assert np.all(m == 0.5)


def test_copy_records_are_prefix_stable_in_the_electron_count(oracle):
    cone = _cone(N_HAT, target=0.05)
    small = oracle.cone_collect(_run(oracle, HBN_50KEV, 5, 120, cone=cone)[3]["cone"])
    large = oracle.cone_collect(_run(oracle, HBN_50KEV, 5, 240, cone=cone)[3]["cone"])
    rows = large["parent"] < 120
    for key in ("parent", "flight", "pos", "dir", "E_keV", "weight", "stream_key"):
        np.testing.assert_array_equal(small[key], large[key][rows])


@pytest.mark.parametrize("options", [{}, {"min_energy_keV": 30.0, "angle_scale": 0.1}])
def test_split_tallies_match_analog_within_statistical_error(oracle, options):
    """Per-history in-cone and total path length: analog vs primary plus
    weighted copies (the copies transported as one analog generation), over
    3 seeds of 2000 histories; the in-cone tally's variance drops."""
    means = {"analog": [], "split": []}
    variances = {"analog": [], "split": []}
    ne = 2000
    for seed in range(3):
        rows, cnt, _, _ = _run(oracle, HBN_50KEV, seed, ne)
        cone_a, len_a = _tallies(rows, cnt, ne)
        rows, cnt, _, st = _run(oracle, HBN_50KEV, seed, ne, cone=_cone(N_HAT, target=0.05, **options))
        cone_b, len_b = _tallies(rows, cnt, ne)
        cp = oracle.cone_collect(st["cone"])
        nc = cp["weight"].size
        if nc:
            rows_c, cnt_c, _, _ = _run(oracle, HBN_50KEV, seed, nc, keys=cp["stream_key"],
                                       state=(cp["pos"], cp["dir"], cp["E_keV"], cp["clock"]))
            cone_c, len_c = _tallies(rows_c, cnt_c, nc)
            cone_b = history_sums(cone_b, cone_c, cp["parent"], cp["weight"], ne)
            len_b = history_sums(len_b, len_c, cp["parent"], cp["weight"], ne)
        for label, values in (("analog", (cone_a, len_a)), ("split", (cone_b, len_b))):
            means[label].append([v.mean() for v in values])
            variances[label].append([v.var(ddof=1) / ne for v in values])
    mean_a, mean_b = np.mean(means["analog"], axis=0), np.mean(means["split"], axis=0)
    se = np.sqrt((np.sum(variances["analog"], 0) + np.sum(variances["split"], 0)) / 9.0)
    assert np.all(np.abs(mean_a - mean_b) < 4.0 * se), (mean_a, mean_b, se)
    if not options:
        # the in-cone tally's variance per history drops (ratio about 0.7 here).  With
        # the energy gate and the angle roulette at u_ref = 0.1 (gamma ~ 1.1) the
        # roulette keeps about 1 copy in 8 on this screened-Rutherford law and the
        # ratio is about 1.1: unbiased, not more efficient, so only the mean is held.
        assert np.sum(variances["split"], 0)[0] < np.sum(variances["analog"], 0)[0]


# ---- the reference's own law: tabulated (ELSEPA) angles ---------------------------------------
def _tab_group(dcs_rows, energies_eV=(1000.0, 2000.0)):
    from pyrite_amd.tables.elsepa import pack_tables, read_dcs
    from tests.transport_cases import ELSEPA_PANEL

    panel = read_dcs(ELSEPA_PANEL)
    E = np.asarray(energies_eV, dtype=float)
    dcs = np.tile(panel.dcs_cm2_sr, (E.size, 1)) if dcs_rows is None else dcs_rows(panel.mu, E)
    return pack_tables([[{"energy_eV": E, "total_elastic_cm2": np.full(E.size, 1e-16), "mu": panel.mu,
                          "dcs_cm2_sr": dcs}]], [[1.0]]), panel


def _tab_scores(oracle, group, direction, E_keV, target, count, theta=THETA_C, angle_scale=0.0):
    keys = [oracle.cone_event_key(918273, i) for i in range(count)]
    made, out = oracle.cone_event_tab(np.tile(direction, (count, 1)), group, 0, E_keV, keys,
                                      _cone(theta=theta, target=target, angle_scale=angle_scale))
    o, w = out[:, :3], out[:, 3]
    assert np.all(w[made] > 0.0)
    assert np.all(np.abs(np.linalg.norm(o[made], axis=1) - 1.0) < 1e-12)
    assert np.all(o[made, 2] >= np.cos(theta) - 1e-12)
    w = np.where(made, w, 0.0)
    return np.stack([w, w * o[:, 2]], axis=1)


@pytest.mark.parametrize("alpha", [0.02, 0.5])
@pytest.mark.parametrize("angle", [0.1, 1.0, 2.5])
@pytest.mark.parametrize("target", [1e-4, 0.2])
def test_tabulated_law_reproduces_the_screened_rutherford_cap(oracle, alpha, angle, target):
    """A table holding the screened-Rutherford density on the vendor grid: the
    tabulated event (band bound from the rows' CDFs, band by bisection on the
    mixed quantile) gives the law's cap probability and first moment."""
    group, _ = _tab_group(lambda mu, E: np.tile(alpha * (1 + alpha) / (mu + alpha) ** 2, (E.size, 1)))
    d = np.array([np.sin(angle), 0.0, np.cos(angle)])
    scores = _tab_scores(oracle, group, d, 1.5, target, 12000)
    p, q = _sr_cap_integrals(alpha, float(np.cos(angle)), THETA_C)
    se = scores.std(axis=0, ddof=1) / np.sqrt(scores.shape[0])
    assert abs(scores[:, 0].mean() - p) <= 6.0 * se[0] + 2e-4 * p, (scores[:, 0].mean(), p, se[0])
    assert abs(scores[:, 1].mean() - q) <= 6.0 * se[1] + 2e-4 * abs(q), (scores[:, 1].mean(), q, se[1])


@pytest.mark.parametrize("angle", [0.3, 1.2])
def test_tabulated_event_weights_are_the_vendor_laws_cap_probability(oracle, angle):
    """The vendor DCS between its two nodes (a quantile mixture in ln E): the
    mean copy weight equals the cap probability computed by stratified
    quadrature over the analog sampler's uniform, sum over xi of the arc
    fraction half(theta(xi))/pi -- no band or bisection involved."""
    group, _ = _tab_group(None)
    E = 1.5
    d = np.array([np.sin(angle), 0.0, np.cos(angle)])
    xi = (np.arange(400_000) + 0.5) / 400_000
    mu = oracle.elsepa_sample_mu(group, 0, E, xi)
    cos_t = 1.0 - 2.0 * mu
    sin_t = np.sqrt(np.maximum(0.0, 1.0 - cos_t**2))
    cos_dn, sin_dn, cos_c = np.cos(angle), np.sin(angle), np.cos(THETA_C)
    with np.errstate(divide="ignore", invalid="ignore"):
        c_star = (cos_c - cos_t * cos_dn) / (sin_t * sin_dn)
    half = np.where(c_star <= -1.0, np.pi, np.where(c_star >= 1.0, 0.0, np.arccos(np.clip(c_star, -1.0, 1.0))))
    p = (half / np.pi).mean()
    assert p > 1e-4
    scores = _tab_scores(oracle, group, d, E, 1e-3, 12000)
    se = scores[:, 0].std(ddof=1) / np.sqrt(scores.shape[0])
    assert abs(scores[:, 0].mean() - p) <= 6.0 * se + 1e-3 * p, (scores[:, 0].mean(), p, se)


def test_split_tallies_match_analog_on_tabulated_laws(oracle):
    """The split-vs-analog tally check on ELSEPA-format tables (N on the
    vendor DCS shape): unbiased, in-cone variance lower."""
    means = {"analog": [], "split": []}
    variances = {"analog": [], "split": []}
    ne = 2000
    case = HBN_50KEV_ELSEPA
    for seed in range(3):
        rows, cnt, _, _ = _run(oracle, case, seed, ne)
        cone_a, len_a = _tallies(rows, cnt, ne)
        rows, cnt, _, st = _run(oracle, case, seed, ne, cone=_cone(N_HAT, target=0.05))
        cone_b, len_b = _tallies(rows, cnt, ne)
        cp = oracle.cone_collect(st["cone"])
        nc = cp["weight"].size
        assert nc > 0
        rows_c, cnt_c, _, _ = _run(oracle, case, seed, nc, keys=cp["stream_key"],
                                   state=(cp["pos"], cp["dir"], cp["E_keV"], cp["clock"]))
        cone_c, len_c = _tallies(rows_c, cnt_c, nc)
        cone_b = history_sums(cone_b, cone_c, cp["parent"], cp["weight"], ne)
        len_b = history_sums(len_b, len_c, cp["parent"], cp["weight"], ne)
        for label, values in (("analog", (cone_a, len_a)), ("split", (cone_b, len_b))):
            means[label].append([v.mean() for v in values])
            variances[label].append([v.var(ddof=1) / ne for v in values])
    mean_a, mean_b = np.mean(means["analog"], axis=0), np.mean(means["split"], axis=0)
    se = np.sqrt((np.sum(variances["analog"], 0) + np.sum(variances["split"], 0)) / 9.0)
    assert np.all(np.abs(mean_a - mean_b) < 4.0 * se), (mean_a, mean_b, se)
    assert np.sum(variances["split"], 0)[0] < np.sum(variances["analog"], 0)[0]
