"""Behaviour-level pins of the oracle's heuristic rules, where the compiled reference cannot
reach (model/samodel.c needs netcdf.h and cpgplot.h).  Each rule is treated as a black box
and checked by a property a plausible transcription mistake (a wrong tier boundary or
threshold, a swapped constant, the wrong band or date, a dropped weight) would break:

* the K penalty (model/samodel.c:2696-2732): its jumps located by bisection in H and in K_min;
  continuity at the thresholds where the formula's reference K equals the threshold, the one
  jump where it does not; the sum over dates; the last date's K_min alone in the K > 0.7 term;
* the start/step rule (:2246-2354): invariances under scaling the reflectances, moving one band,
  moving the depth start;
* the outputs (:1992-2079): IOD's limits at H = 0 and H -> infinity, K_min's independence of
  H and the bottom, the bottom percentages' invariance, the origin weight sqrt(N_r) of the
  region depth."""
import numpy as np
import pytest

import _oracle as O

WL = O.WL4
EG = np.exp(-0.015 * (np.array(WL) - 440.0))


def _meas():
    return np.array([0.010, 0.011, 0.008, 0.002], np.float32)


def _kerr(H, G=0.0, P=1e-300, X=0.0, dates=1):
    """K penalty of the mix-layout vector (H, B, P, G, X) over `dates` identical dates."""
    x = [H, 0.5] + [100 * P, 100 * G, 100 * X] * dates
    return O.evaluate(x, np.tile(_meas(), (dates, 1)))["k_error"]


def _kmin(G=0.0, P=1e-300, X=0.0):
    x = [5.0, 0.5, 100 * P, 100 * G, 100 * X]
    return O.evaluate(x, _meas())["K"].min()


def _bisect(f, lo, hi, it=200):
    """The boundary between f's value at lo and at hi (f piecewise constant, one jump)."""
    flo = f(lo)
    assert f(hi) != flo
    for _ in range(it):
        mid = 0.5 * (lo + hi)
        if mid == lo or mid == hi:
            break
        if f(mid) == flo:
            lo = mid
        else:
            hi = mid
    return lo, hi


# ---- the K penalty ----------------------------------------------------------------------

@pytest.mark.parametrize("t", [1.0, 2.0, 3.0, 4.0, 5.0])
def test_k_penalty_tiers_change_at_whole_metres(t):
    """With clear water (K_min far below every threshold) the penalty is piecewise constant in
    H; each jump sits exactly at a whole metre, and the test is strict: H = t belongs to the
    tier above."""
    assert _kmin() < 0.1
    lo, hi = _bisect(lambda h: _kerr(h), t - 0.5, t + 0.5)
    assert hi == t and lo == np.nextafter(t, 0.0)
    assert _kerr(t) == _kerr(t + 0.3) and _kerr(lo) == _kerr(t - 0.3)


def test_k_penalty_constant_within_tiers_and_off_beyond_5m():
    for a, b in ((0.01, 0.99), (1.0, 1.99), (2.0, 2.99), (3.0, 3.99), (4.0, 4.99)):
        vals = {_kerr(h) for h in np.linspace(a, b, 17)}
        assert len(vals) == 1 and vals.pop() > 0.0
    assert all(_kerr(h) == 0.0 for h in (5.0, 5.5, 8.0, 25.0))
    # tiers 2 and 4 share their threshold; the others differ from one another
    v = [_kerr(h) for h in (0.5, 1.5, 2.5, 3.5, 4.5)]
    assert v[1] == v[3] and len({v[0], v[1], v[2], v[4]}) == 4


# The tiers' K_min thresholds, formed as the reference forms them (in double): 0.275 below
# 1 m, then 0.5 (w 0.185 + (2 - w) 0.275) with w = 1.5, 1.25, 1.5, 1.75 below 2, 3, 4, 5 m.
THR = [0.275, 0.5 * (1.5 * 0.185 + 0.5 * 0.275), 0.5 * (1.25 * 0.185 + 0.75 * 0.275),
       0.5 * (1.5 * 0.185 + 0.5 * 0.275), 0.5 * (1.75 * 0.185 + 0.25 * 0.275)]


@pytest.mark.parametrize("H", [0.5, 1.5, 2.5, 3.5, 4.5])
def test_k_penalty_thresholds_in_k_min(H):
    """Raising gelbstoff raises K_min through the thresholds.  The tiers are one if / else-if
    chain, so at depth H the penalty is on while K_min is below ANY threshold of a tier deeper
    than H, and switches off at the largest of them (1.5 m: the 3 m tier's 0.21875, not the 2 m
    tier's 0.2075).  Below 1 m the formula's reference K (0.185) is not the threshold, so the
    penalty jumps from 100 (1/0.285 - 1/0.195)^2 to zero; deeper it fades to zero
    continuously."""
    thr = max(t for depth, t in zip((1, 2, 3, 4, 5), THR) if H < depth)
    assert 0.3 < _kmin(G=2.0) < 0.7      # past every threshold, short of the turbid term
    lo, hi = _bisect(lambda g: _kerr(H, G=g) > 0.0, 0.0, 2.0)
    k_lo, k_hi = _kmin(G=lo), _kmin(G=hi)
    assert k_lo < thr <= k_hi and k_hi - k_lo < 1e-12
    just_below = _kerr(H, G=lo)
    if H < 1.0:
        assert np.isclose(just_below, 100.0 * (1 / 0.285 - 1 / 0.195) ** 2, rtol=1e-9)
    else:
        assert just_below < 1e-9


def test_k_penalty_falls_through_to_the_next_tier():
    """At 1.5 m a K_min between the 2 m tier's threshold (0.2075) and the 3 m tier's
    (0.21875) fails the first test and takes the 3 m tier's penalty: exactly the penalty the
    same water has at 2.5 m.  Below 0.2075 the two depths differ."""
    g_mid = _bisect(lambda g: _kmin(G=g) < 0.5 * (THR[1] + THR[2]), 0.0, 2.0)[0]
    assert THR[1] < _kmin(G=g_mid) < THR[2]
    assert _kerr(1.5, G=g_mid) == _kerr(2.5, G=g_mid) > 0.0
    g_low = _bisect(lambda g: _kmin(G=g) < 0.15, 0.0, 2.0)[0]
    assert _kerr(1.5, G=g_low) != _kerr(2.5, G=g_low)


def test_k_penalty_sums_over_dates():
    for H in (0.5, 2.5, 4.5):
        one = _kerr(H)
        assert _kerr(H, dates=2) == one + one
        assert np.isclose(_kerr(H, dates=3), 3 * one, rtol=1e-15)


def test_turbid_term_reads_the_last_date():
    """The K_min > 0.7 term reads K_min of the last date the tier loop visited: a turbid date
    costs it only when it comes last."""
    turbid = [100 * 1e-300, 100 * 12.0, 0.0]       # P, G, X: K_min > 0.7
    clear = [100 * 1e-300, 0.0, 0.0]
    assert _kmin(G=12.0) > 0.7
    meas = np.tile(_meas(), (2, 1))
    e_tc = O.evaluate([8.0, 0.5] + turbid + clear, meas)["k_error"]
    e_ct = O.evaluate([8.0, 0.5] + clear + turbid, meas)["k_error"]
    e_t = O.evaluate([8.0, 0.5] + turbid, _meas())["k_error"]
    assert e_tc == 0.0
    assert e_ct > 0.0 and e_ct == e_t
    k = _kmin(G=12.0)
    assert np.isclose(e_t, 100.0 * (4.0 * (k - 0.7)) ** 2, rtol=1e-12)


def test_objective_weights_the_penalty_15_of_120():
    """objective = (80 E_Rrs + 15 K_error) / 120 with one region: holding E_Rrs fixed, the
    objective moves by 15/120 of the penalty (here across the 1 m boundary)."""
    a = O.evaluate([1.0, 0.5, 1e-298, 0.0, 0.0], _meas())
    b = O.evaluate([np.nextafter(1.0, 0.0), 0.5, 1e-298, 0.0, 0.0], _meas())
    assert np.isclose(b["rrs_error"], a["rrs_error"], rtol=1e-12)
    d_obj = b["objective"] - a["objective"]
    d_pen = b["k_error"] - a["k_error"]
    assert np.isclose(d_obj, 15.0 / 120.0 * d_pen, rtol=1e-9)
    assert np.isclose(a["objective"], (80.0 * a["rrs_error"] + 15.0 * a["k_error"]) / 120.0,
                      rtol=1e-14)


# ---- the start/step rule ------------------------------------------------------------------

R0 = np.array([0.0110, 0.0123, 0.0087, 0.0031])


@pytest.mark.parametrize("mode,nbot", [(1, 2), (0, 1), (0, 2), (0, 3)])
def test_start_scales_with_reflectance(mode, nbot):
    """Doubling every reflectance (exact in binary) leaves H, q and the P and G starts
    (functions of the 490/550 ratio) unchanged and doubles the X start (Rrs at 640 nm) and the
    bottom amplitudes (the mean Rrs at 490 nm), steps alike."""
    s1, t1 = O.start_step(R0, 10.0, mode=mode, n_bottoms=nbot)
    s2, t2 = O.start_step(2 * R0, 10.0, mode=mode, n_bottoms=nbot)
    off = 2 if mode == 1 else 1 + 2 * nbot
    same = [0, off, off + 1] + ([1] if mode == 1 else list(range(1 + nbot, 1 + 2 * nbot)))
    twice = [off + 2] + ([] if mode == 1 else list(range(1, 1 + nbot)))
    for i in same:
        assert s1[i] == s2[i] and t1[i] == t2[i], i
    for i in twice:
        assert s2[i] == 2 * s1[i] and t2[i] == 2 * t1[i], i
    assert sorted(same + twice) == list(range(len(s1)))


@pytest.mark.parametrize("mode,nbot", [(1, 2), (0, 2)])
def test_start_reads_bands_through_490_550_640(mode, nbot):
    """With bands at 443/482/561/655 nm the rule reads Rrs at 490 and 550 nm (between 482 and
    561) and at 640 nm (between 561 and 655): moving the 443 nm band changes nothing (the rule
    uses 490, not the paper's 440, for P), moving the 655 nm band changes only the X start."""
    base, tb = O.start_step(R0, 10.0, mode=mode, n_bottoms=nbot)
    r = R0.copy()
    r[0] *= 3.0
    s, t = O.start_step(r, 10.0, mode=mode, n_bottoms=nbot)
    assert np.array_equal(s, base) and np.array_equal(t, tb)
    r = R0.copy()
    r[3] *= 1.5
    s, t = O.start_step(r, 10.0, mode=mode, n_bottoms=nbot)
    off = 2 if mode == 1 else 1 + 2 * nbot
    changed = np.nonzero(s != base)[0].tolist()
    assert changed == [off + 2] and np.nonzero(t != tb)[0].tolist() == [off + 2]
    r = R0.copy()
    r[1] *= 1.2                          # the 482 nm band moves P, G and the amplitudes
    s, _ = O.start_step(r, 10.0, mode=mode, n_bottoms=nbot)
    moved = set(np.nonzero(s != base)[0].tolist())
    assert {off, off + 1} <= moved and (mode == 1 or set(range(1, 1 + nbot)) <= moved)


def test_start_depth_is_the_depth_start():
    s1, t1 = O.start_step(R0, 7.5, mode=0, n_bottoms=2)
    s2, t2 = O.start_step(R0, 15.0, mode=0, n_bottoms=2)
    assert s1[0] == 7.5 and s2[0] == 15.0 and t2[0] == 2 * t1[0] and t1[0] > s1[0]
    assert np.array_equal(s1[1:], s2[1:]) and np.array_equal(t1[1:], t2[1:])


# ---- the outputs ----------------------------------------------------------------------------

def _x(H, nbot=3, q=(1.0, 2.0, 3.0), P=0.02, G=0.01, X=0.005):
    return [H] + [30.0] * nbot + list(q[:nbot]) + [100 * P, 100 * G, 100 * X]


def test_iod_limits_and_monotone():
    """IOD = 100 mean(rrs_bottom / rrs): every band's signal is the bottom's at H = 0 (IOD =
    100 exactly) and none of it as H grows without bound; in between it falls with H."""
    e0 = O.evaluate(_x(0.0), _meas(), mode=0, n_bottoms=3)
    assert e0["iod"] == 100.0
    deep = O.evaluate(_x(1.0e4), _meas(), mode=0, n_bottoms=3)
    assert 0.0 <= deep["iod"] < 1e-12
    iods = [O.evaluate(_x(h), _meas(), mode=0, n_bottoms=3)["iod"]
            for h in (0.1, 0.5, 1, 2, 5, 10, 20, 40)]
    assert all(a > b for a, b in zip(iods, iods[1:]))


def test_k_min_is_the_water_column_minimum():
    """K_min is the smallest band attenuation: independent of depth and bottom, growing with
    each constituent, and with clear water the smallest a_w + b_bw of the band tables."""
    a = O.evaluate(_x(3.0), _meas(), mode=0, n_bottoms=3)
    b = O.evaluate(_x(17.0, q=(5.0, 1.0, 0.1)), _meas(), mode=0, n_bottoms=3)
    assert a["k_min"] == b["k_min"] == a["K"].min()
    for kw in (dict(P=0.05), dict(G=0.05), dict(X=0.02)):
        assert O.evaluate(_x(3.0, **kw), _meas(), mode=0, n_bottoms=3)["k_min"] > a["k_min"]
    t, _ = O.band_tables(WL)            # rows a0, a1, a_w, b_bw, then the bottoms
    clear = O.evaluate(_x(3.0, P=1e-300, G=0.0, X=0.0), _meas(), mode=0, n_bottoms=3)
    assert np.isclose(clear["k_min"], np.min(t[2] + t[3]), rtol=1e-12)


def test_bottom_percentages_and_type():
    """Bottom percentages depend on the weights q only (scale-free), sum to 100, and the type
    is 1 + the table index of the largest; the amplitudes do not enter."""
    e = O.evaluate(_x(3.0, q=(1.0, 2.0, 3.0)), _meas(), mode=0, n_bottoms=3)
    f = O.evaluate([3.0, 10.0, 70.0, 5.0, 0.5, 1.0, 1.5] + [2.0, 1.0, 0.5], _meas(), mode=0,
                   n_bottoms=3)
    assert np.allclose(e["b_pct"][:3], f["b_pct"][:3], rtol=1e-14)
    assert np.isclose(e["b_pct"].sum(), 100.0, rtol=1e-14)
    assert e["bottom_type"] == 3.0
    g = O.evaluate(_x(3.0, nbot=2, q=(4.0, 1.0)), _meas(), mode=0, n_bottoms=2,
                   bottom_idx=(2, 0, 1))
    assert g["bottom_type"] == 3.0 and g["b_pct"][2] > g["b_pct"][0] and g["b_pct"][1] == 0.0


@pytest.mark.parametrize("R,origin", [(9, 4), (9, 0), (4, 2), (25, 12)])
def test_region_depth_weights_the_origin_by_sqrt_regions(R, origin):
    """The region depth is a weighted mean of the regions' depths: equal depths give that
    depth, and moving the origin's depth moves the result sqrt(N_r) times as much as moving
    any other region's."""
    nbot = 1
    rng = np.random.default_rng(R + origin)
    rrs = rng.uniform(0.003, 0.012, (R, 1, 4))
    base = [8.0] * R + [30.0] * R + [1.0] * R + [2.0, 1.0, 0.5]
    _, d0, _, _ = O.region_evaluate(base, rrs, origin=origin, n_bottoms=nbot)
    assert np.isclose(d0, 8.0, rtol=1e-14)
    delta = 0.5
    sens = []
    for r in range(R):
        x = list(base)
        x[r] += delta
        sens.append(O.region_evaluate(x, rrs, origin=origin, n_bottoms=nbot)[1] - d0)
    sens = np.array(sens)
    others = np.delete(sens, origin)
    assert np.allclose(others, others[0], rtol=1e-9)
    assert np.isclose(sens[origin] / others[0], np.sqrt(R), rtol=1e-9)
    assert np.isclose(sens.sum(), delta, rtol=1e-9)


# ---- the ordered inversion's spectral rules (model/samodel.c:701-770, 2795-2842) -------------

def test_spectral_diff_mean_properties():
    """model_spectral_diff_mean over bands 1..3: zero for identical spectra, scale-free, blind
    to band 0, and 100 |d| for a uniform relative offset d."""
    m = np.array([0.011, 0.013, 0.009, 0.004])
    assert O.spectral_diff_mean(m, m) == 0.0
    l = m * np.array([0.7, 1.02, 0.97, 1.05])
    d = O.spectral_diff_mean(m, l)
    assert d > 0.0
    assert O.spectral_diff_mean(2 * m, 2 * l) == d
    l0 = l.copy()
    l0[0] = 5.0
    assert O.spectral_diff_mean(m, l0) == d
    for off in (0.01, -0.03, 0.2):
        assert np.isclose(O.spectral_diff_mean(m, m * (1 + off)), 100 * abs(off), rtol=1e-6)
    # sqrt of the mean over three bands: one band off by d gives 100 d / sqrt(3)
    l1 = m.copy()
    l1[2] *= 1.06
    assert np.isclose(O.spectral_diff_mean(m, l1), 100 * 0.06 / np.sqrt(3), rtol=1e-6)


def test_order_key_is_the_norm_times_the_angle():
    """The spectral index: the norm of the Rrs at bands 1, 2, 3 (band 0 ignored); with a noise
    spectrum, times the angle between the two, so a pixel whose spectrum is parallel to the
    noise keys to zero and an orthogonal one to norm pi / 2."""
    r = np.array([0.5, 0.003, 0.004, 0.012], np.float32)
    k = O.order_key(r)
    assert np.isclose(k, np.linalg.norm(r[1:].astype(np.float64)), rtol=1e-7)
    r2 = r.copy()
    r2[0] = 9.0
    assert O.order_key(r2) == k
    noise = np.array([1e-4, 3e-4, 4e-4, 12e-4])
    assert O.order_key(r, noise) < 1e-3 * k          # parallel: angle ~ 0
    ko = O.order_key(np.array([0.0, 0.004, 0.003, 0.0], np.float32), noise)
    v = np.array([0.004, 0.003, 0.0])
    ang = np.arccos(v @ noise[1:] / (np.linalg.norm(v) * np.linalg.norm(noise[1:])))
    assert np.isclose(ko, np.float32(np.linalg.norm(v)) * ang, rtol=1e-6)
