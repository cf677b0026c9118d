"""The oracle's forward model, objective, start rules and multi-start, pinned where the
reference's compiled code cannot reach them (model/samodel.c needs netCDF and PGPLOT headers
this image lacks): against the reference's own tables (read from its source text into
tests/golden/reference_tables.json), against special cases in which the model reduces to a
closed form, and against a brute-force restatement of the multi-start.  The product's band
coefficients are checked against the same tables."""
import json
import os

import numpy as np
import pytest

import _oracle as O
import photic_amd as P

HERE = os.path.dirname(os.path.abspath(__file__))
TAB = {k: np.array(v["values"]) for k, v in
       json.load(open(os.path.join(HERE, "golden", "reference_tables.json"))).items()}
NODES = (440.0, 490.0, 550.0, 640.0)  # table nodes: interpolation returns the node value
SAND, SEAGRASS = TAB["bottom_0_sand"], TAB["bottom_1_seagrass"]


def tab(name, wl):
    return np.interp(np.asarray(wl, float), TAB["ref_wl"], TAB[name])


def to_r(Rrs):
    """Below-surface r from above-surface Rrs = 0.5 r / (1 - 1.5 r)."""
    return Rrs / (0.5 + 1.5 * Rrs)


def fwd(H, Pc, G, X, B, wl=NODES, ss=1.0, Y=1.0):
    p = np.array([[H, Pc, G, X, B]], float)
    return O.forward(p, np.array([ss]), np.array([Y]), wl=wl)[0]


# --- tables ----------------------------------------------------------------------------

def test_oracle_tables_equal_reference_text():
    for mine, ref in (("ref_wl", "ref_wl"), ("aw", "aw"), ("bbw", "bbw"), ("a0", "a0"),
                      ("a1", "a1")):
        assert np.array_equal(O.table(mine), TAB[ref]), mine
    bottoms = O.table("bottom")
    keys = sorted(k for k in TAB if k.startswith("bottom_"))
    for k, key in enumerate(keys):
        assert np.array_equal(bottoms[k], TAB[key]), key


def test_product_band_coefficients_equal_reference_text():
    wl = TAB["ref_wl"]
    names = ["a0", "a1", "aw", "bbw"] + sorted(k for k in TAB if k.startswith("bottom_"))
    for g in range(0, 41, 8):
        w = tuple(wl[g:g + 8])
        if len(w) < 2:
            w = tuple(wl[g - 1:g + 1])
        coef, aw640 = P.band_coefficients(P.cfg_default(wavelengths=w))
        idx = [int(np.where(wl == x)[0][0]) for x in w]
        for r, name in enumerate(names):
            assert np.array_equal(coef[r], TAB[name][idx]), (name, w)
        assert aw640 == TAB["aw"][24]  # 640 nm is a node
    # between nodes: the product's interpolation equals the oracle's bit for bit
    w = (443.0, 482.0, 561.0, 655.0)
    coef, _ = P.band_coefficients(P.cfg_default(wavelengths=w))
    ocoef, _ = O.band_tables(w)
    assert np.array_equal(coef, ocoef)
    assert np.allclose(coef[2], tab("aw", w), rtol=1e-14)


# --- forward model: closed forms ---------------------------------------------------------

def test_forward_zero_depth_is_the_bottom():
    """H = 0: the water column vanishes, r = rho / pi (model/samodel.c:2919-2934)."""
    for B in (0.0, 0.3, 1.0):
        Rrs = fwd(0.0, 0.05, 0.02, 0.01, B, Y=1.3)
        rho = B * tab("bottom_0_sand", NODES) + (1 - B) * tab("bottom_1_seagrass", NODES)
        assert np.allclose(to_r(Rrs), rho / np.pi, rtol=1e-14, atol=0)


def _deep_r(a, bb):
    u = bb / (a + bb)
    return (0.084 + 0.170 * u) * u


def test_forward_deep_pure_water():
    """P, G, X -> 0 and H -> infinity: r = r_dp of pure water, u = b_bw / (a_w + b_bw)."""
    Rrs = fwd(1.0e4, 1e-300, 0.0, 0.0, 0.5)
    r = _deep_r(tab("aw", NODES), tab("bbw", NODES))
    assert np.allclose(to_r(Rrs), r, rtol=1e-13, atol=0)


def test_forward_deep_single_constituents():
    aw, bbw = tab("aw", NODES), tab("bbw", NODES)
    # gelbstoff: a_g = G exp(-0.015 (lambda - 440))
    Rrs = fwd(1.0e4, 1e-300, 0.2, 0.0, 0.5)
    ag = 0.2 * np.exp(-0.015 * (np.array(NODES) - 440.0))
    assert np.allclose(to_r(Rrs), _deep_r(aw + ag, bbw), rtol=1e-13)
    # particles: b_bp = X (440 / lambda)^Y
    for Y in (0.0, 1.7):
        Rrs = fwd(1.0e4, 1e-300, 0.0, 0.03, 0.5, Y=Y)
        bbp = 0.03 * (440.0 / np.array(NODES)) ** Y
        assert np.allclose(to_r(Rrs), _deep_r(aw, bbw + bbp), rtol=1e-13)
    # phytoplankton: a_phi = (a0 + a1 ln P) P; at 440 nm a0 = 1, a1 = 0
    a0, a1 = tab("a0", NODES), tab("a1", NODES)
    for Pc in (1.0, np.e, 0.04):
        Rrs = fwd(1.0e4, Pc, 0.0, 0.0, 0.5)
        aphi = (a0 + a1 * np.log(Pc)) * Pc
        assert np.allclose(to_r(Rrs), _deep_r(aw + aphi, bbw), rtol=1e-13), Pc
    assert a0[0] == 1.0 and a1[0] == 0.0


def test_forward_is_affine_in_the_bottom_mix():
    """r is affine in rho, and rho is affine in B."""
    r = [to_r(fwd(3.0, 0.05, 0.02, 0.01, B)) for B in (0.0, 0.25, 1.0)]
    assert np.allclose(r[1], 0.75 * r[0] + 0.25 * r[2], rtol=1e-12)


@pytest.mark.parametrize("G,ss", [(0.0, 1.0), (0.0, 1.4), (10.0, 1.2)])
def test_forward_attenuation_coefficients(G, ss):
    """dr/dH at H = 0 is K (r_dp M_C - (rho/pi) M_B) with M_C = sec_sun + 1.03 sqrt(1 + 2.4 u)
    and M_B = sec_sun + 1.04 sqrt(1 + 5.4 u) (nadir view); with G = 10, a + b_b exceeds 2.5
    at every band below 640 nm and K is clamped to 2.5."""
    aw, bbw = tab("aw", NODES), tab("bbw", NODES)
    a = aw + G * np.exp(-0.015 * (np.array(NODES) - 440.0))
    u = bbw / (a + bbw)
    K = np.minimum(a + bbw, 2.5)
    rdp = _deep_r(a, bbw)
    h = 1e-6
    d = {}
    for B in (0.0, 1.0):
        rp = to_r(fwd(h, 1e-300, G, 0.0, B, ss=ss))
        rm = to_r(fwd(-h, 1e-300, G, 0.0, B, ss=ss))
        d[B] = (rp - rm) / (2 * h)
    rho0, rho1 = tab("bottom_1_seagrass", NODES), tab("bottom_0_sand", NODES)
    sel = np.abs(rho1 - rho0) > 1e-3  # both spectra are normalised to 1 at 550 nm
    K, rdp, u, rho0, rho1 = K[sel], rdp[sel], u[sel], rho0[sel], rho1[sel]
    MB = (d[0.0][sel] - d[1.0][sel]) / (K * (rho1 - rho0) / np.pi)
    MC = (d[0.0][sel] + K * rho0 / np.pi * MB) / (K * rdp)
    assert sel.sum() == 3
    assert np.allclose(MC, ss + 1.03 * np.sqrt(1 + 2.4 * u), rtol=2e-6)
    assert np.allclose(MB, ss + 1.04 * np.sqrt(1 + 5.4 * u), rtol=2e-6)
    if G > 0:
        assert np.all(K[:2] == 2.5)


def test_Y_of_chi_clamps_and_value():
    """Y = 3.44 (1 - 3.17 exp(-2.01 chi)) clamped to [0, 2.5] (model/samodel.c:2898-2901)."""
    assert O.Y_of_chi(0.0) == 0.0
    assert O.Y_of_chi(10.0) == 2.5
    assert abs(O.Y_of_chi(1.0) - 1.9788803) < 1e-7


# --- objective -----------------------------------------------------------------------------

def _objective(x, rrs_meas, ss=1.0):
    return O.objective(np.array([x], float), np.array([rrs_meas], np.float32),
                       np.array([ss], np.float32), wl=NODES)


def test_objective_e_rrs_and_weights():
    """E_Rrs = 100 sqrt(mean residual^2) / mean measured; objective = 80 E_Rrs / 120 when no
    K penalty applies (H >= 5, K_min <= 0.7) (model/samodel.c:2545-2579, 2736-2744)."""
    H, Pc, G, X, B = 8.0, 0.05, 0.03, 0.01, 0.6
    m0 = fwd(H, Pc, G, X, B)
    meas = (m0 * np.array([1.01, 0.98, 1.03, 1.0])).astype(np.float32)
    Y = O.Y_of_chi(float(meas[0]) / float(meas[1]))
    m = fwd(H, Pc, G, X, B, Y=Y)
    e = 100.0 * np.sqrt(np.mean((m - meas.astype(float)) ** 2)) / np.mean(meas.astype(float))
    obj, err = _objective([H, B, 100 * Pc, 100 * G, 100 * X], meas)
    assert np.isclose(err[0], e, rtol=1e-12)
    assert np.isclose(obj[0], 80.0 * e / 120.0, rtol=1e-12)


# K-penalty tiers of model/samodel.c:2696-2732 (k_max_min 0.275, k_min_min 0.185): below
# depth H_t a K_min under the tier's threshold costs 100 (1/(0.01+K_min) - 1/(0.01+thr))^2.
TIERS = [(0.5, 0.185), (1.5, 0.5 * (1.5 * 0.185 + 0.5 * 0.275)),
         (2.5, 0.5 * (1.25 * 0.185 + 0.75 * 0.275)), (3.5, 0.5 * (1.5 * 0.185 + 0.5 * 0.275)),
         (4.5, 0.5 * (1.75 * 0.185 + 0.25 * 0.275)), (6.0, None)]


@pytest.mark.parametrize("H,thr", TIERS)
def test_objective_k_penalty_tiers(H, thr):
    meas = fwd(H, 1e-300, 0.0, 0.0, 0.5).astype(np.float32)
    obj, err = _objective([H, 0.5, 1e-298, 0.0, 0.0], meas)
    kmin = np.min(tab("aw", NODES) + tab("bbw", NODES))
    kerr = 0.0 if thr is None else 100.0 * (1 / (0.01 + kmin) - 1 / (0.01 + thr)) ** 2
    assert np.isclose(obj[0], (80.0 * err[0] + 15.0 * kerr) / 120.0, rtol=1e-12)
    assert kerr == 0.0 or kerr > 1e4


def test_objective_turbid_penalty():
    """K_min > 0.7 costs 100 (4 (K_min - 0.7))^2."""
    G = 10.0
    meas = fwd(8.0, 1e-300, G, 0.0, 0.5).astype(np.float32)
    obj, err = _objective([8.0, 0.5, 1e-298, 100 * G, 0.0], meas)
    a = tab("aw", NODES) + G * np.exp(-0.015 * (np.array(NODES) - 440.0))
    kmin = np.min(np.minimum(a + tab("bbw", NODES), 2.5))
    assert kmin > 0.7
    kerr = 100.0 * (4.0 * (kmin - 0.7)) ** 2
    assert np.isclose(obj[0], (80.0 * err[0] + 15.0 * kerr) / 120.0, rtol=1e-12)


# --- start and step (model/samodel.c:2246-2354) ---------------------------------------------

def test_start_step_hand_values():
    R = np.array([0.010, 0.012, 0.009, 0.003])  # at 440, 490, 550, 640 nm
    aw640 = 0.3108                              # model/samodel.c:130, 640 nm
    P0 = 100 * 0.072 * (0.012 / 0.009) ** -1.7  # P from Rrs(490)/Rrs(550)
    X0 = 100 * 30 * aw640 * 0.003
    start, step = O.start_step(R, 40.0, mode=1, wl=NODES)
    assert np.allclose(start, [40.0, 0.5, P0, 1.5 * P0, X0], rtol=1e-15)
    assert np.allclose(step, [50.0, 0.5, 2 * P0, 3 * P0, 2 * X0], rtol=1e-15)
    start, step = O.start_step(R, 7.5, mode=0, n_bottoms=2, wl=NODES)
    B0 = 100 * 4 * 0.012
    assert np.allclose(start, [7.5, B0, B0, 1.0, 1.0, P0, 1.5 * P0, X0], rtol=1e-15)
    assert np.allclose(step, [9.375, 1.5 * B0, 1.5 * B0, 0.5, 0.5, 2 * P0, 3 * P0, 2 * X0],
                       rtol=1e-15)


# --- multi-start (model/samodel.c:2119-2427) -------------------------------------------------

def test_multistart_is_the_first_good_start_or_the_best():
    """The depth starts run in order; the run stops after the first start whose best misfit is
    below 2.5, and keeps the lowest misfit seen (the first on ties).  Restated by brute force
    from independent single-start runs."""
    starts = (40.0, 30.0, 20.0, 15.0, 10.0, 7.5, 2.5, 1.0)
    rrs, ss, truth, nodata, _ = O.gen(21, 48, deep_frac=0.3)
    kw = dict(reqmin=1e-10, konvge=100, kcount=2000)
    full = O.invert(rrs, ss, **kw)
    single = [O.invert(rrs, ss, hstart=(h,), **kw) for h in starts]
    n_checked = 0
    for p in np.where(full["status"] >= 0)[0]:
        errs = [s["misfit"][p] for s in single]
        best, n_used = 0, len(starts)
        for k, e in enumerate(errs):
            if e < errs[best]:
                best = k
            if errs[best] < 2.5:
                n_used = k + 1
                break
        assert full["n_starts"][p] == n_used, p
        assert np.array_equal(full["x"][:, p], single[best]["x"][:, p]), p
        n_checked += 1
    assert n_checked > 40
    assert len(set(full["n_starts"][full["status"] >= 0])) > 1


# --- outputs (model/samodel.c:1992-2079) ----------------------------------------------------

def test_outputs_from_the_best_vector():
    """Depth = H; bottom % = 100 q_k / sum q; bottom type = 1 + the largest; albedo =
    sum (q_k / sum q) B_k / 100, from the native vector [H, B_0..2, q_0..2, P, G, X]."""
    rrs, ss, _, _, _ = O.gen(22, 24)
    o = O.invert(rrs, ss, mode=0, n_bottoms=3, reqmin=1e-10, konvge=100, kcount=2000)
    v = o["status"] >= 0
    x = o["x"][:, v]
    q = np.abs(x[4:7])
    pct = 100.0 * q / q.sum(0)
    assert np.allclose(o["depth"][v], x[0], rtol=1e-7)
    assert np.allclose(o["b_pct"][:, v], pct, rtol=1e-6, atol=1e-5)
    assert np.array_equal(o["bottom_type"][v], 1.0 + np.argmax(pct, axis=0))
    alb = np.sum(q / q.sum(0) * 0.01 * np.abs(x[1:4]), axis=0)
    assert np.allclose(o["albedo"][v], alb, rtol=1e-6)
