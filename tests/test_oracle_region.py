"""Pins of the oracle's spatial-region restatement (oracle/orc_region.c), the reference's
n_spatial > 1 run (model/samodel.c:2957-3027 gather, :2432-2759 objective, :2246-2354 starts,
:1992-2079 outputs).

* With one region every function must reduce bit for bit to the single-region oracle, which
  is pinned elsewhere (closed forms, the reference's compiled nelmin).
* The continuity penalties against values worked out by hand.
* The K penalty reads the origin's depth and the last region's attenuation.
* The gather on hand-built grids: clamped duplicates at edges, dropped no-data neighbours,
  kept negative neighbours, the origin's index.
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import oracle_bindings as O  # noqa: E402
from conftest import needs_ref  # noqa: E402


def _pixels(seed, n):
    rrs = O.gen(seed, n)[0]  # [nb][n]
    return rrs.T.astype(np.float64)  # [n][nb]


def _native_x(rng, nbot, nd):
    x = [rng.uniform(0.5, 20.0)]
    x += list(rng.uniform(5.0, 60.0, nbot))
    x += list(rng.uniform(0.1, 2.0, nbot))
    for _ in range(nd):
        x += [rng.uniform(0.5, 8.0), rng.uniform(0.2, 6.0), rng.uniform(0.05, 2.0)]
    return np.array(x) * rng.choice([-1.0, 1.0], len(x))


@pytest.mark.parametrize("nbot,nd", [(1, 1), (2, 1), (3, 1), (3, 2), (2, 4)])
def test_one_region_objective_is_the_single_region_objective(nbot, nd):
    rng = np.random.default_rng(10 * nbot + nd)
    px = _pixels(100 + nbot, 6 * nd).reshape(6, nd, 4)
    ss = rng.uniform(1.05, 1.5, nd).astype(np.float32)
    for t in range(6):
        x = _native_x(rng, nbot, nd)
        r32 = px[t].astype(np.float32)
        rrs_err = np.zeros(1)
        w = np.asarray(O.WL4)
        f1 = O.lib().orc_py_objective_stack_ex(w.ctypes.data, 4, nd, 0, nbot, x.ctypes.data,
                                                np.ascontiguousarray(r32).ctypes.data,
                                                ss.ctypes.data, rrs_err.ctypes.data)
        f2, terms = O.region_objective(x, r32.astype(np.float64)[None], origin=0,
                                       sec_sun=ss.astype(np.float64), n_bottoms=nbot)
        assert f1 == f2
        assert terms[0] == rrs_err[0]
        assert terms[1] == 0.0 and terms[2] == 0.0


@pytest.mark.parametrize("nbot,nd", [(1, 1), (3, 1), (2, 3)])
def test_one_region_start_step_is_the_single_region_rule(nbot, nd):
    px = _pixels(7 + nbot, nd).reshape(1, nd, 4).astype(np.float32).astype(np.float64)
    s1, t1 = np.zeros(64), np.zeros(64)
    w = np.asarray(O.WL4)
    r = np.ascontiguousarray(px[0])
    n = O.lib().orc_py_start_step_stack(0, nbot, w.ctypes.data, 4, nd, r.ctypes.data, 15.0,
                                        s1.ctypes.data, t1.ctypes.data)
    s2, t2 = O.region_start_step(px, 15.0, n_bottoms=nbot)
    assert n == len(s2)
    assert np.array_equal(s1[:n], s2) and np.array_equal(t1[:n], t2)


@pytest.mark.parametrize("nbot,nd", [(1, 1), (2, 1), (3, 2)])
def test_one_region_inversion_is_the_single_region_inversion(nbot, nd):
    """A 1 x n grid with n_spatial = 1: every pixel is its own single region, so the region
    inversion must give the single-region native inversion of the same pixel bit for bit."""
    n = 24
    rrs, ss, valid, _ = O.gen_stack(40 + nbot, n, n_dates=nd, k_min=nd)
    valid[:] = 1
    rrs = np.maximum(rrs, 1e-4).astype(np.float32)
    sec_sun = np.full(nd, 1.2, np.float32).astype(np.float64)
    sec_view = np.full(nd, 1.05, np.float32).astype(np.float64)
    ss_px = np.broadcast_to(sec_sun.astype(np.float32)[:, None], (nd, n)).copy()
    sv_px = np.broadcast_to(sec_view.astype(np.float32)[:, None], (nd, n)).copy()
    a = O.invert_stack_native(rrs, ss_px, valid, n_bottoms=nbot, sec_view=sv_px)
    b = O.invert_regions(rrs, 1, n, np.arange(n), -9999.0, sec_sun, sec_view, n_spatial=1,
                         n_bottoms=nbot)
    npar = 1 + 2 * nbot + 3 * nd
    assert np.array_equal(b["n_regions"], np.ones(n))
    assert np.array_equal(a["x"][:npar], b["x"][:npar])
    for k in ("depth", "rrs_error", "model_error", "k_min", "iod", "albedo", "bottom_type",
              "b_pct", "n_evals", "n_starts", "n_iter", "status"):
        assert np.array_equal(a[k], b[k]), k


def _region_rrs(R, nd=1, seed=3):
    return _pixels(seed, R * nd).reshape(R, nd, 4).astype(np.float32).astype(np.float64)


def _region_x(H, B, q, pgx):
    """Native region vector [H_r], [B_{r,k}], [q_{r,k}], [(P, G, X)_s] from nested lists."""
    return np.array(list(H) + [v for b in B for v in b] + [v for w in q for v in w] +
                    [v for t in pgx for v in t], dtype=np.float64)


def test_depth_continuity_penalty_by_hand():
    # H = 2, 2, 5: mean 3 < 4 so the band is +-40% (1.8 .. 4.2); only 5 lies outside:
    # penalty = 100 sqrt((5 - 3)^2 / 1) / 3 = 200 / 3
    r = _region_rrs(3)
    x = _region_x([2.0, 2.0, 5.0], [[30.0]] * 3, [[1.0]] * 3, [(2.0, 3.0, 0.5)])
    _, t = O.region_objective(x, r, n_bottoms=1)
    assert t[1] == pytest.approx(200.0 / 3.0, rel=1e-15)
    # H = 10, 11, 12, 13 (negative signs read through |.|): mean 11.5 < 12, band +-10%
    # (10.35 .. 12.65): 10 and 13 are outside, (1.5^2 + 1.5^2) / 2 = 2.25 -> 100 * 1.5 / 11.5
    x = _region_x([10.0, -11.0, 12.0, -13.0], [[30.0]] * 4, [[1.0]] * 4, [(2.0, 3.0, 0.5)])
    _, t = O.region_objective(x, _region_rrs(4), n_bottoms=1)
    assert t[1] == pytest.approx(150.0 / 11.5, rel=1e-15)
    # equal depths: no outlier, no penalty
    x = _region_x([7.0] * 5, [[30.0]] * 5, [[1.0]] * 5, [(2.0, 3.0, 0.5)])
    _, t = O.region_objective(x, _region_rrs(5), n_bottoms=1)
    assert t[1] == 0.0 and t[2] == 0.0


def test_bottom_continuity_penalty_by_hand():
    # two regions at H = 10 (mean 10: band +-5%), two bottoms.
    # region 0: B = (100, 50), q = (1, 1): q_k B_k / sum q = 50, 25
    # region 1: B = (100, 100), q = (1, 3):                  25, 75
    # bottom 0: mean 37.5, both outside +-5%: 2 x 12.5^2;  bottom 1: mean 50, 2 x 25^2
    # penalty = 100 sqrt((2 * 156.25 + 2 * 625) / 4) / ((37.5 + 50) / 2)
    r = _region_rrs(2)
    x = _region_x([10.0, 10.0], [[100.0, 50.0], [100.0, 100.0]], [[1.0, 1.0], [1.0, 3.0]],
                  [(2.0, 3.0, 0.5)])
    f, t = O.region_objective(x, r, n_bottoms=2, bottom_idx=(0, 1))
    assert t[2] == pytest.approx(100.0 * np.sqrt(1562.5 / 4.0) / 43.75, rel=1e-15)
    # the weighted sum: (80 E_Rrs + 15 E_H + 10 E_B + 15 K) / 120
    assert f == pytest.approx((80 * t[0] + 15 * t[1] + 10 * t[2] + 15 * t[3]) / 120, rel=1e-15)
    # the same fractions in both regions: no penalty, whatever the amplitudes' scale
    x = _region_x([10.0, 10.0], [[100.0, 50.0], [100.0, 50.0]], [[2.0, 2.0], [1.0, 1.0]],
                  [(2.0, 3.0, 0.5)])
    _, t = O.region_objective(x, r, n_bottoms=2, bottom_idx=(0, 1))
    assert t[2] == 0.0


def test_k_penalty_uses_the_origin_depth_and_the_last_region():
    # clear water (small P, G, X) so that K is low and the shallow tiers apply
    r = _region_rrs(2, seed=11)
    x = _region_x([0.5, 20.0], [[30.0], [30.0]], [[1.0], [1.0]], [(0.2, 0.1, 0.02)])
    _, t0 = O.region_objective(x, r, origin=0, n_bottoms=1)
    _, t1 = O.region_objective(x, r, origin=1, n_bottoms=1)
    assert t0[3] > 0.0 and t1[3] == 0.0
    # the K tested is the last region's: equal depths, then the two-region K penalty is the
    # single-region one on the last region's data, not the first's
    r = np.stack([_region_rrs(1, seed=12)[0], _region_rrs(1, seed=13)[0]])
    x2 = _region_x([0.8, 0.8], [[30.0], [30.0]], [[1.0], [1.0]], [(0.2, 0.1, 0.02)])
    x1 = _region_x([0.8], [[30.0]], [[1.0]], [(0.2, 0.1, 0.02)])
    _, t = O.region_objective(x2, r, n_bottoms=1)
    _, a = O.region_objective(x1, r[1:], n_bottoms=1)
    _, b = O.region_objective(x1, r[:1], n_bottoms=1)
    assert a[3] != b[3]
    assert t[3] == a[3]
    # the albedo output is the last region's sum q_k B_k
    x2 = _region_x([5.0, 5.0], [[30.0], [70.0]], [[1.0], [1.0]], [(2.0, 3.0, 0.5)])
    _, t = O.region_objective(x2, r, n_bottoms=1)
    assert t[4] == pytest.approx(0.7, rel=1e-15)


def _grid(n_rows, n_cols, nd=1, seed=0):
    rng = np.random.default_rng(seed)
    return rng.uniform(0.001, 0.02, (nd, 4, n_rows * n_cols)).astype(np.float32)


def test_gather_interior_and_edges():
    g = _grid(4, 5)
    v = g.reshape(1, 4, 4, 5)
    # interior pixel: the 3 x 3 block in row-major order, origin in the middle
    reg, org = O.region_gather(g, 4, 5, 2, 2, 2, -9999.0)
    assert reg.shape[0] == 9 and org == 4
    k = 0
    for ii in (-1, 0, 1):
        for jj in (-1, 0, 1):
            assert np.array_equal(reg[k, 0], v[0, :, 2 + ii, 2 + jj].astype(np.float64))
            k += 1
    # corner (0, 0): clamped duplicates are kept; the centre is visited at candidates 0, 1, 3
    # and 4, and the last visit (4) is the origin
    reg, org = O.region_gather(g, 4, 5, 0, 0, 2, -9999.0)
    assert reg.shape[0] == 9 and org == 4
    assert np.array_equal(reg[0], reg[4]) and np.array_equal(reg[1], reg[4])
    # n_spatial 1 (and 0, read as 1): the pixel alone
    for ns in (0, 1):
        reg, org = O.region_gather(g, 4, 5, 3, 4, ns, -9999.0)
        assert reg.shape[0] == 1 and org == 0
        assert np.array_equal(reg[0, 0], v[0, :, 3, 4].astype(np.float64))
    # n_spatial 3: a 5 x 5 block
    reg, org = O.region_gather(g, 4, 5, 2, 2, 3, -9999.0)
    assert reg.shape[0] == 25 and org == 12


def test_gather_nodata_and_negative_neighbours():
    g = _grid(3, 3, nd=2, seed=1)
    nod = np.full((2, 4), -9999.0, np.float32)
    # (0, 1) no-data on scene 1's band 2 only: that candidate is dropped
    g[1, 2, 0 * 3 + 1] = -9999.0
    # (2, 2) negative: kept (the neighbour test is approx_equal to no-data only)
    g[0, 0, 2 * 3 + 2] = -0.001
    reg, org = O.region_gather(g, 3, 3, 1, 1, 2, nod)
    assert reg.shape[0] == 8 and org == 3
    assert reg[7, 0, 0] == np.float64(np.float32(-0.001))
    # the centre itself no-data: the origin is the index it would have had, which the next
    # valid candidate then takes
    g2 = _grid(3, 3, seed=2)
    g2[0, 1, 1 * 3 + 1] = -9999.0
    reg, org = O.region_gather(g2, 3, 3, 1, 1, 2, -9999.0)
    assert reg.shape[0] == 8 and org == 4
    assert np.array_equal(reg[4, 0], g2[0, :, 1 * 3 + 2].astype(np.float64))
    # per-band no-data values
    g3 = _grid(3, 3, seed=3)
    g3[0, 3, 0] = 7.0
    reg, _ = O.region_gather(g3, 3, 3, 1, 1, 2, np.array([[-9999.0, -9999.0, -9999.0, 7.0]]))
    assert reg.shape[0] == 8


def test_region_start_step_by_hand():
    # two regions, one scene: amplitudes at 100 * 4 * mean Rrs(490) over both regions, every
    # weight 1 (step 0.5), every depth at h0 (step 1.25 h0); P from the mean 490/550 ratio
    r = np.array([[[0.010, 0.012, 0.008, 0.002]], [[0.014, 0.016, 0.012, 0.004]]])
    s, t = O.region_start_step(r, 10.0, n_bottoms=2)
    assert len(s) == 2 * 5 + 3
    assert np.array_equal(s[:2], [10.0, 10.0]) and np.array_equal(t[:2], [12.5, 12.5])
    w490 = (490.0 - 482.0) / (561.0 - 482.0)
    r490 = np.array([0.012 * (1 - w490) + 0.008 * w490, 0.016 * (1 - w490) + 0.012 * w490])
    assert s[2:6] == pytest.approx([400.0 * r490.mean()] * 4, rel=1e-14)
    assert np.array_equal(s[6:10], [1.0] * 4) and np.array_equal(t[6:10], [0.5] * 4)
    assert t[2:6] == pytest.approx(1.5 * s[2:6], rel=1e-15)
    assert np.array_equal(t[10:], 2.0 * s[10:])


@needs_ref
@pytest.mark.parametrize("r", [1, 2, 3])
def test_smoothing_matches_reference(r):
    """The oracle's smoothing against the reference's own compiled smoothed_index_2d
    (model/common.c:232-276) at every pixel of a grid with no-data values, bit for bit."""
    rng = np.random.default_rng(r)
    g = rng.uniform(0.0005, 0.03, (9, 11)).astype(np.float32)
    g[rng.random(g.shape) < 0.2] = -9999.0
    g[0:3, 0:3] = -9999.0  # a corner with no data at all in radius 2
    s = O.smooth_grid(g, r, -9999.0)
    for i in range(g.shape[0]):
        for j in range(g.shape[1]):
            assert s[i, j] == O.ref_smoothed(g, i, j, r, -9999.0), (i, j)
    if r == 1:
        assert np.array_equal(s, g)
    if r == 2:
        assert s[0, 0] == -9999.0


WL5 = (443.0, 490.0, 560.0, 620.0, 665.0)


def test_scene_band_sets_match_the_shared_band_path():
    """Every scene given the same four bands through the per-scene entry gives the shared-band
    objective bit for bit."""
    rng = np.random.default_rng(3)
    px = _pixels(30, 9 * 2).reshape(9, 2, 4).astype(np.float32).astype(np.float64)
    assert np.all(px > 0.0)
    for t in range(5):
        x = _native_x(rng, 2, 2)
        f1, t1 = O.region_objective(x, px, origin=4, sec_sun=[1.1, 1.2], sec_view=[1.0, 1.05],
                                    n_bottoms=2)
        f2, t2 = O.region_objective_bands(x, px.reshape(9, 8), [O.WL4, O.WL4], origin=4,
                                          sec_sun=[1.1, 1.2], sec_view=[1.0, 1.05], n_bottoms=2)
        assert np.array_equal([f1], [f2], equal_nan=True)
        assert np.array_equal(t1, t2, equal_nan=True)


@pytest.mark.parametrize("H", [1.5, 3.5, 20.0])
def test_scene_band_sets_enter_the_misfit_scene_by_scene(H):
    """Two scenes with their own band sets (four and five bands): the joint misfit is formed
    from each scene alone. E_Rrs = 100 sqrt(err / n) / (tot / n) over all (scene, band) pairs,
    so each scene's squared residual sum follows from its own E_Rrs and the joint E_Rrs from
    their sum over nine pairs; the K penalty is the sum of each scene's own (each over its own
    bands' K). A scene evaluated with the other's band table or band count fails this."""
    a = O.gen(41, 1)[0][:, 0].astype(np.float64)
    b = O.gen(41, 1, wl=WL5)[0][:, 0].astype(np.float64)
    rng = np.random.default_rng(int(H * 10))
    for _ in range(4):
        x = np.concatenate([[H], rng.uniform(5, 60, 1), rng.uniform(0.1, 2.0, 1)])
        pa = [rng.uniform(0.2, 3.0), rng.uniform(0.1, 2.0), rng.uniform(0.02, 0.6)]
        pb = [rng.uniform(0.2, 3.0), rng.uniform(0.1, 2.0), rng.uniform(0.02, 0.6)]
        _, ta = O.region_objective_bands(np.concatenate([x, pa]), a[None], [O.WL4],
                                         sec_sun=1.1, n_bottoms=1)
        _, tb = O.region_objective_bands(np.concatenate([x, pb]), b[None], [WL5],
                                         sec_sun=1.2, sec_view=1.04, n_bottoms=1)
        _, tab = O.region_objective_bands(np.concatenate([x, pa, pb]),
                                          np.concatenate([a, b])[None], [O.WL4, WL5],
                                          sec_sun=[1.1, 1.2], sec_view=[1.0, 1.04], n_bottoms=1)

        def err(e, r):
            return (e / 100.0 * r.sum() / len(r)) ** 2 * len(r)

        n = 9
        e_ab = 100.0 * np.sqrt((err(ta[0], a) + err(tb[0], b)) / n) / ((a.sum() + b.sum()) / n)
        assert tab[0] == pytest.approx(e_ab, rel=1e-12)
        assert tab[3] == pytest.approx(ta[3] + tb[3], rel=1e-14, abs=1e-300)
