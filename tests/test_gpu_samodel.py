"""The drop-in samodel() (include/photic_samodel.h) on the GPU, called with the reference's
structures and row-pointer grids, against the FP64 oracle's native-mode inversion with the
reference's Nelder-Mead settings, pixel by pixel."""
import ctypes as C

import numpy as np
import pytest

import _oracle as O
import photic_amd as P
from photic_amd.samodel_types import Geogrid, RowGrid, Scene

pytestmark = pytest.mark.gpu


def _call(bands, nodata, theta_w, n_bottoms, emp=None, emp_nodata=-99999.0, smoothing=1):
    nb, nrows, ncols = bands.shape
    grids = [RowGrid(bands[b]) for b in range(nb)]
    gd = (Geogrid * (nb + 1))()
    for b in range(nb):
        gd[b].nrows, gd[b].ncols, gd[b].nodata_value = nrows, ncols, nodata
        gd[b].array = grids[b].pp
    scenes = (Scene * 1)()
    sc = scenes[0]
    sc.n_bands, sc.nrows, sc.ncols = nb, nrows, ncols
    for b, wl in enumerate((443, 482, 561, 655)):
        sc.band_indexes[b] = b
        sc.wavelengths[b] = wl
    sc.theta_w, sc.theta_v = theta_w, 0.0
    idx = (C.c_int * 1)(0)
    empg = Geogrid()
    keep = None
    if emp is not None:
        keep = RowGrid(emp)
        empg.nrows, empg.ncols, empg.nodata_value, empg.array = nrows, ncols, emp_nodata, keep.pp
    outs = {k: RowGrid(np.zeros((nrows, ncols))) for k in
            ("depth", "sigma", "err", "alb", "sand", "grass", "coral", "kmin", "btype", "iod")}
    P.lib.photic_samodel.restype = C.c_int
    rc = P.lib.photic_samodel(scenes, gd, idx, 1, int(emp is not None), empg, smoothing, 1,
                              n_bottoms, *[outs[k].pp for k in ("depth", "sigma", "err", "alb",
                                                                 "sand", "grass", "coral", "kmin",
                                                                 "btype", "iod")])
    return rc, {k: v.array() for k, v in outs.items()}


def _scene(seed, nrows, ncols):
    rrs, ss, truth, nodata, _ = O.gen(seed, nrows * ncols)
    return rrs.reshape(4, nrows, ncols), truth


@pytest.mark.parametrize("nbot", [1, 2, 3])
def test_samodel_dropin_matches_oracle(nbot):
    nrows, ncols = 24, 40
    bands, truth = _scene(21 + nbot, nrows, ncols)
    bands[:, 3, 5] = -9999.0           # nodata on every band
    bands[2, 7, 7] = -9999.0           # nodata on one band
    bands[1, 9, 11] = -0.0005          # negative reflectance
    rc, out = _call(bands, -9999.0, 35.0, nbot)
    assert rc == 0
    ss = 1.0 / np.cos(35.0 * np.pi / 180.0)
    o = O.invert(bands.reshape(4, -1), None, mode=0, n_bottoms=nbot, reqmin=1e-2, konvge=100,
                 kcount=5000, sec_sun0=ss)
    depth = -out["depth"].reshape(-1)
    valid = o["status"] >= 0
    assert np.array_equal(out["btype"].reshape(-1) != -9999.0, valid)
    for (i, j) in ((3, 5), (7, 7), (9, 11)):
        assert out["depth"][i, j] == 0 and out["sand"][i, j] == -9999.0
    ok = np.abs(depth[valid] - o["depth"][valid]) < 0.01
    err = out["err"].reshape(-1)[valid]
    flat = np.abs(err - o["rrs_error"][valid]) <= 1e-6 * np.abs(o["rrs_error"][valid])
    print("nbot %d: %d valid pixels, %.4f with depth within 1 cm of the oracle, identical %.4f" %
          (nbot, valid.sum(), ok.mean(), (depth[valid] == o["depth"][valid]).mean()))
    # outside 1 cm only where the misfit is flat in depth (both end points fit equally well)
    assert np.all(ok | flat) and ok.mean() > 0.99
    same = depth[valid] == o["depth"][valid]
    for k_out, k_or in (("err", "rrs_error"), ("kmin", "k_min"), ("iod", "iod"),
                        ("alb", "albedo"), ("btype", "bottom_type")):
        a = out[k_out].reshape(-1)[valid][same]
        b = o[k_or][valid][same]
        assert np.allclose(a, b, rtol=1e-5, atol=1e-6), k_out
    assert np.allclose(out["sand"].reshape(-1)[valid][same], o["b_pct"][0][valid][same], rtol=1e-5)


def test_samodel_dropin_empirical_depths():
    nrows, ncols = 16, 32
    bands, truth = _scene(31, nrows, ncols)
    rng = np.random.default_rng(5)
    emp = -rng.uniform(0.2, 20.0, (nrows, ncols)).astype(np.float32)
    emp[0, :8] = -99999.0    # no empirical depth here: all depth starts
    emp[1, :8] = -0.5        # > -1: "very shallow", start at 1 m
    rc, out = _call(bands, -9999.0, 30.0, 2, emp=emp)
    assert rc == 0
    ss = 1.0 / np.cos(30.0 * np.pi / 180.0)
    flat = bands.reshape(4, -1)
    e = emp.reshape(-1)
    depth = -out["depth"].reshape(-1)
    agree = []
    for p in range(flat.shape[1]):
        if e[p] == -99999.0:
            o = O.invert(flat[:, p:p + 1], None, mode=0, n_bottoms=2, reqmin=1e-2, kcount=5000,
                         sec_sun0=ss)
        else:
            h = 1.0 if e[p] > -1.0 else abs(e[p])
            nbot = 1 if h > 8.0 else 2
            o = O.invert(flat[:, p:p + 1], None, mode=0, n_bottoms=nbot, bottom_idx=(0, 1, 2),
                         reqmin=1e-2, kcount=5000, hstart=(float(np.float32(h)),), sec_sun0=ss)
        if o["status"][0] >= 0:
            same_fit = abs(out["err"].reshape(-1)[p] - o["rrs_error"][0]) <= 1e-6 * o["rrs_error"][0]
            agree.append((abs(depth[p] - o["depth"][0]) < 0.01, same_fit))
    agree = np.array(agree)
    print("empirical: %d pixels, %.4f within 1 cm" % (len(agree), agree[:, 0].mean()))
    assert np.all(agree[:, 0] | agree[:, 1]) and agree[:, 0].mean() > 0.99


def test_samodel_rejects_unsupported():
    bands, _ = _scene(41, 4, 4)
    rc, out = _call(bands, -9999.0, 30.0, 4)
    assert rc == P.EUNSUPPORTED
    assert np.all(out["btype"] == -9999.0)


def _call_scenes(bands, theta_w, theta_v, n_bottoms, nodata=-9999.0, emp=None,
                 emp_nodata=-99999.0, n_spatial=1, smoothing=1, rsigma=0.0):
    """samodel() over several scenes: bands [nd][4][nrows][ncols], one grid per (scene, band)."""
    nd, nb, nrows, ncols = bands.shape
    grids = [RowGrid(bands[d, b]) for d in range(nd) for b in range(nb)]
    gd = (Geogrid * (nd * nb))()
    for q in range(nd * nb):
        gd[q].nrows, gd[q].ncols, gd[q].nodata_value = nrows, ncols, nodata
        gd[q].array = grids[q].pp
    scenes = (Scene * nd)()
    for d in range(nd):
        sc = scenes[d]
        sc.n_bands, sc.nrows, sc.ncols = nb, nrows, ncols
        for b, wl in enumerate((443, 482, 561, 655)):
            sc.band_indexes[b] = d * nb + b
            sc.wavelengths[b] = wl
        sc.theta_w, sc.theta_v = theta_w[d], theta_v[d]
        for b in range(nb):
            sc.R_sigma[b] = rsigma
    idx = (C.c_int * nd)(*range(nd))
    empg = Geogrid()
    keep = None
    if emp is not None:
        keep = RowGrid(emp)
        empg.nrows, empg.ncols, empg.nodata_value, empg.array = nrows, ncols, emp_nodata, keep.pp
    outs = {k: RowGrid(np.zeros((nrows, ncols))) for k in
            ("depth", "sigma", "err", "alb", "sand", "grass", "coral", "kmin", "btype", "iod")}
    P.lib.photic_samodel.restype = C.c_int
    rc = P.lib.photic_samodel(scenes, gd, idx, nd, int(emp is not None), empg, smoothing,
                              n_spatial, n_bottoms, *[outs[k].pp for k in ("depth", "sigma", "err", "alb",
                                                                 "sand", "grass", "coral", "kmin",
                                                                 "btype", "iod")])
    return rc, {k: v.array() for k, v in outs.items()}


@pytest.mark.parametrize("nbot,with_emp,nd", [(2, False, 3), (1, True, 3), (1, False, 10)])
def test_samodel_dropin_several_scenes(nbot, with_emp, nd):
    """samodel() with several scenes: the reference's N_s = nd run in its native layout, each
    scene with its own sun and view angles, a pixel unmodelled when any scene has no data,
    optionally an empirical depth grid (its own depth start, sand only beyond 8 m) -- against
    the oracle's native multi-date inversion, pixel by pixel.  Up to eight scenes go through
    the stack kernel, more through the region kernel with one region."""
    nrows, ncols = (16, 24) if nd <= 8 else (8, 10)
    n = nrows * ncols
    rrs, _, _, _ = O.gen_stack(71 + nbot + nd, n, n_dates=nd, p_cloud=0.0, k_min=nd)
    bands = rrs.reshape(nd, 4, nrows, ncols).copy()
    bands[1, :, 2, 3] = -9999.0        # no data on one scene
    bands[2, 1, 4, 5] = -0.0004        # negative on one band of one scene
    tw = tuple(30.0 + 15.0 * d / max(nd - 1, 1) for d in range(nd))
    tv = tuple(5.0 * d for d in range(nd))
    emp = None
    if with_emp:
        emp = np.random.default_rng(5).uniform(-15.0, -0.5, (nrows, ncols)).astype(np.float32)
        emp[0, :6] = -99999.0          # no empirical depth here: all depth starts
    rc, out = _call_scenes(bands, tw, tv, nbot, emp=emp)
    assert rc == 0
    ss = np.repeat((1.0 / np.cos(np.radians(tw)))[:, None], n, 1).astype(np.float32)
    sv = np.repeat((1.0 / np.cos(np.radians(tv)))[:, None], n, 1).astype(np.float32)
    r = bands.reshape(nd, 4, n)
    valid = np.ones((nd, n), np.uint8)
    o_depth = np.zeros(n, np.float32)
    o_err = np.zeros(n, np.float32)
    o_status = np.full(n, -1, np.int32)
    groups = [(np.arange(n), None, nbot)]
    if with_emp:
        e = emp.reshape(-1)
        has = e != -99999.0
        h = np.where(e > -1.0, 1.0, np.abs(e)).astype(np.float32)
        groups = [(np.nonzero(~has)[0], None, nbot), (np.nonzero(has & (h <= 8.0))[0], h, nbot),
                  (np.nonzero(has & (h > 8.0))[0], h, 1)]
    for idx, h, nb_ in groups:
        if len(idx) == 0:
            continue
        o = O.invert_stack_native(r[:, :, idx], ss[:, idx], valid[:, idx], n_bottoms=nb_,
                                  sec_view=sv[:, idx], hstart_px=None if h is None else h[idx])
        o_depth[idx], o_err[idx], o_status[idx] = o["depth"], o["rrs_error"], o["status"]
    depth = -out["depth"].reshape(-1)
    ok = o_status >= 0
    assert not ok[2 * ncols + 3] and not ok[4 * ncols + 5]
    assert np.array_equal(out["btype"].reshape(-1) != -9999.0, ok)
    within = np.abs(depth[ok] - o_depth[ok]) < 0.01
    err = out["err"].reshape(-1)[ok]
    flat = np.abs(err - o_err[ok]) <= 1e-6 * np.abs(o_err[ok])
    print("%d scenes, %d bottoms%s: %d pixels, within 1 cm %.4f, identical %.4f" % (
        nd, nbot, " + empirical depths" if with_emp else "", ok.sum(), within.mean(),
        (depth[ok] == o_depth[ok]).mean()))
    assert (within | flat).mean() > 0.995 and within.mean() > 0.98


@pytest.mark.parametrize("nbot,smoothing,with_emp,nd", [(3, 1, False, 1), (2, 2, True, 1),
                                                        (1, 1, False, 2)])
def test_samodel_dropin_regions(nbot, smoothing, with_emp, nd):
    """samodel() with n_spatial = 2, the reference's default: every pixel inverted with its
    3 x 3 neighbourhood of smoothed grids (model/samodel.c:2957-3027), screened on its own raw
    values, optionally an empirical depth grid (its own depth start, sand only beyond 8 m) --
    against the oracle's region inversion of the oracle-smoothed grids, pixel by pixel."""
    nrows, ncols = 10, 12
    n = nrows * ncols
    if nd == 1:
        rrs = O.gen(90 + nbot, n)[0][None]
    else:
        rrs = O.gen_stack(91 + nbot, n, n_dates=nd, p_cloud=0.0, k_min=nd)[0]
    bands = np.ascontiguousarray(rrs.reshape(nd, 4, nrows, ncols), np.float32).copy()
    bands[0, :, 2, 3] = -9999.0         # no data on every band of a scene
    bands[nd - 1, 1, 4, 5] = -0.0004    # negative on one band
    bands[0, 2, 6, 0] = -9999.0         # no data on one band, at the edge
    tw, tv = (30.0, 41.0)[:nd], (0.0, 6.0)[:nd]
    emp = None
    if with_emp:
        emp = np.random.default_rng(6).uniform(-15.0, -0.5, (nrows, ncols)).astype(np.float32)
        emp[0, :6] = -99999.0           # no empirical depth here: all depth starts
        emp[1, :3] = -0.5               # "very shallow": start at 1 m
    rc, out = _call_scenes(bands, tw, tv, nbot, emp=emp, n_spatial=2, smoothing=smoothing)
    assert rc == 0
    ss = np.array([O.sec_deg(t) for t in tw])
    sv = np.array([O.sec_deg(t) for t in tv])
    raw = bands.reshape(nd, 4, n)
    sm = np.stack([np.stack([O.smooth_grid(bands[d, b], smoothing, -9999.0) for b in range(4)])
                   for d in range(nd)]).reshape(nd, 4, n)
    groups = [(np.arange(n), None, nbot)]
    if with_emp:
        e = emp.reshape(-1)
        has = e != -99999.0
        h = np.where(e > -1.0, 1.0, np.abs(e)).astype(np.float32)
        groups = [(np.nonzero(~has)[0], None, nbot), (np.nonzero(has & (h <= 8.0))[0], h, nbot),
                  (np.nonzero(has & (h > 8.0))[0], h, 1)]
    o_depth = np.zeros(n, np.float32)
    o_err = np.zeros(n, np.float32)
    o_status = np.full(n, -1, np.int32)
    o_btype = np.full(n, -9999.0, np.float32)
    o_kmin = np.zeros(n, np.float32)
    o_iod = np.zeros(n, np.float32)
    o_alb = np.zeros(n, np.float32)
    o_sand = np.full(n, -9999.0, np.float32)
    for idx, h, nb_ in groups:
        if len(idx) == 0:
            continue
        o = O.invert_regions(sm, nrows, ncols, idx, -9999.0, ss, sv, n_spatial=2, n_bottoms=nb_,
                             bottom_idx=(0, 1, 2), hstart_px=None if h is None else h[idx],
                             screen_rrs=raw)
        v = o["status"] >= 0
        sel = idx[v]
        o_status[idx] = o["status"]
        o_depth[sel], o_err[sel], o_btype[sel] = o["depth"][v], o["rrs_error"][v], o["bottom_type"][v]
        o_kmin[sel], o_iod[sel], o_alb[sel] = o["k_min"][v], o["iod"][v], o["albedo"][v]
        o_sand[sel] = o["b_pct"][0][v]
    ok = o_status >= 0
    assert not ok[2 * ncols + 3] and not ok[4 * ncols + 5] and not ok[6 * ncols + 0]
    assert np.array_equal(out["btype"].reshape(-1) != -9999.0, ok)
    depth = -out["depth"].reshape(-1)
    same = depth[ok] == o_depth[ok]
    err = out["err"].reshape(-1)[ok]
    flat = np.abs(err - o_err[ok]) <= 1e-6 * np.abs(o_err[ok])
    print("regions drop-in: %d bottoms, radius %d, %d scenes%s: %d modelled, identical depth "
          "%.4f" % (nbot, smoothing, nd, ", empirical depths" if with_emp else "", ok.sum(),
                    same.mean()))
    assert np.all(same | flat) and same.mean() >= 0.98
    for a, b in ((out["err"], o_err), (out["kmin"], o_kmin), (out["iod"], o_iod),
                 (out["alb"], o_alb), (out["btype"], o_btype), (out["sand"], o_sand)):
        assert np.array_equal(a.reshape(-1)[ok][same], b[ok][same])


@pytest.mark.parametrize("nbot,ns,nd,nrows,ncols,with_emp", [(2, 1, 1, 12, 16, False),
                                                             (3, 2, 1, 6, 7, False),
                                                             (1, 1, 2, 10, 10, True)])
def test_samodel_depth_sigma(nbot, ns, nd, nrows, ncols, with_emp):
    """The depth-error pass (model/samodel.c:1376-1482) of the drop-in against the oracle's
    restatement, given the same retrieved depths: the same Philox-drawn samples, each
    perturbed by n_sigma R_sigma and inverted again -- the first with the main pass's start
    rules, every later one (start_at_previous) from the depth of the main pass's last modelled
    pixel, or from its own empirical depth -- per-interval standard deviations of the sampled
    depths in double, written to the float grid.  At most two intervals not bit-identical, each
    within 10 %: a sample whose Nelder-Mead path splits from the oracle's (the oracle takes
    exp / log / pow from the host C library, which misrounds about one argument in 10^4; the
    kernels round correctly) moves its interval's sigma, the more the fewer samples found a
    pixel in that interval; the single-start samples split more readily than multi-start ones
    (seen: one interval of 28 off by 5.5 % with 3 bottoms and 3 x 3 regions)."""
    n = nrows * ncols
    if nd == 1:
        rrs = O.gen(120 + nbot, n)[0][None]
    else:
        rrs = O.gen_stack(121 + nbot, n, n_dates=nd, p_cloud=0.0, k_min=nd)[0]
    bands = np.ascontiguousarray(rrs.reshape(nd, 4, nrows, ncols), np.float32).copy()
    bands[0, :, 1, 2] = -9999.0
    tw, tv = (33.0, 44.0)[:nd], (0.0, 4.0)[:nd]
    emp = None
    if with_emp:
        emp = np.random.default_rng(8).uniform(-14.0, -0.5, (nrows, ncols)).astype(np.float32)
        emp[0, :4] = -99999.0
    rsig = 3e-4
    rc, out = _call_scenes(bands, tw, tv, nbot, emp=emp, n_spatial=ns, rsigma=rsig)
    assert rc == 0
    depth = -out["depth"].reshape(-1)
    ss = np.array([O.sec_deg(t) for t in tw])
    sv = np.array([O.sec_deg(t) for t in tv])
    table, sigma = O.depth_sigma(depth, bands.reshape(nd, 4, n), nrows, ncols, -9999.0,
                                 np.full((nd, 4), rsig), ss, sv, n_spatial=ns, n_bottoms=nbot,
                                 emp=None if emp is None else emp.reshape(-1))
    got = out["sigma"].reshape(-1)
    print("depth sigma: %d intervals, %d non-zero, max %.3f m" %
          (len(table), int((table > 0).sum()), float(table.max()) if len(table) else 0.0))
    assert len(table) > 0 and (table > 0).sum() > 0
    k = np.ceil(depth / 0.25).astype(np.int64) - 1  # the pixel's interval (d, d + 0.25]
    m = (depth > 0) & (k < len(table))
    split = np.unique(k[m & (got != sigma)])
    print("  intervals not bit-identical: %s" % split.tolist())
    assert len(split) <= 2
    assert np.allclose(got, sigma, rtol=0.1, atol=1e-6)
    assert np.all(got[depth <= 0] == 0.0) and np.all(got[depth > 0] >= 0.0)


def test_samodel_smoothed_single_region_keeps_negative_means():
    """n_smoothing_radius 2 with one region: pixels are screened on their own raw values
    (model/samodel.c:933-947) and inverted with the smoothed ones (:2997-3007), so a pixel
    whose raw bands are valid but whose 3 x 3 mean goes negative next to a strongly negative
    neighbour is still modelled -- against the oracle's one-region inversion of the
    oracle-smoothed grid screened on the raw grid."""
    nrows, ncols = 12, 16
    n = nrows * ncols
    bands = np.ascontiguousarray(O.gen(131, n)[0].reshape(4, nrows, ncols), np.float32).copy()
    bands[0, 5, 7] = -0.5               # raw negative: this pixel is not modelled
    bands[:, 9, 2] = -9999.0            # no data on every band
    rc, out = _call(bands, -9999.0, 35.0, 2, smoothing=2)
    assert rc == 0
    sm = np.stack([O.smooth_grid(bands[b], 2, -9999.0) for b in range(4)]).reshape(1, 4, n)
    neigh = [(5 + di) * ncols + 7 + dj for di in (-1, 0, 1) for dj in (-1, 0, 1) if di or dj]
    assert np.all(sm[0, 0, neigh] < 0.0)   # the case: valid raw values, negative means
    o = O.invert_regions(sm, nrows, ncols, np.arange(n), -9999.0,
                         np.array([O.sec_deg(35.0)]), np.array([1.0]), n_spatial=1,
                         n_bottoms=2, bottom_idx=(0, 1, 2), screen_rrs=bands.reshape(1, 4, n))
    ok = o["status"] >= 0
    assert np.all(ok[neigh]) and not ok[5 * ncols + 7] and not ok[9 * ncols + 2]
    assert np.array_equal(out["btype"].reshape(-1) != -9999.0, ok)
    depth = -out["depth"].reshape(-1)
    same = depth[ok] == o["depth"][ok]
    err = out["err"].reshape(-1)[ok]
    flat = np.abs(err - o["rrs_error"][ok]) <= 1e-6 * np.abs(o["rrs_error"][ok])
    print("smoothed, one region: %d modelled, identical depth %.4f" % (ok.sum(), same.mean()))
    assert np.all(same | flat) and same.mean() >= 0.98


@pytest.mark.parametrize("nd,n_spatial", [(1, 1), (2, 1), (1, 2)])
def test_samodel_watchdog_fails_the_call(nd, n_spatial, monkeypatch):
    """A kernel wavefront that finishes no pixel within the watchdog's limit gives up; the
    drop-in then returns PHOTIC_EDEVICE and leaves every output grid at the reference's
    defaults (model/samodel.c:819-841).  The limit is forced to 1 ns from the host, so every
    wave leaves its loop after its first pass and each kernel ends normally."""
    nrows, ncols = 8, 64
    n = nrows * ncols
    if nd == 1:
        rrs = O.gen(141, n)[0][None]
    else:
        rrs = O.gen_stack(142, n, n_dates=nd, p_cloud=0.0, k_min=nd)[0]
    bands = np.ascontiguousarray(rrs.reshape(nd, 4, nrows, ncols), np.float32).copy()
    monkeypatch.setenv("PHOTIC_WATCHDOG_S", "1e-9")
    rc, out = _call_scenes(bands, (30.0, 40.0)[:nd], (0.0, 0.0)[:nd], 2, n_spatial=n_spatial)
    assert rc == P.EDEVICE
    assert "watchdog" in P.last_error()
    for k in ("depth", "sigma", "err", "alb", "kmin", "iod"):
        assert np.all(out[k] == 0.0), k
    for k in ("sand", "grass", "coral", "btype"):
        assert np.all(out[k] == -9999.0), k
    monkeypatch.delenv("PHOTIC_WATCHDOG_S")
    rc, out = _call_scenes(bands, (30.0, 40.0)[:nd], (0.0, 0.0)[:nd], 2, n_spatial=n_spatial)
    assert rc == 0 and np.any(out["btype"] != -9999.0)


def _call_band_sets(scene_grids, wls, theta_w, theta_v, n_bottoms, n_spatial, nodata=-9999.0):
    """samodel() over scenes with their own band sets: scene_grids[s] [nb_s][nrows][ncols] at
    wavelengths wls[s] (integers, as the reference's scene structure holds them)."""
    nd = len(scene_grids)
    nrows, ncols = scene_grids[0].shape[1:]
    flat = [g for sg in scene_grids for g in sg]
    grids = [RowGrid(np.ascontiguousarray(g, np.float32)) for g in flat]
    gd = (Geogrid * len(flat))()
    for q in range(len(flat)):
        gd[q].nrows, gd[q].ncols, gd[q].nodata_value = nrows, ncols, nodata
        gd[q].array = grids[q].pp
    scenes = (Scene * nd)()
    at = 0
    for d in range(nd):
        sc = scenes[d]
        sc.n_bands, sc.nrows, sc.ncols = len(wls[d]), nrows, ncols
        for b, wl in enumerate(wls[d]):
            sc.band_indexes[b] = at
            sc.wavelengths[b] = int(wl)
            at += 1
        sc.theta_w, sc.theta_v = theta_w[d], theta_v[d]
    idx = (C.c_int * nd)(*range(nd))
    empg = Geogrid()
    outs = {k: RowGrid(np.zeros((nrows, ncols))) for k in
            ("depth", "sigma", "err", "alb", "sand", "grass", "coral", "kmin", "btype", "iod")}
    P.lib.photic_samodel.restype = C.c_int
    rc = P.lib.photic_samodel(scenes, gd, idx, nd, 0, empg, 1, n_spatial, n_bottoms,
                              *[outs[k].pp for k in ("depth", "sigma", "err", "alb", "sand",
                                                     "grass", "coral", "kmin", "btype", "iod")])
    return rc, {k: v.array() for k, v in outs.items()}


@pytest.mark.parametrize("wls,n_spatial", [(((443, 482, 561, 655), (443, 490, 560, 620, 665)), 1),
                                           (((412, 443, 490, 531, 555, 670),), 2)])
def test_samodel_dropin_band_sets(wls, n_spatial):
    """samodel() with each scene's own band set (a five- beside a four-band scene, or one
    six-band scene): every pixel through the region kernel, against the oracle's region
    inversion with per-scene bands, pixel by pixel."""
    nrows, ncols = 6, 8
    n = nrows * ncols
    nd = len(wls)
    scene_grids = [O.gen(60 + s, n, wl=tuple(float(w) for w in wls[s]))[0]
                   .reshape(len(wls[s]), nrows, ncols).copy() for s in range(nd)]
    scene_grids[0][:, 2, 3] = -9999.0    # no data on every band of a scene
    scene_grids[-1][1, 4, 5] = -0.0004   # negative on one band
    tw = tuple(30.0 + 8.0 * s for s in range(nd))
    tv = tuple(3.0 * s for s in range(nd))
    rc, out = _call_band_sets(scene_grids, wls, tw, tv, 2, n_spatial)
    assert rc == 0
    rows = np.concatenate([g.reshape(len(g), n) for g in scene_grids]).astype(np.float32)
    ss = np.array([O.sec_deg(t) for t in tw])
    sv = np.array([O.sec_deg(t) for t in tv])
    o = O.invert_regions_bands(rows, [tuple(float(w) for w in ws) for ws in wls], nrows, ncols,
                               np.arange(n), -9999.0, ss, sv, n_spatial=n_spatial, n_bottoms=2,
                               bottom_idx=(0, 1, 2), screen=True)
    ok = o["status"] >= 0
    assert not ok[2 * ncols + 3] and not ok[4 * ncols + 5]
    assert np.array_equal(out["btype"].reshape(-1) != -9999.0, ok)
    depth = -out["depth"].reshape(-1)
    same = depth[ok] == o["depth"][ok]
    err = out["err"].reshape(-1)[ok]
    flat = np.abs(err - o["rrs_error"][ok]) <= 1e-6 * np.abs(o["rrs_error"][ok])
    print("band sets %s, n_spatial %d: %d modelled, identical depth %.4f" %
          ([len(w) for w in wls], n_spatial, ok.sum(), same.mean()))
    assert np.all(same | flat) and same.mean() >= 0.98
