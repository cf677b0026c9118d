"""The oracle pinned against the reference's own compiled code (oracle/_ref: model/asa047.c
and model/common.c built as they lie) and against known answers."""
import numpy as np
import pytest

import _oracle as O

from conftest import needs_ref


# --- Philox4x32-10 known-answer vectors (Random123 kat_vectors, Salmon et al. 2011) ---
@pytest.mark.parametrize("ctr,key,expect", [
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
])
def test_philox_kat(ctr, key, expect):
    c = np.array(ctr, np.uint32)
    k = np.array(key, np.uint32)
    out = np.zeros(4, np.uint32)
    O.lib().orc_py_philox(c.ctypes.data, k.ctypes.data, out.ctypes.data)
    assert tuple(int(v) for v in out) == expect


# --- interp_1d / approx_equal: bit-identical to the reference's common.c ---
@needs_ref
def test_interp_matches_reference():
    rng = np.random.default_rng(7)
    X = np.arange(400.0, 801.0, 10.0)
    Y = rng.uniform(0, 3, X.size)
    qs = list(rng.uniform(380, 820, 2000)) + list(X) + [440.0, 490.0, 550.0, 640.0, 750.0,
                                                        443.0, 482.0, 561.0, 655.0, 399.999999]
    for x in qs:
        a = O.lib().orc_py_interp(X.ctypes.data, Y.ctypes.data, X.size, float(x))
        b = O.ref().interp_1d(X.ctypes.data, Y.ctypes.data, X.size, float(x))
        assert a == b, x
    # the band set of the configs, extrapolating at 440 and 750
    Xb = np.array(O.WL4)
    for _ in range(500):
        Yb = rng.uniform(0, 0.05, 4)
        for x in (440.0, 490.0, 550.0, 640.0, 750.0):
            assert (O.lib().orc_py_interp(Xb.ctypes.data, Yb.ctypes.data, 4, x) ==
                    O.ref().interp_1d(Xb.ctypes.data, Yb.ctypes.data, 4, x))


@needs_ref
def test_approx_equal_matches_reference():
    rng = np.random.default_rng(3)
    a = rng.normal(0, 1, 3000).astype(np.float32)
    b = (a * (1 + rng.normal(0, 1e-6, a.size))).astype(np.float32)
    for x, y in zip(a, b):
        for eps in (1e-6, 1e-5, 1e-4):
            assert O.lib().orc_py_approx_equal(x, y, eps) == O.ref().approx_equal(x, y, eps)
    assert O.lib().orc_py_approx_equal(np.float32(0.0), np.float32(0.0), 1e-6) == 1


# --- nelmin: the restatement against the reference's compiled routine ---
def _same(a, b):
    return (np.array_equal(a["xmin"], b["xmin"]) and a["ynewlo"] == b["ynewlo"]
            and a["icount"] == b["icount"] and a["numres"] == b["numres"]
            and a["ifault"] == b["ifault"] and np.array_equal(a["start_after"], b["start_after"]))


@needs_ref
@pytest.mark.parametrize("kind", [0, 1, 2])
@pytest.mark.parametrize("n", [1, 2, 5, 8])
def test_nelmin_bit_identical(kind, n):
    rng = np.random.default_rng(100 * kind + n)
    for trial in range(6):
        start = rng.normal(0, 2, n)
        step = np.abs(rng.normal(1, 0.5, n)) + 0.1
        w = rng.uniform(0.5, 4, n)
        c = rng.normal(0, 1, n)
        for konvge, kcount, reqmin in ((10, 5000, 1e-10), (100, 2000, 1e-10), (5, 300, 1e-2)):
            a = O.nelmin_test(kind, start, step, w, c, reqmin, konvge, kcount, use_ref=False)
            b = O.nelmin_test(kind, start, step, w, c, reqmin, konvge, kcount, use_ref=True)
            assert _same(a, b), (kind, n, trial, konvge, a, b)
            assert a["evals"] == a["icount"]


def test_nelmin_known_minima():
    # quadratic: the minimum is c; Rosenbrock: (1, ..., 1)
    c = np.array([0.3, -1.2, 2.0])
    r = O.nelmin_test(0, [1.0, 1.0, 1.0], [0.5, 0.5, 0.5], [1, 2, 3], c, reqmin=1e-16, konvge=10,
                      kcount=20000)
    assert r["ifault"] == 0
    assert np.allclose(r["xmin"], c, atol=1e-5)
    r = O.nelmin_test(1, [-1.2, 1.0], [0.5, 0.5], reqmin=1e-16, konvge=10, kcount=20000)
    assert np.allclose(r["xmin"], [1.0, 1.0], atol=1e-3)
    # kcount exhausted -> ifault 2, icount > kcount
    r = O.nelmin_test(1, [-1.2, 1.0, 0.5, 0.3], [0.5] * 4, reqmin=1e-30, konvge=10, kcount=50)
    assert r["ifault"] == 2 and r["icount"] > 50


@needs_ref
def test_nelmin_restart_path_is_exercised_and_identical():
    # Rosenbrock from huge simplex steps with a loose reqmin stops short of the minimum often
    # enough that the (long)(10 f) local-minimum test fails and the restart branch of
    # asa047.c:488-493 runs
    rng = np.random.default_rng(11)
    restarts = 0
    for trial in range(60):
        n = 3
        start, step = rng.normal(0, 3, n), np.full(n, 1.0e4)
        for reqmin in (1e-1, 1e-2):
            a = O.nelmin_test(1, start, step, None, None, reqmin, 5, 3000, use_ref=False)
            b = O.nelmin_test(1, start, step, None, None, reqmin, 5, 3000, use_ref=True)
            assert _same(a, b)
            restarts += a["numres"]
    assert restarts > 5


# --- the full per-pixel inversion with the reference's nelmin vs the restatement ---
@needs_ref
@pytest.mark.parametrize("mode,nbot", [(1, 2), (0, 1), (0, 2), (0, 3)])
def test_pixel_inversion_reference_nelmin(mode, nbot):
    rrs, ss, truth, nodata, _ = O.gen(1, 96)
    kw = dict(mode=mode, n_bottoms=nbot, reqmin=1e-10, konvge=100, kcount=2000)
    a = O.invert(rrs, ss, **kw)
    b = O.invert(rrs, ss, use_ref_nm=True, **kw)
    for k in ("x", "depth", "rrs_error", "model_error", "n_iter", "converged", "n_starts",
              "status", "k_min", "iod", "b_pct", "bottom_type"):
        assert np.array_equal(a[k], b[k]), k
    assert (a["status"] == -1).sum() == nodata.sum()


# --- the scene-scale no-data discs: geometry from the KAT-pinned Philox ---
def _uniforms(seed, ctr0, block):
    c = np.array([ctr0 & 0xFFFFFFFF, ctr0 >> 32, block, 0], np.uint32)
    k = np.array([seed & 0xFFFFFFFF, seed >> 32], np.uint32)
    o = np.zeros(4, np.uint32)
    O.lib().orc_py_philox(c.ctypes.data, k.ctypes.data, o.ctypes.data)
    a = ((int(o[0]) << 32) | int(o[1])) >> 11
    b = ((int(o[2]) << 32) | int(o[3])) >> 11
    return a * 2.0 ** -53, b * 2.0 ** -53


def test_nodata_discs_geometry():
    """Every pixel of a small scene: masked exactly when its centre is strictly inside one of
    the discs whose centres and radii come from the Philox draws (blocks 250 and 251 of counter
    = disc index), and the masked fraction near 1 - exp(-n pi E[r^2] / area)."""
    seed, rows, cols, nd, r0, r1 = 9, 180, 240, 30, 6.0, 14.0
    yy, xx = np.mgrid[0:rows, 0:cols]
    yc, xc = yy + 0.5, xx + 0.5
    expect = np.zeros((rows, cols), bool)
    for i in range(nd):
        u0, u1 = _uniforms(seed, i, 250)
        u2, _ = _uniforms(seed, i, 251)
        cx, cy, r = u0 * cols, u1 * rows, r0 + (r1 - r0) * u2
        expect |= (xc - cx) ** 2 + (yc - cy) ** 2 < r * r
    got = O.disc_mask(seed, rows, cols, nd, r0, r1, np.arange(rows * cols)).reshape(rows, cols)
    assert np.array_equal(got.astype(bool), expect)
    mean_r2 = (r0 * r0 + r0 * r1 + r1 * r1) / 3.0
    frac = 1.0 - np.exp(-nd * np.pi * mean_r2 / (rows * cols))
    assert abs(got.mean() - frac) < 0.06
