"""ctypes bindings of the oracle (oracle/liboracle.so) and of the reference's own compiled
nelmin / interp_1d / approx_equal (oracle/_ref/libphotic_ref.so) -- test infrastructure."""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))  # repo root
ORACLE = os.path.join(ROOT, "oracle", "liboracle.so")
REF = os.path.join(ROOT, "oracle", "_ref", "libphotic_ref.so")
HOSTSIM = os.path.join(ROOT, "tests", "hostsim", "libph_hostsim.so")

WL4 = (443.0, 482.0, 561.0, 655.0)


def _p(a):
    return None if a is None else a.ctypes.data


class PyOut(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("x", "depth", "rrs_error", "model_error", "k_min", "iod",
                                          "albedo", "bottom_type", "b_pct", "n_iter",
                                          "converged", "n_evals", "n_starts", "status",
                                          "misfit")]


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE):
            raise RuntimeError("oracle not built: make -C oracle")
        _lib = C.CDLL(ORACLE)
        v, i, l, d = C.c_void_p, C.c_int, C.c_long, C.c_double
        _lib.orc_py_invert.restype = l
        _lib.orc_py_invert.argtypes = [i, i, v, v, i, l, v, v, d, d, i, i, i, v, d, C.c_float, v, v,
                                       v, v, C.POINTER(PyOut)]
        _lib.orc_py_forward.argtypes = [v, i, l, v, v, v, v]
        _lib.orc_py_objective.argtypes = [v, i, l, v, v, v, v, v]
        _lib.orc_py_gen.argtypes = [C.c_uint64, l, l, v, i, d, d, v, v, v, v, v]
        _lib.orc_py_disc_mask.argtypes = [C.c_uint64, l, l, i, d, d, v, l, v]
        _lib.orc_py_band_tables.argtypes = [v, i, v, v]
        _lib.orc_py_start_step.argtypes = [i, i, v, i, v, d, v, v]
        _lib.orc_py_objective_stack.argtypes = [v, i, i, v, v, v, v]
        _lib.orc_py_objective_stack_ex.restype = d
        _lib.orc_py_objective_stack_ex.argtypes = [v, i, i, i, i, v, v, v, v]
        _lib.orc_py_start_step_stack.argtypes = [i, i, v, i, i, v, d, v, v]
        _lib.orc_py_objective_stack.restype = d
        _lib.orc_py_gen_stack.argtypes = [C.c_uint64, l, l, v, i, d, i, d, i, v, v, v, v]
        _lib.orc_py_invert_stack.restype = l
        _lib.orc_py_invert_stack_ex.argtypes = [v, i, i, l, v, v, v, v, i, i, v, d, i, i, i, v,
                                                v, C.c_float, v, v, v, C.POINTER(PyOut)]
        _lib.orc_py_invert_stack.argtypes = [v, i, i, l, v, v, v, d, i, i, i, v, C.c_float, v,
                                             v, v, C.POINTER(PyOut)]
        _lib.orc_py_start_step.restype = i
        _lib.orc_py_Y.argtypes = [d]
        _lib.orc_py_Y.restype = d
        _lib.orc_py_nelmin_test.argtypes = [v, i, i, v, v, v, v, d, i, i, v, v, v, v, v, v]
        _lib.orc_py_interp.restype = d
        _lib.orc_py_interp.argtypes = [v, v, i, d]
        _lib.orc_py_approx_equal.restype = i
        _lib.orc_py_approx_equal.argtypes = [C.c_float, C.c_float, C.c_float]
        _lib.orc_py_philox.argtypes = [v, v, v]
        _lib.orc_py_mlp_forward.argtypes = [v, i, v, v]
        _lib.orc_py_mlp_loss_grad.restype = d
        _lib.orc_py_mlp_loss_grad.argtypes = [v, i, v, v, v]
        _lib.orc_py_mlp_nparam.restype = i
        _lib.orc_py_set_threads.restype = i
        _lib.orc_py_set_threads.argtypes = [i]
        _lib.orc_py_invert_regions.restype = l
        _lib.orc_py_invert_regions.argtypes = [v, i, i, i, i, v, v, v, v, i, i, v, d, i, i, i, v,
                                               v, l, v, v, i, v, v, C.POINTER(PyOut)]
        _lib.orc_py_region_objective_bands.restype = d
        _lib.orc_py_region_objective_bands.argtypes = [v, v, i, i, i, v, v, v, i, v, v, v]
        _lib.orc_py_invert_regions_bands.restype = l
        _lib.orc_py_invert_regions_bands.argtypes = [v, v, i, i, i, v, v, v, v, i, i, v, d, i,
                                                     i, i, v, v, l, v, v, i, v, v,
                                                     C.POINTER(PyOut)]
        _lib.orc_py_eval.argtypes = [i, i, v, v, i, i, v, v, v, v, v, v]
        _lib.orc_py_region_eval.argtypes = [v, i, i, i, i, v, v, v, i, v, v, v]
        _lib.orc_py_invert_ordered.restype = l
        _lib.orc_py_invert_ordered.argtypes = [i, i, v, v, i, l, v, v, d, d, i, i, C.c_float,
                                               i, i, i, d, d, d, v, v, v, v, v, v, v, v, v]
        _lib.orc_py_spectral_diff_mean.restype = C.c_float
        _lib.orc_py_spectral_diff_mean.argtypes = [v, v]
        _lib.orc_py_order_key.restype = C.c_float
        _lib.orc_py_order_key.argtypes = [v, v]
        _lib.orc_sec_deg.restype = d
        _lib.orc_sec_deg.argtypes = [d]
        _lib.orc_py_region_objective.restype = d
        _lib.orc_py_region_objective.argtypes = [v, i, i, i, i, v, v, v, i, v, v, v]
        _lib.orc_py_region_start_step.restype = i
        _lib.orc_py_region_start_step.argtypes = [v, i, i, i, v, i, d, v, v]
        _lib.orc_py_region_optimise.restype = i
        _lib.orc_py_region_optimise.argtypes = [v, i, i, i, i, v, v, v, i, v, d, i, i, i, v, v, v]
        _lib.orc_py_smooth_grid.argtypes = [v, i, i, i, C.c_float, v]
        _lib.orc_py_depth_sigma.restype = i
        _lib.orc_py_depth_sigma.argtypes = [v, i, i, i, i, v, v, v, v, v, v, i, i, v, C.c_float,
                                            d, d, i, i, v, v]
        _lib.orc_py_gen_field.argtypes = [C.c_uint64, l, l, d, v, i, d, v, v, v]
        _lib.orc_py_region_gather.restype = i
        _lib.orc_py_region_gather.argtypes = [v, v, i, i, i, i, i, i, i, v, v]
    return _lib


def set_threads(n):
    return lib().orc_py_set_threads(n)


def ref():
    """The reference's compiled asa047.c + common.c, or None when not built."""
    global _ref
    if _ref is None and os.path.exists(REF):
        _ref = C.CDLL(REF)
        _ref.interp_1d.restype = C.c_double
        _ref.interp_1d.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_double]
        _ref.approx_equal.restype = C.c_int
        _ref.approx_equal.argtypes = [C.c_float, C.c_float, C.c_float]
        _ref.smoothed_index_2d.restype = C.c_float
        _ref.smoothed_index_2d.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int,
                                           C.c_int, C.c_float]
    return _ref


def ref_nelmin_ptr():
    r = ref()
    if r is None:
        return None
    return C.cast(r.nelmin, C.c_void_p).value


def invert(rrs, sec_sun=None, mode=1, n_bottoms=2, bottom_idx=(0, 1, 2), wl=WL4, reqmin=1e-10,
           konvge=100, kcount=2000, hstart=(40.0, 30.0, 20.0, 15.0, 10.0, 7.5, 2.5, 1.0),
           early_exit=0.0, nodata=-9999.0, use_ref_nm=False, seed=None,
           seed_rel=(0.3, 0.0, 1.0, 1.0, 1.0), seed_min=(0.5, 0.25, 0.1, 0.1, 0.02),
           sec_sun0=1.0):
    """Oracle inversion of single-date pixels; rrs [nb][n] float32."""
    L = lib()
    rrs = np.ascontiguousarray(rrs, np.float32)
    nb, n = rrs.shape
    npar = 5 if mode == 1 else 1 + 2 * n_bottoms + 3
    out = {
        "x": np.zeros((npar, n)), "depth": np.zeros(n, np.float32),
        "rrs_error": np.zeros(n, np.float32), "model_error": np.zeros(n, np.float32),
        "k_min": np.zeros(n, np.float32), "iod": np.zeros(n, np.float32),
        "albedo": np.zeros(n, np.float32), "bottom_type": np.zeros(n, np.float32),
        "b_pct": np.zeros((3, n), np.float32), "n_iter": np.zeros(n, np.int32),
        "converged": np.zeros(n, np.int32), "n_evals": np.zeros(n, np.int32),
        "n_starts": np.zeros(n, np.int32), "status": np.zeros(n, np.int32),
        "misfit": np.zeros(n),
    }
    o = PyOut(**{k: v.ctypes.data for k, v in out.items()})
    wl = np.asarray(wl, np.float64)
    hs = np.asarray(hstart, np.float64)
    bi = np.asarray(bottom_idx, np.int32)
    ss = None if sec_sun is None else np.ascontiguousarray(sec_sun, np.float32)
    sd = None if seed is None else np.ascontiguousarray(seed, np.float32)
    sr = np.asarray(seed_rel, np.float64)
    sm = np.asarray(seed_min, np.float64)
    nm = ref_nelmin_ptr() if use_ref_nm else None
    if use_ref_nm and nm is None:
        raise RuntimeError("oracle/_ref not built")
    L.orc_py_invert(mode, n_bottoms, bi.ctypes.data, wl.ctypes.data, nb, n, rrs.ctypes.data,
                    _p(ss), sec_sun0, reqmin, konvge, kcount, len(hs), hs.ctypes.data, early_exit,
                    nodata, nm, _p(sd), sr.ctypes.data, sm.ctypes.data, C.byref(o))
    return out


def forward(params, sec_sun, Y, wl=WL4):
    params = np.ascontiguousarray(params, np.float64)
    n = params.shape[0]
    out = np.zeros((n, len(wl)))
    w = np.asarray(wl, np.float64)
    lib().orc_py_forward(w.ctypes.data, len(wl), n, params.ctypes.data,
                         np.ascontiguousarray(sec_sun, np.float64).ctypes.data,
                         np.ascontiguousarray(Y, np.float64).ctypes.data, out.ctypes.data)
    return out


def objective(x, rrs, sec_sun, wl=WL4):
    x = np.ascontiguousarray(x, np.float64)
    n = x.shape[0]
    out, err = np.zeros(n), np.zeros(n)
    w = np.asarray(wl, np.float64)
    r = np.ascontiguousarray(rrs, np.float32)
    s = np.ascontiguousarray(sec_sun, np.float32)
    lib().orc_py_objective(w.ctypes.data, len(wl), n, x.ctypes.data, r.ctypes.data, s.ctypes.data,
                           out.ctypes.data, err.ctypes.data)
    return out, err


def gen(seed, n, pix0=0, sigma=2e-4, deep_frac=0.0, wl=WL4):
    nb = len(wl)
    rrs = np.zeros((nb, n), np.float32)
    ss = np.zeros(n, np.float32)
    truth = np.zeros((6, n))
    nodata = np.zeros(n, np.int32)
    Y = np.zeros(n)
    w = np.asarray(wl, np.float64)
    lib().orc_py_gen(seed, pix0, n, w.ctypes.data, nb, sigma, deep_frac, rrs.ctypes.data,
                     ss.ctypes.data, truth.ctypes.data, nodata.ctypes.data, Y.ctypes.data)
    return rrs, ss, truth, nodata, Y


def disc_mask(seed, n_rows, n_cols, n_discs, r_min, r_max, idx):
    """The no-data disc mask (uint8) at pixel indices idx of an n_rows x n_cols scene."""
    idx = np.ascontiguousarray(idx, np.int64)
    mask = np.zeros(idx.shape[0], np.uint8)
    lib().orc_py_disc_mask(seed, n_rows, n_cols, n_discs, r_min, r_max, idx.ctypes.data,
                           idx.shape[0], mask.ctypes.data)
    return mask


def band_tables(wl=WL4):
    nb = len(wl)
    out = np.zeros((12, nb))
    aw640 = C.c_double()
    w = np.asarray(wl, np.float64)
    lib().orc_py_band_tables(w.ctypes.data, nb, out.ctypes.data, C.byref(aw640))
    return out, aw640.value


def nelmin_test(kind, start, step, w=None, c=None, reqmin=1e-10, konvge=10, kcount=5000,
                use_ref=False):
    n = len(start)
    st = np.array(start, np.float64)
    sp = np.array(step, np.float64)
    xmin = np.zeros(n)
    yn = C.c_double()
    ic, nr, ifa = C.c_int(), C.c_int(), C.c_int()
    ev = C.c_long()
    wv = None if w is None else np.asarray(w, np.float64)
    cv = None if c is None else np.asarray(c, np.float64)
    nm = ref_nelmin_ptr() if use_ref else None
    lib().orc_py_nelmin_test(nm, kind, n, _p(wv), _p(cv), st.ctypes.data, sp.ctypes.data, reqmin,
                             konvge, kcount, xmin.ctypes.data, C.byref(yn), C.byref(ic),
                             C.byref(nr), C.byref(ifa), C.byref(ev))
    return dict(xmin=xmin, ynewlo=yn.value, icount=ic.value, numres=nr.value, ifault=ifa.value,
                evals=ev.value, start_after=st)


def hostsim_invert(cfg_const, n_bottoms, rrs, sec_sun=None, seed=None, math="libm"):
    """The product's lane engine on the host (tests/hostsim).  math="libm": exp/log/pow from
    the host C library, as in the oracle; math="cr": the correctly rounded functions the GPU
    kernels use."""
    path = HOSTSIM if math == "libm" else HOSTSIM.replace("libph_hostsim", "libph_hostsim_cr")
    if not os.path.exists(path):
        raise RuntimeError("host build of the engine missing: tests/hostsim/build.sh")
    h = C.CDLL(path)
    rrs = np.ascontiguousarray(rrs, np.float32)
    nb, n = rrs.shape
    npar = 5 if n_bottoms == 0 else 1 + 2 * n_bottoms + 3
    out = {"x": np.zeros((npar, n)), "depth": np.zeros(n, np.float32),
           "rrs_error": np.zeros(n, np.float32), "model_error": np.zeros(n, np.float32),
           "status": np.zeros(n, np.int32), "n_evals": np.zeros(n, np.int32),
           "n_iter": np.zeros(n, np.int32), "n_starts": np.zeros(n, np.int32)}
    ss = None if sec_sun is None else np.ascontiguousarray(sec_sun, np.float32)
    sd = None if seed is None else np.ascontiguousarray(seed, np.float32)
    h.phs_invert.argtypes = [C.c_void_p, C.c_int, C.c_long] + [C.c_void_p] * 11
    rc = h.phs_invert(C.cast(cfg_const, C.c_void_p), n_bottoms, n, rrs.ctypes.data, _p(ss), _p(sd),
                      *[out[k].ctypes.data for k in ("x", "depth", "rrs_error", "model_error",
                                                      "status", "n_evals", "n_iter", "n_starts")])
    assert rc == 0
    return out


def Y_of_chi(chi):
    return lib().orc_py_Y(chi)


def start_step(rrs, h0, mode=1, n_bottoms=2, wl=WL4):
    """Start point and steps of one depth start for a single-date pixel (rrs [nb])."""
    r = np.ascontiguousarray(rrs, np.float64)
    w = np.asarray(wl, np.float64)
    start, step = np.zeros(64), np.zeros(64)
    n = lib().orc_py_start_step(mode, n_bottoms, w.ctypes.data, len(w), r.ctypes.data, h0,
                                start.ctypes.data, step.ctypes.data)
    return start[:n], step[:n]


def table(name):
    """One of the oracle's 41-node tables: ref_wl, aw, bbw, a0, a1, or bottom (8 x 41)."""
    if name == "bottom":
        return np.ctypeslib.as_array((C.c_double * (8 * 41)).in_dll(lib(), "orc_bottom")).reshape(8, 41).copy()
    return np.ctypeslib.as_array((C.c_double * 41).in_dll(lib(), "orc_" + name)).copy()


def gen_stack(seed, n, n_dates=8, pix0=0, sigma=2e-4, p_cloud=0.3, k_min=3, wl=WL4):
    """Multi-date stack scene: rrs [D][nb][n] float32, sec_sun [D][n] float32, valid [D][n]
    uint8, truth [2 + 4D][n] (H, B, then P, G, X, theta_sun per date)."""
    nb = len(wl)
    rrs = np.zeros((n_dates, nb, n), np.float32)
    ss = np.zeros((n_dates, n), np.float32)
    valid = np.zeros((n_dates, n), np.uint8)
    truth = np.zeros((2 + 4 * n_dates, n))
    w = np.asarray(wl, np.float64)
    lib().orc_py_gen_stack(seed, pix0, n, w.ctypes.data, nb, sigma, n_dates, p_cloud, k_min,
                           rrs.ctypes.data, ss.ctypes.data, valid.ctypes.data, truth.ctypes.data)
    return rrs, ss, valid, truth


def invert_stack(rrs, sec_sun, valid, reqmin=1e-10, konvge=100, kcount=2000,
                 hstart=(40.0, 30.0, 20.0, 15.0, 10.0, 7.5, 2.5, 1.0), nodata=-9999.0,
                 seed=None, seed_rel=(0.3, 0.0, 1.0, 1.0, 1.0),
                 seed_min=(0.5, 0.25, 0.1, 0.1, 0.02), wl=WL4):
    """Oracle inversion of multi-date stacks (mix layout, valid dates only)."""
    rrs = np.ascontiguousarray(rrs, np.float32)
    nd, nb, n = rrs.shape
    out = {
        "x": np.zeros((2 + 3 * nd, n)), "depth": np.zeros(n, np.float32),
        "rrs_error": np.zeros(n, np.float32), "model_error": np.zeros(n, np.float32),
        "k_min": np.zeros(n, np.float32), "iod": np.zeros(n, np.float32),
        "albedo": np.zeros(n, np.float32), "bottom_type": np.zeros(n, np.float32),
        "b_pct": np.zeros((3, n), np.float32), "n_iter": np.zeros(n, np.int32),
        "converged": np.zeros(n, np.int32), "n_evals": np.zeros(n, np.int32),
        "n_starts": np.zeros(n, np.int32), "status": np.zeros(n, np.int32),
        "misfit": np.zeros(n),
    }
    o = PyOut(**{k: v.ctypes.data for k, v in out.items()})
    w = np.asarray(wl, np.float64)
    hs = np.asarray(hstart, np.float64)
    ss = np.ascontiguousarray(sec_sun, np.float32)
    va = np.ascontiguousarray(valid, np.uint8)
    sd = None if seed is None else np.ascontiguousarray(seed, np.float32)
    sr = np.asarray(seed_rel, np.float64)
    sm = np.asarray(seed_min, np.float64)
    lib().orc_py_invert_stack(w.ctypes.data, nb, nd, n, rrs.ctypes.data, ss.ctypes.data,
                              va.ctypes.data, reqmin, konvge, kcount, len(hs), hs.ctypes.data,
                              nodata, _p(sd), sr.ctypes.data, sm.ctypes.data, C.byref(o))
    return out


def invert_stack_native(rrs, sec_sun, valid, n_bottoms=2, bottom_idx=(0, 1, 2), sec_view=None,
                        hstart_px=None, reqmin=1e-2, konvge=100, kcount=5000,
                        hstart=(40.0, 30.0, 20.0, 15.0, 10.0, 7.5, 2.5, 1.0), nodata=-9999.0,
                        wl=WL4):
    """Oracle inversion of multi-date stacks in the reference's native layout
    [H, B_k, q_k, (P, G, X) per date]; sec_view [D][n] or None; hstart_px [n] or None."""
    rrs = np.ascontiguousarray(rrs, np.float32)
    nd, nb, n = rrs.shape
    npar = 1 + 2 * n_bottoms + 3 * nd
    out = {
        "x": np.zeros((npar, n)), "depth": np.zeros(n, np.float32),
        "rrs_error": np.zeros(n, np.float32), "model_error": np.zeros(n, np.float32),
        "k_min": np.zeros(n, np.float32), "iod": np.zeros(n, np.float32),
        "albedo": np.zeros(n, np.float32), "bottom_type": np.zeros(n, np.float32),
        "b_pct": np.zeros((3, n), np.float32), "n_iter": np.zeros(n, np.int32),
        "converged": np.zeros(n, np.int32), "n_evals": np.zeros(n, np.int32),
        "n_starts": np.zeros(n, np.int32), "status": np.zeros(n, np.int32),
        "misfit": np.zeros(n),
    }
    o = PyOut(**{k: v.ctypes.data for k, v in out.items()})
    w = np.asarray(wl, np.float64)
    hs = np.asarray(hstart, np.float64)
    bi = np.asarray(bottom_idx, np.int32)
    ss = np.ascontiguousarray(sec_sun, np.float32)
    sv = None if sec_view is None else np.ascontiguousarray(sec_view, np.float32)
    hp = None if hstart_px is None else np.ascontiguousarray(hstart_px, np.float32)
    va = np.ascontiguousarray(valid, np.uint8)
    sr = np.zeros(5)
    lib().orc_py_invert_stack_ex(w.ctypes.data, nb, nd, n, rrs.ctypes.data, ss.ctypes.data,
                                 _p(sv), va.ctypes.data, 0, n_bottoms, bi.ctypes.data, reqmin,
                                 konvge, kcount, len(hs), hs.ctypes.data, _p(hp), nodata, None,
                                 sr.ctypes.data, sr.ctypes.data, C.byref(o))
    return out


def hostsim_invert_stack(cfg_const, rrs, sec_sun, valid, seed=None, math="libm"):
    """The product's stack engine on the host (tests/hostsim); cfg_const from a single-date
    configuration (the engine sets the date count, reqmin * n and the early exit per pixel)."""
    path = HOSTSIM if math == "libm" else HOSTSIM.replace("libph_hostsim", "libph_hostsim_cr")
    h = C.CDLL(path)
    rrs = np.ascontiguousarray(rrs, np.float32)
    nd, nb, n = rrs.shape
    out = {"x": np.zeros((2 + 3 * nd, n)), "depth": np.zeros(n, np.float32),
           "rrs_error": np.zeros(n, np.float32), "model_error": np.zeros(n, np.float32),
           "status": np.zeros(n, np.int32), "n_evals": np.zeros(n, np.int32),
           "n_iter": np.zeros(n, np.int32), "n_starts": np.zeros(n, np.int32)}
    ss = np.ascontiguousarray(sec_sun, np.float32)
    va = np.ascontiguousarray(valid, np.uint8)
    sd = None if seed is None else np.ascontiguousarray(seed, np.float32)
    h.phs_invert_stack.argtypes = [C.c_void_p, C.c_int, C.c_long] + [C.c_void_p] * 12
    rc = h.phs_invert_stack(C.cast(cfg_const, C.c_void_p), nd, n, rrs.ctypes.data, ss.ctypes.data,
                            va.ctypes.data, _p(sd),
                            *[out[k].ctypes.data for k in ("x", "depth", "rrs_error",
                                                            "model_error", "status", "n_evals",
                                                            "n_iter", "n_starts")])
    assert rc == 0
    return out


def objective_stack(x, rrs, sec_sun, wl=WL4, n_bottoms=None):
    """Objective and E_Rrs of one vector x over k dates (rrs [k][nb]): the mix layout
    [2 + 3k], or the native one [1 + 2 n_bottoms + 3k] when n_bottoms is given."""
    x = np.ascontiguousarray(x, np.float64)
    r = np.ascontiguousarray(rrs, np.float32)
    s = np.ascontiguousarray(sec_sun, np.float32)
    k, nb = r.shape
    e = C.c_double()
    w = np.asarray(wl, np.float64)
    if n_bottoms is None:
        f = lib().orc_py_objective_stack(w.ctypes.data, nb, k, x.ctypes.data, r.ctypes.data,
                                         s.ctypes.data, C.byref(e))
    else:
        f = lib().orc_py_objective_stack_ex(w.ctypes.data, nb, k, 0, n_bottoms, x.ctypes.data,
                                            r.ctypes.data, s.ctypes.data, C.byref(e))
    return f, e.value


def start_step_stack(rrs, h0, mode=1, n_bottoms=2, wl=WL4):
    """Start point and steps of one depth start for a k-date pixel (rrs [k][nb])."""
    r = np.ascontiguousarray(rrs, np.float64)
    k, nb = r.shape
    w = np.asarray(wl, np.float64)
    start, step = np.zeros(64), np.zeros(64)
    n = lib().orc_py_start_step_stack(mode, n_bottoms, w.ctypes.data, nb, k, r.ctypes.data, h0,
                                      start.ctypes.data, step.ctypes.data)
    return start[:n], step[:n]


HSTART8 = (40.0, 30.0, 20.0, 15.0, 10.0, 7.5, 2.5, 1.0)


def _region_out(npar, n):
    return {
        "x": np.zeros((npar, n)), "depth": np.zeros(n, np.float32),
        "rrs_error": np.zeros(n, np.float32), "model_error": np.zeros(n, np.float32),
        "k_min": np.zeros(n, np.float32), "iod": np.zeros(n, np.float32),
        "albedo": np.zeros(n, np.float32), "bottom_type": np.zeros(n, np.float32),
        "b_pct": np.zeros((3, n), np.float32), "n_iter": np.zeros(n, np.int32),
        "converged": np.zeros(n, np.int32), "n_evals": np.zeros(n, np.int32),
        "n_starts": np.zeros(n, np.int32), "status": np.zeros(n, np.int32),
        "misfit": np.zeros(n),
    }


def invert_regions(rrs, n_rows, n_cols, idx, nodata, sec_sun, sec_view, n_spatial=2,
                   n_bottoms=3, bottom_idx=(0, 1, 2), hstart_px=None, reqmin=1e-2, konvge=100,
                   kcount=5000, hstart=HSTART8, screen=False, screen_rrs=None, wl=WL4):
    """Oracle inversion of the pixels idx (row * n_cols + col) of a scene rrs [nd][nb][n_rows *
    n_cols] with their (2 n_spatial - 1)^2 regions; nodata [nd][nb]; sec_sun, sec_view [nd];
    screen: apply the reference's own-pixel no-data rule first, to rrs or to screen_rrs (the
    grids before smoothing) when given.  x is [R_max (1 + 2 n_bottoms) + 3 nd][n] in the
    reference's layout, zero-padded."""
    rrs = np.ascontiguousarray(rrs, np.float32)
    nd, nb, npix = rrs.shape
    assert npix == n_rows * n_cols
    ix = np.ascontiguousarray(idx, np.int64)
    n = len(ix)
    rmax = (2 * max(n_spatial, 1) - 1) ** 2
    npar = rmax * (1 + 2 * n_bottoms) + 3 * nd
    out = _region_out(npar, n)
    out["n_regions"] = np.zeros(n, np.int32)
    out["origin"] = np.zeros(n, np.int32)
    o = PyOut(**{k: v.ctypes.data for k, v in out.items() if k not in ("n_regions", "origin")})
    w = np.asarray(wl, np.float64)
    hs = np.asarray(hstart, np.float64)
    bi = np.asarray(bottom_idx, np.int32)
    nod = np.ascontiguousarray(np.broadcast_to(np.asarray(nodata, np.float32), (nd, nb)))
    ss = np.ascontiguousarray(np.broadcast_to(np.asarray(sec_sun, np.float64), (nd,)))
    sv = np.ascontiguousarray(np.broadcast_to(np.asarray(sec_view, np.float64), (nd,)))
    hp = None if hstart_px is None else np.ascontiguousarray(hstart_px, np.float32)
    sc = None
    if screen_rrs is not None:
        sc = np.ascontiguousarray(screen_rrs, np.float32).reshape(rrs.shape)
    elif screen:
        sc = rrs
    lib().orc_py_invert_regions(w.ctypes.data, nb, nd, n_rows, n_cols, rrs.ctypes.data,
                                nod.ctypes.data, ss.ctypes.data, sv.ctypes.data, n_spatial,
                                n_bottoms, bi.ctypes.data, reqmin, konvge, kcount, len(hs),
                                hs.ctypes.data, _p(hp), n, ix.ctypes.data, _p(sc), npar,
                                out["n_regions"].ctypes.data, out["origin"].ctypes.data,
                                C.byref(o))
    return out


def invert_regions_bands(rrs_rows, wls, n_rows, n_cols, idx, nodata, sec_sun, sec_view,
                         n_spatial=2, n_bottoms=3, bottom_idx=(0, 1, 2), hstart_px=None,
                         reqmin=1e-2, konvge=100, kcount=5000, hstart=HSTART8, screen=False,
                         screen_rrs=None):
    """invert_regions with a band set per scene: wls[s] the wavelengths of scene s; rrs_rows
    [sum len(wls[s])][n_rows * n_cols] the scenes' band grids one after another, nodata one
    value per band grid (or a scalar)."""
    rrs = np.ascontiguousarray(rrs_rows, np.float32)
    nbs = np.asarray([len(w) for w in wls], np.int32)
    nd = len(nbs)
    rows = int(nbs.sum())
    assert rrs.shape == (rows, n_rows * n_cols)
    w = np.ascontiguousarray(np.concatenate([np.asarray(x, np.float64) for x in wls]))
    ix = np.ascontiguousarray(idx, np.int64)
    n = len(ix)
    rmax = (2 * max(n_spatial, 1) - 1) ** 2
    npar = rmax * (1 + 2 * n_bottoms) + 3 * nd
    out = _region_out(npar, n)
    out["n_regions"] = np.zeros(n, np.int32)
    out["origin"] = np.zeros(n, np.int32)
    o = PyOut(**{k: v.ctypes.data for k, v in out.items() if k not in ("n_regions", "origin")})
    hs = np.asarray(hstart, np.float64)
    bi = np.asarray(bottom_idx, np.int32)
    nod = np.ascontiguousarray(np.broadcast_to(np.asarray(nodata, np.float32), (rows,)))
    ss = np.ascontiguousarray(np.broadcast_to(np.asarray(sec_sun, np.float64), (nd,)))
    sv = np.ascontiguousarray(np.broadcast_to(np.asarray(sec_view, np.float64), (nd,)))
    hp = None if hstart_px is None else np.ascontiguousarray(hstart_px, np.float32)
    sc = None
    if screen_rrs is not None:
        sc = np.ascontiguousarray(screen_rrs, np.float32).reshape(rrs.shape)
    elif screen:
        sc = rrs
    lib().orc_py_invert_regions_bands(w.ctypes.data, nbs.ctypes.data, nd, n_rows, n_cols,
                                      rrs.ctypes.data, nod.ctypes.data, ss.ctypes.data,
                                      sv.ctypes.data, n_spatial, n_bottoms, bi.ctypes.data,
                                      reqmin, konvge, kcount, len(hs), hs.ctypes.data, _p(hp), n,
                                      ix.ctypes.data, _p(sc), npar, out["n_regions"].ctypes.data,
                                      out["origin"].ctypes.data, C.byref(o))
    return out


def region_objective(x, rrs, origin=0, sec_sun=1.0, sec_view=1.0, n_bottoms=3,
                     bottom_idx=(0, 1, 2), wl=WL4):
    """Region objective of one vector x for region data rrs [R][nd][nb]; returns (E, terms)
    with terms = E_Rrs, depth penalty, bottom penalty, K penalty, albedo, model error."""
    r = np.ascontiguousarray(rrs, np.float64)
    R, nd, nb = r.shape
    xv = np.ascontiguousarray(x, np.float64)
    w = np.asarray(wl, np.float64)
    bi = np.asarray(bottom_idx, np.int32)
    ss = np.ascontiguousarray(np.broadcast_to(np.asarray(sec_sun, np.float64), (nd,)))
    sv = np.ascontiguousarray(np.broadcast_to(np.asarray(sec_view, np.float64), (nd,)))
    terms = np.zeros(6)
    f = lib().orc_py_region_objective(w.ctypes.data, nb, nd, R, origin, r.ctypes.data,
                                      ss.ctypes.data, sv.ctypes.data, n_bottoms, bi.ctypes.data,
                                      xv.ctypes.data, terms.ctypes.data)
    return f, terms


def region_objective_bands(x, rrs_rows, wls, origin=0, sec_sun=1.0, sec_view=1.0,
                           n_bottoms=3, bottom_idx=(0, 1, 2)):
    """region_objective with a band set per scene: wls[s] the wavelengths of scene s, rrs_rows
    [R][sum len(wls[s])] each region's values scene by scene; returns (E, terms)."""
    r = np.ascontiguousarray(rrs_rows, np.float64)
    nbs = np.asarray([len(w) for w in wls], np.int32)
    nd = len(nbs)
    R = r.shape[0]
    assert r.shape[1] == int(nbs.sum())
    w = np.ascontiguousarray(np.concatenate([np.asarray(q, np.float64) for q in wls]))
    xv = np.ascontiguousarray(x, np.float64)
    bi = np.asarray(bottom_idx, np.int32)
    ss = np.ascontiguousarray(np.broadcast_to(np.asarray(sec_sun, np.float64), (nd,)))
    sv = np.ascontiguousarray(np.broadcast_to(np.asarray(sec_view, np.float64), (nd,)))
    terms = np.zeros(6)
    f = lib().orc_py_region_objective_bands(w.ctypes.data, nbs.ctypes.data, nd, R, origin,
                                            r.ctypes.data, ss.ctypes.data, sv.ctypes.data,
                                            n_bottoms, bi.ctypes.data, xv.ctypes.data,
                                            terms.ctypes.data)
    return f, terms


def evaluate(x, rrs, sec_sun=1.0, sec_view=1.0, mode=1, n_bottoms=2, bottom_idx=(0, 1, 2),
             wl=WL4):
    """The objective of one vector x over k dates (rrs [k][nb]) and the outputs derived from
    it, as a dict (objective, rrs_error, k_error, depth, k_min, iod, albedo, bottom_type,
    b_pct [8]) plus K [k][nb]."""
    r = np.ascontiguousarray(np.atleast_2d(rrs), np.float32)
    k, nb = r.shape
    xv = np.ascontiguousarray(x, np.float64)
    w = np.asarray(wl, np.float64)
    bi = np.asarray(bottom_idx, np.int32)
    ss = np.ascontiguousarray(np.broadcast_to(np.asarray(sec_sun, np.float64), (k,)))
    sv = np.ascontiguousarray(np.broadcast_to(np.asarray(sec_view, np.float64), (k,)))
    out = np.zeros(16)
    K = np.zeros((k, nb))
    lib().orc_py_eval(mode, n_bottoms, bi.ctypes.data, w.ctypes.data, nb, k, xv.ctypes.data,
                      r.ctypes.data, ss.ctypes.data, sv.ctypes.data, out.ctypes.data,
                      K.ctypes.data)
    names = ("objective", "rrs_error", "k_error", "depth", "k_min", "iod", "albedo",
             "bottom_type")
    d = {n: out[i] for i, n in enumerate(names)}
    d["b_pct"] = out[8:16].copy()
    d["K"] = K
    return d


def region_evaluate(x, rrs, origin=0, sec_sun=1.0, sec_view=1.0, n_bottoms=3,
                    bottom_idx=(0, 1, 2), wl=WL4):
    """The region objective of one vector x for region data rrs [R][nd][nb] and the region
    outputs: (objective, depth, K_min, IOD)."""
    r = np.ascontiguousarray(rrs, np.float64)
    R, nd, nb = r.shape
    xv = np.ascontiguousarray(x, np.float64)
    w = np.asarray(wl, np.float64)
    bi = np.asarray(bottom_idx, np.int32)
    ss = np.ascontiguousarray(np.broadcast_to(np.asarray(sec_sun, np.float64), (nd,)))
    sv = np.ascontiguousarray(np.broadcast_to(np.asarray(sec_view, np.float64), (nd,)))
    out = np.zeros(4)
    lib().orc_py_region_eval(w.ctypes.data, nb, nd, R, origin, r.ctypes.data, ss.ctypes.data,
                             sv.ctypes.data, n_bottoms, bi.ctypes.data, xv.ctypes.data,
                             out.ctypes.data)
    return tuple(out)


def region_start_step(rrs, h0, n_bottoms=3, wl=WL4):
    """Start point and steps of one depth start for region data rrs [R][nd][nb]."""
    r = np.ascontiguousarray(rrs, np.float64)
    R, nd, nb = r.shape
    w = np.asarray(wl, np.float64)
    start, step = np.zeros(256), np.zeros(256)
    n = lib().orc_py_region_start_step(w.ctypes.data, nb, nd, R, r.ctypes.data, n_bottoms, h0,
                                       start.ctypes.data, step.ctypes.data)
    return start[:n], step[:n]


def region_optimise(rrs, origin=0, sec_sun=1.0, sec_view=1.0, n_bottoms=3, bottom_idx=(0, 1, 2),
                    reqmin=1e-2, konvge=100, kcount=5000, hstart=HSTART8, wl=WL4):
    """One pixel's region inversion from region data rrs [R][nd][nb]: (x, outputs) with
    outputs = depth, E_Rrs, model error, K_min, albedo, IOD, bottom type, evaluations."""
    r = np.ascontiguousarray(rrs, np.float64)
    R, nd, nb = r.shape
    w = np.asarray(wl, np.float64)
    bi = np.asarray(bottom_idx, np.int32)
    hs = np.asarray(hstart, np.float64)
    ss = np.ascontiguousarray(np.broadcast_to(np.asarray(sec_sun, np.float64), (nd,)))
    sv = np.ascontiguousarray(np.broadcast_to(np.asarray(sec_view, np.float64), (nd,)))
    x, out = np.zeros(256), np.zeros(8)
    n = lib().orc_py_region_optimise(w.ctypes.data, nb, nd, R, origin, r.ctypes.data,
                                     ss.ctypes.data, sv.ctypes.data, n_bottoms, bi.ctypes.data,
                                     reqmin, konvge, kcount, len(hs), hs.ctypes.data,
                                     x.ctypes.data, out.ctypes.data)
    return x[:n], out


def region_gather(rrs, n_rows, n_cols, i, j, n_spatial, nodata):
    """The regions of pixel (i, j): (rrs_regions [R][nd][nb], origin)."""
    g = np.ascontiguousarray(rrs, np.float32)
    nd, nb, npix = g.shape
    nod = np.ascontiguousarray(np.broadcast_to(np.asarray(nodata, np.float32), (nd, nb)))
    out = np.zeros((25, nd, nb))
    org = C.c_int(0)
    R = lib().orc_py_region_gather(g.ctypes.data, nod.ctypes.data, nd, nb, n_rows, n_cols, i, j,
                                   n_spatial, C.byref(org), out.ctypes.data)
    return out[:R], org.value


def smooth_grid(grid, r, spval):
    """The oracle's smoothing of every pixel of a 2-D float grid."""
    g = np.ascontiguousarray(grid, np.float32)
    out = np.empty_like(g)
    lib().orc_py_smooth_grid(g.ctypes.data, g.shape[0], g.shape[1], r, spval, out.ctypes.data)
    return out


def ref_smoothed(grid, i, j, r, spval):
    """The reference's own smoothed_index_2d on a row-pointer copy of grid."""
    g = np.ascontiguousarray(grid, np.float32)
    rows = (C.c_void_p * g.shape[0])(*[g.ctypes.data + k * g.strides[0] for k in range(g.shape[0])])
    return ref().smoothed_index_2d(rows, i, j, g.shape[0], g.shape[1], r, spval)


def sec_deg(theta_deg):
    """1 / cos of an angle in degrees, as the reference converts scene angles."""
    return lib().orc_sec_deg(float(theta_deg))


def gen_field(seed, n_rows, n_cols, theta=35.0, sigma=2e-4, wl=WL4):
    """The region workload's coherent scene: rrs [nb][n] float32 (negative bands kept as
    drawn), truth [6][n], nodata [n] uint8."""
    nb = len(wl)
    n = n_rows * n_cols
    rrs = np.zeros((nb, n), np.float32)
    truth = np.zeros((6, n))
    nod = np.zeros(n, np.uint8)
    w = np.asarray(wl, np.float64)
    lib().orc_py_gen_field(seed, n_rows, n_cols, theta, w.ctypes.data, nb, sigma, rrs.ctypes.data,
                           truth.ctypes.data, nod.ctypes.data)
    return rrs, truth, nod


def depth_sigma(depth, grid, n_rows, n_cols, nodata, rsig, sec_sun, sec_view, n_spatial=1,
                n_bottoms=3, emp=None, emp_nodata=-99999.0, h_prev=None, reqmin=1e-2, konvge=100,
                kcount=5000, wl=WL4):
    """The oracle's depth-error pass: depth [npix] positive (0 unmodelled), grid [nd][nb][npix]
    smoothed, nodata / rsig [nd][nb], sec_sun / sec_view [nd], emp [npix] or None; h_prev the
    start of every sample after the first without empirical depths (None: the depth of the
    last modelled pixel in row-major order).  Returns (table [n_int], sigma [npix] float32)."""
    g = np.ascontiguousarray(grid, np.float32)
    nd, nb, npix = g.shape
    dp = np.ascontiguousarray(depth, np.float32).reshape(-1)
    nod = np.ascontiguousarray(np.broadcast_to(np.asarray(nodata, np.float32), (nd, nb)))
    rs = np.ascontiguousarray(np.broadcast_to(np.asarray(rsig, np.float64), (nd, nb)))
    ss = np.ascontiguousarray(np.broadcast_to(np.asarray(sec_sun, np.float64), (nd,)))
    sv = np.ascontiguousarray(np.broadcast_to(np.asarray(sec_view, np.float64), (nd,)))
    em = None if emp is None else np.ascontiguousarray(emp, np.float32).reshape(-1)
    table = np.zeros(128)
    sigma = np.zeros(npix, np.float32)
    w = np.asarray(wl, np.float64)
    if h_prev is None:
        modelled = np.nonzero(dp > 0)[0]
        h_prev = float(dp[modelled[-1]]) if len(modelled) else 0.0
    n_int = lib().orc_py_depth_sigma(w.ctypes.data, nb, nd, n_rows, n_cols, dp.ctypes.data,
                                     g.ctypes.data, nod.ctypes.data, rs.ctypes.data,
                                     ss.ctypes.data, sv.ctypes.data, n_spatial, n_bottoms,
                                     _p(em), emp_nodata, h_prev, reqmin, konvge, kcount,
                                     table.ctypes.data, sigma.ctypes.data)
    return table[:n_int], sigma


def invert_ordered(rrs, sec_sun=None, mode=1, n_bottoms=2, bottom_idx=(0, 1, 2), reqmin=1e-10,
                   konvge=100, kcount=2000, nodata=-9999.0, run=64, window=16, use_lut=1,
                   match_max=1.5, entry_max=1.5, restart=0.0, noise=None, sec_sun0=1.0, wl=WL4):
    """The oracle's ordered inversion (ordering, look-up table, hot start) of single-date
    pixels rrs [4][n]; noise [4] R_sigma or None."""
    rrs = np.ascontiguousarray(rrs, np.float32)
    nb, n = rrs.shape
    npar = 5 if mode == 1 else 1 + 2 * n_bottoms + 3
    out = {"x": np.zeros((npar, n)), "depth": np.zeros(n, np.float32),
           "rrs_error": np.zeros(n, np.float32), "model_error": np.zeros(n, np.float32),
           "status": np.zeros(n, np.int32), "n_evals": np.zeros(n, np.int32),
           "n_starts": np.zeros(n, np.int32), "role": np.zeros(n, np.int32)}
    w = np.asarray(wl, np.float64)
    bi = np.asarray(bottom_idx, np.int32)
    ss = None if sec_sun is None else np.ascontiguousarray(sec_sun, np.float32)
    nz = None if noise is None else np.ascontiguousarray(noise, np.float64)
    lib().orc_py_invert_ordered(mode, n_bottoms, bi.ctypes.data, w.ctypes.data, nb, n,
                                rrs.ctypes.data, _p(ss), sec_sun0, reqmin, konvge, kcount,
                                nodata, run, window, use_lut, match_max, entry_max, restart,
                                _p(nz), *[out[k].ctypes.data for k in
                                          ("x", "depth", "rrs_error", "model_error", "status",
                                           "n_evals", "n_starts", "role")])
    return out


def spectral_diff_mean(meas, model):
    m = np.ascontiguousarray(meas, np.float64)
    l = np.ascontiguousarray(model, np.float64)
    return float(lib().orc_py_spectral_diff_mean(m.ctypes.data, l.ctypes.data))


def order_key(r, noise=None):
    rr = np.ascontiguousarray(r, np.float32)
    nz = None if noise is None else np.ascontiguousarray(noise, np.float64)
    return float(lib().orc_py_order_key(rr.ctypes.data, _p(nz)))
