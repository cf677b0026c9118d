"""photic_amd -- MI355X inversion of photic's per-pixel semi-analytical model.

Python mirror of the C ABI in include/photic.h (libphotic.so, built in-tree by
photic_amd/build.py).  The library is the product; this module only marshals numpy arrays to
device buffers and back.  It raises ImportError when libphotic.so is missing and
PhoticError when the library reports an error -- there is no CPU fallback.
"""
import ctypes as C  # noqa: N812
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libphotic.so")

MAX_BANDS = 8
MAX_HSTARTS = 16
MODE_NATIVE, MODE_MIX = 0, 1
FP64, FP32 = 0, 1
OK, EINVAL, EUNSUPPORTED, EDEVICE, ENOMEM = 0, 1, 2, 3, 4


class PhoticError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("photic error %d: %s" % (code, msg))
        self.code = code


class Cfg(C.Structure):
    _fields_ = [
        ("n_bands", C.c_int),
        ("wavelengths", C.c_double * MAX_BANDS),
        ("mode", C.c_int),
        ("n_bottoms", C.c_int),
        ("bottom_idx", C.c_int * 8),
        ("reqmin", C.c_double),
        ("konvge", C.c_int),
        ("kcount", C.c_int),
        ("n_hstarts", C.c_int),
        ("hstart", C.c_double * MAX_HSTARTS),
        ("early_exit", C.c_double),
        ("nodata", C.c_float),
        ("sec_sun", C.c_double),
        ("sec_view", C.c_double),
        ("precision", C.c_int),
        ("seed_rel_step", C.c_double * 5),
        ("seed_min_step", C.c_double * 5),
        ("block", C.c_int),
        ("grid", C.c_int),
        ("chunk", C.c_int),
        ("watchdog_s", C.c_double),
        ("check_watchdog", C.c_int),
        ("seed_abort_misfit", C.c_double),
        ("prescreened", C.c_int),
        ("seed_konvge", C.c_int),
    ]


class Outputs(C.Structure):
    _fields_ = [
        ("depth", C.c_void_p),
        ("rrs_error", C.c_void_p),
        ("model_error", C.c_void_p),
        ("albedo", C.c_void_p),
        ("k_min", C.c_void_p),
        ("iod", C.c_void_p),
        ("b_pct", C.c_void_p),
        ("bottom_type", C.c_void_p),
        ("x", C.c_void_p),
        ("status", C.c_void_p),
        ("n_evals", C.c_void_p),
        ("n_iter", C.c_void_p),
        ("n_starts", C.c_void_p),
        ("rrs_model", C.c_void_p),
    ]


# progress callback of photic_invert_stack_stream: (tiles done, tiles, user)
PROGRESS_FN = C.CFUNCTYPE(None, C.c_long, C.c_long, C.c_void_p)


class OrderCfg(C.Structure):
    _fields_ = [("run", C.c_int), ("window", C.c_int), ("use_lut", C.c_int),
                ("match_max", C.c_double), ("entry_max", C.c_double),
                ("restart_misfit", C.c_double), ("r_sigma", C.c_double * 4)]


class OrderStats(C.Structure):
    _fields_ = [(k, C.c_long) for k in ("n_valid", "n_leaders", "n_lut", "n_hot",
                                        "n_restarted")]


class Tuning(C.Structure):
    _fields_ = [("block", C.c_int), ("blocks_per_cu", C.c_int), ("chunk", C.c_int),
                ("ms", C.c_float), ("n_tried", C.c_int)]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError("libphotic.so is not built: run photic_amd/build.py (or "
                          "__graft_entry__.build())")
    lib = C.CDLL(LIB_PATH)
    vp, i, l, d = C.c_void_p, C.c_int, C.c_long, C.c_double
    sig = {
        "photic_cfg_default": (None, [C.POINTER(Cfg), i, C.POINTER(C.c_double)]),
        "photic_n_params": (i, [C.POINTER(Cfg)]),
        "photic_invert": (i, [C.POINTER(Cfg), l, vp, vp, vp, vp, C.POINTER(Outputs), vp]),
        "photic_gen_scene": (i, [C.POINTER(Cfg), C.c_uint64, l, l, d, d, vp, vp, vp, vp]),
        "photic_gen_nodata_discs": (i, [C.c_uint64, l, l, i, d, d, i, C.c_float, vp, vp]),
        "photic_forward": (i, [C.POINTER(Cfg), l, vp, vp, vp, vp, i, vp]),
        "photic_objective": (i, [C.POINTER(Cfg), l, vp, vp, vp, vp, vp, i, vp]),
        "photic_kernel_info": (i, [C.POINTER(Cfg), C.POINTER(i), C.POINTER(i), C.POINTER(i),
                                   C.POINTER(i)]),
        "photic_device_count": (i, []),
        "photic_device_malloc": (i, [C.POINTER(vp), C.c_size_t]),
        "photic_device_free": (i, [vp]),
        "photic_host_malloc": (i, [C.POINTER(vp), C.c_size_t]),
        "photic_host_free": (i, [vp]),
        "photic_memcpy_htod": (i, [vp, vp, C.c_size_t]),
        "photic_memcpy_dtoh": (i, [vp, vp, C.c_size_t]),
        "photic_memset_zero": (i, [vp, C.c_size_t]),
        "photic_synchronize": (i, []),
        "photic_device_copy": (i, [vp, vp, C.c_size_t, vp]),
        "photic_stream_create": (i, [C.POINTER(vp)]),
        "photic_stream_destroy": (i, [vp]),
        "photic_event_create": (i, [C.POINTER(vp)]),
        "photic_event_destroy": (i, [vp]),
        "photic_event_record": (i, [vp, vp]),
        "photic_event_elapsed_ms": (i, [vp, vp, C.POINTER(C.c_float)]),
        "photic_last_error": (C.c_char_p, []),
        "photic_mlp_create": (i, [C.POINTER(vp), C.c_uint64]),
        "photic_mlp_destroy": (None, [vp]),
        "photic_mlp_n_params": (i, []),
        "photic_mlp_set_weights": (i, [vp, vp]),
        "photic_mlp_get_weights": (i, [vp, vp]),
        "photic_mlp_trainset": (i, [C.POINTER(Cfg), C.c_uint64, l, l, d, vp, vp, vp]),
        "photic_mlp_gradient": (i, [vp, vp, vp, l, vp, C.POINTER(C.c_float)]),
        "photic_mlp_train": (i, [vp, vp, vp, l, l, i, C.c_float, C.c_float,
                                 C.POINTER(C.c_float), vp]),
        "photic_mlp_forward": (i, [vp, vp, l, vp, vp]),
        "photic_mlp_seed": (i, [vp, C.POINTER(Cfg), l, vp, vp, vp, vp]),
        "photic_mlp_seed_stack": (i, [vp, C.POINTER(Cfg), l, i, vp, vp, vp, vp, vp]),
        "photic_invert_stack_stream": (i, [C.POINTER(Cfg), vp, l, l, i, vp, vp, vp, l,
                                           C.POINTER(Outputs), PROGRESS_FN, vp]),
        "photic_order_cfg_default": (None, [C.POINTER(OrderCfg)]),
        "photic_mlp_set_seed_skip": (i, [vp, d]),
        "photic_invert_ordered": (i, [C.POINTER(Cfg), C.POINTER(OrderCfg), l, vp, vp,
                                      C.POINTER(Outputs), vp, C.POINTER(OrderStats), vp]),
        "photic_invert_h": (i, [C.POINTER(Cfg), l, vp, vp, vp, vp, vp, C.POINTER(Outputs),
                                vp]),
        "photic_band_coefficients": (i, [C.POINTER(Cfg), vp, C.POINTER(d)]),
        "photic_invert_stack": (i, [C.POINTER(Cfg), l, i, vp, vp, vp, vp, C.POINTER(Outputs),
                                    vp]),
        "photic_invert_stack_h": (i, [C.POINTER(Cfg), l, i, vp, vp, vp, vp, vp,
                                      C.POINTER(Outputs), vp]),
        "photic_gen_field": (i, [C.POINTER(Cfg), C.c_uint64, l, l, C.c_double, C.c_double, vp, vp,
                                 vp]),
        "photic_invert_regions": (i, [C.POINTER(Cfg), i, i, i, vp, vp, vp, vp, i, l, vp, vp,
                                      C.POINTER(Outputs), vp, vp]),
        "photic_invert_regions_bands": (i, [C.POINTER(Cfg), i, i, i, vp, vp, vp, vp, vp, vp, i,
                                            l, vp, vp, C.POINTER(Outputs), vp, vp]),
        "photic_autotune": (i, [C.POINTER(Cfg), l, vp, vp, vp, C.POINTER(Tuning), vp]),
        "photic_autotune_stack": (i, [C.POINTER(Cfg), l, i, vp, vp, vp, vp, C.POINTER(Tuning),
                                      vp]),
        "photic_gen_stack": (i, [C.POINTER(Cfg), C.c_uint64, l, l, i, d, d, i, vp, vp, vp, vp,
                                 vp]),
        "ph_build_const": (i, [C.POINTER(Cfg), i, i, vp]),
        "ph_const_size": (i, []),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


lib = _load()


def last_error():
    """The message of the library's last failure on this thread."""
    return lib.photic_last_error().decode()


def check(rc):
    if rc != OK:
        raise PhoticError(rc, lib.photic_last_error().decode())


def device_count():
    return lib.photic_device_count()


def cfg_default(wavelengths=(443.0, 482.0, 561.0, 655.0), **kw):
    """The reference's settings (model/samodel.c:2142-2147) in MIX mode; keyword overrides."""
    cfg = Cfg()
    wl = (C.c_double * MAX_BANDS)(*wavelengths)
    lib.photic_cfg_default(C.byref(cfg), len(wavelengths), wl)
    for k, v in kw.items():
        if k == "hstart":
            cfg.n_hstarts = len(v)
            for j, h in enumerate(v):
                cfg.hstart[j] = h
        elif k == "bottom_idx":
            for j, b in enumerate(v):
                cfg.bottom_idx[j] = b
        elif k in ("seed_rel_step", "seed_min_step"):
            arr = getattr(cfg, k)
            for j, s in enumerate(v):
                arr[j] = s
        else:
            setattr(cfg, k, v)
    return cfg


def n_params(cfg):
    return lib.photic_n_params(C.byref(cfg))


class DeviceBuffer:
    """A device allocation; frees itself."""

    def __init__(self, nbytes):
        self.ptr = C.c_void_p()
        self.nbytes = int(nbytes)
        check(lib.photic_device_malloc(C.byref(self.ptr), max(self.nbytes, 1)))

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        buf = cls(a.nbytes)
        if a.nbytes:
            check(lib.photic_memcpy_htod(buf.ptr, a.ctypes.data, a.nbytes))
        return buf

    def to_array(self, dtype, shape):
        return self.copy_to(np.empty(shape, dtype=dtype))

    def copy_to(self, out):
        """Copy the buffer's first out.nbytes bytes into the C-contiguous host array out."""
        if not out.flags.c_contiguous or out.nbytes > self.nbytes:
            raise ValueError("copy_to needs a C-contiguous array of at most %d bytes"
                             % self.nbytes)
        if out.nbytes:
            check(lib.photic_memcpy_dtoh(out.ctypes.data, self.ptr, out.nbytes))
        return out

    @property
    def value(self):
        return self.ptr.value

    def free(self):
        if self.ptr is not None and self.ptr.value:
            lib.photic_device_free(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class _HostMemory:
    """One page-locked host allocation; frees itself."""

    def __init__(self, nbytes):
        self.ptr = C.c_void_p()
        check(lib.photic_host_malloc(C.byref(self.ptr), max(nbytes, 1)))

    def __del__(self):
        try:
            if self.ptr is not None and self.ptr.value:
                lib.photic_host_free(self.ptr)
            self.ptr = None
        except Exception:
            pass


class PinnedBuffer:
    """Page-locked host memory seen as a numpy array (.array); device-to-host copies into it
    skip the runtime's staging copy. The memory is freed once no array taken from it is
    left."""

    def __init__(self, dtype, shape):
        dtype = np.dtype(dtype)
        shape = tuple(int(v) for v in np.atleast_1d(shape))
        nbytes = dtype.itemsize * int(np.prod(shape))
        mem = _HostMemory(nbytes)
        raw = (C.c_char * max(nbytes, 1)).from_address(mem.ptr.value)
        raw._owner = mem  # every array viewing raw keeps the allocation alive
        self.array = np.frombuffer(raw, dtype=np.uint8, count=nbytes).view(dtype).reshape(shape)


_OUT_FIELDS = {
    # name: (dtype, rows)
    "depth": (np.float32, 1), "rrs_error": (np.float32, 1), "model_error": (np.float32, 1),
    "albedo": (np.float32, 1), "k_min": (np.float32, 1), "iod": (np.float32, 1),
    "b_pct": (np.float32, 3), "bottom_type": (np.float32, 1), "x": (np.float64, None),
    "status": (np.int32, 1), "n_evals": (np.int32, 1), "n_iter": (np.int32, 1),
    "n_starts": (np.int32, 1), "rrs_model": (np.float64, 4),
}
# fields the stack and region kernels do not write
_SINGLE_DATE_ONLY = ("rrs_model",)


class DeviceOutputs:
    """Device output buffers for n pixels (all fields, or those named)."""

    def __init__(self, cfg, n, fields=None):
        self.n = n
        self.np = n_params(cfg)
        self.bufs = {}
        self.struct = Outputs()
        for name, (dt, rows) in _OUT_FIELDS.items():
            if fields is not None and name not in fields:
                continue
            rows = self.np if rows is None else rows
            b = DeviceBuffer(np.dtype(dt).itemsize * rows * n)
            self.bufs[name] = (b, dt, rows)
            setattr(self.struct, name, b.value)

    def _shape(self, rows):
        return (self.n,) if rows == 1 else (rows, self.n)

    def pinned(self):
        """Page-locked host arrays for every field, shaped as fetch() returns them."""
        return {name: PinnedBuffer(dt, self._shape(rows)).array
                for name, (b, dt, rows) in self.bufs.items()}

    def fetch(self, into=None):
        """The fields as host arrays: new ones, or copied into the arrays of `into` (a dict
        such as pinned() returns)."""
        out = {}
        for name, (b, dt, rows) in self.bufs.items():
            if into is None:
                out[name] = b.to_array(dt, self._shape(rows))
            else:
                dst = into[name]
                if dst.dtype != np.dtype(dt) or dst.shape != self._shape(rows):
                    raise ValueError("fetch into %s: need %s %s" % (name, np.dtype(dt),
                                                                    self._shape(rows)))
                out[name] = b.copy_to(dst)
        return out


def invert_device(cfg, n, d_rrs, d_sec_sun=None, d_seed=None, outs=None, stream=None,
                  d_sec_view=None):
    """photic_invert on device pointers (ints or None); returns the DeviceOutputs."""
    if outs is None:
        outs = DeviceOutputs(cfg, n)
    check(lib.photic_invert(C.byref(cfg), n, d_rrs, d_sec_sun, d_sec_view, d_seed,
                            C.byref(outs.struct), stream))
    return outs


def autotune(cfg, n, d_rrs, d_sec_sun=None, d_seed=None, stream=None):
    """Choose and keep the launch shape for cfg's kernel instance on a device-resident sample
    of n pixels; returns the winner as a dict."""
    t = Tuning()
    check(lib.photic_autotune(C.byref(cfg), n, d_rrs, d_sec_sun, d_seed, C.byref(t), stream))
    return {"block": t.block, "blocks_per_cu": t.blocks_per_cu, "chunk": t.chunk,
            "ms": t.ms, "n_tried": t.n_tried}


def autotune_stack(cfg, n, n_dates, d_rrs, d_sec_sun, d_valid, d_seed=None, stream=None):
    """The same for the stack kernels, on a sample of n stacks."""
    t = Tuning()
    check(lib.photic_autotune_stack(C.byref(cfg), n, n_dates, d_rrs, d_sec_sun, d_valid, d_seed,
                                    C.byref(t), stream))
    return {"block": t.block, "blocks_per_cu": t.blocks_per_cu, "chunk": t.chunk,
            "ms": t.ms, "n_tried": t.n_tried}


def invert(cfg, rrs, sec_sun=None, seed=None, fields=None):
    """Invert host arrays: rrs [n_bands][n] float32, sec_sun [n], seed [n_params][n]."""
    rrs = np.ascontiguousarray(rrs, dtype=np.float32)
    n = rrs.shape[1]
    d_rrs = DeviceBuffer.from_array(rrs)
    d_ss = DeviceBuffer.from_array(np.asarray(sec_sun, np.float32)) if sec_sun is not None else None
    d_seed = DeviceBuffer.from_array(np.asarray(seed, np.float32)) if seed is not None else None
    outs = DeviceOutputs(cfg, n, fields)
    invert_device(cfg, n, d_rrs.value, d_ss.value if d_ss else None,
                  d_seed.value if d_seed else None, outs)
    check(lib.photic_synchronize())
    return outs.fetch()


def order_cfg(**kw):
    """The ordered inversion's settings (photic_order_cfg_default), keyword overrides."""
    o = OrderCfg()
    lib.photic_order_cfg_default(C.byref(o))
    for k, v in kw.items():
        if k == "r_sigma":
            for j, x in enumerate(v):
                o.r_sigma[j] = x
        else:
            setattr(o, k, v)
    return o


def invert_ordered_device(cfg, ocfg, n, d_rrs, d_sec_sun=None, outs=None, d_role=None,
                          stream=None):
    """photic_invert_ordered on device pointers; returns (DeviceOutputs, stats dict)."""
    if outs is None:
        outs = DeviceOutputs(cfg, n)
    st = OrderStats()
    check(lib.photic_invert_ordered(C.byref(cfg), C.byref(ocfg), n, d_rrs, d_sec_sun,
                                    C.byref(outs.struct), d_role, C.byref(st), stream))
    return outs, {k: getattr(st, k) for k, _ in OrderStats._fields_}


def invert_ordered(cfg, ocfg, rrs, sec_sun=None, fields=None):
    """The ordered inversion of host arrays rrs [4][n], sec_sun [n]; returns (outputs with
    "role", stats)."""
    rrs = np.ascontiguousarray(rrs, dtype=np.float32)
    n = rrs.shape[1]
    d_rrs = DeviceBuffer.from_array(rrs)
    d_ss = DeviceBuffer.from_array(np.asarray(sec_sun, np.float32)) if sec_sun is not None else None
    d_role = DeviceBuffer(4 * max(n, 1))
    outs, stats = invert_ordered_device(cfg, ocfg, n, d_rrs.value, d_ss.value if d_ss else None,
                                        DeviceOutputs(cfg, n, fields), d_role.value)
    check(lib.photic_synchronize())
    res = outs.fetch()
    res["role"] = d_role.to_array(np.int32, (n,))
    return res, stats


class StackOutputs(DeviceOutputs):
    """Device output buffers for n stacks of n_dates dates (x is [off + 3 n_dates][n], off = 2
    in the mix layout, 1 + 2 n_bottoms in the native one)."""

    def __init__(self, cfg, n, n_dates, fields=None):
        self.n = n
        self.np = (2 if cfg.mode == MODE_MIX else 1 + 2 * cfg.n_bottoms) + 3 * n_dates
        self.bufs = {}
        self.struct = Outputs()
        for name, (dt, rows) in _OUT_FIELDS.items():
            if name in _SINGLE_DATE_ONLY:
                continue
            if fields is not None and name not in fields:
                continue
            rows = self.np if rows is None else rows
            b = DeviceBuffer(np.dtype(dt).itemsize * rows * n)
            self.bufs[name] = (b, dt, rows)
            setattr(self.struct, name, b.value)


def invert_stack_device(cfg, n, n_dates, d_rrs, d_sec_sun, d_valid, d_seed=None, outs=None,
                        stream=None):
    """photic_invert_stack on device pointers (ints or None); returns the StackOutputs."""
    if outs is None:
        outs = StackOutputs(cfg, n, n_dates)
    check(lib.photic_invert_stack(C.byref(cfg), n, n_dates, d_rrs, d_sec_sun, d_valid, d_seed,
                                  C.byref(outs.struct), stream))
    return outs


def invert_stack(cfg, rrs, sec_sun, valid, seed=None, fields=None, sec_view=None, hstart=None):
    """Invert host arrays: rrs [D][4][n] float32, sec_sun [D][n], valid [D][n] uint8,
    seed [2 + 3D][n] or None; sec_view [D][n] and hstart [n] (photic_invert_stack_h) or
    None."""
    rrs = np.ascontiguousarray(rrs, dtype=np.float32)
    nd, _, n = rrs.shape
    d_rrs = DeviceBuffer.from_array(rrs)
    d_ss = DeviceBuffer.from_array(np.asarray(sec_sun, np.float32))
    d_v = DeviceBuffer.from_array(np.asarray(valid, np.uint8))
    d_seed = DeviceBuffer.from_array(np.asarray(seed, np.float32)) if seed is not None else None
    outs = StackOutputs(cfg, n, nd, fields)
    if sec_view is not None or hstart is not None:
        assert seed is None
        d_sv = DeviceBuffer.from_array(np.asarray(sec_view, np.float32)) if sec_view is not None else None
        d_h = DeviceBuffer.from_array(np.asarray(hstart, np.float32)) if hstart is not None else None
        check(lib.photic_invert_stack_h(C.byref(cfg), n, nd, d_rrs.value, d_ss.value,
                                        d_sv.value if d_sv else None, d_v.value,
                                        d_h.value if d_h else None, C.byref(outs.struct), None))
    else:
        invert_stack_device(cfg, n, nd, d_rrs.value, d_ss.value, d_v.value,
                            d_seed.value if d_seed is not None else None, outs)
    check(lib.photic_synchronize())
    return outs.fetch()


class RegionOutputs(DeviceOutputs):
    """Device output buffers for n pixels of a region run (x is [R_max (1 + 2 n_bottoms) +
    3 n_scenes][n], R_max = (2 n_spatial - 1)^2), plus the region count per pixel."""

    def __init__(self, cfg, n, n_scenes, n_spatial, fields=None):
        self.n = n
        ns = max(n_spatial, 1)
        self.np = (2 * ns - 1) ** 2 * (1 + 2 * cfg.n_bottoms) + 3 * n_scenes
        self.bufs = {}
        self.struct = Outputs()
        for name, (dt, rows) in _OUT_FIELDS.items():
            if name in _SINGLE_DATE_ONLY:
                continue
            if fields is not None and name not in fields:
                continue
            rows = self.np if rows is None else rows
            b = DeviceBuffer(np.dtype(dt).itemsize * rows * n)
            self.bufs[name] = (b, dt, rows)
            setattr(self.struct, name, b.value)
        b = DeviceBuffer(4 * n)
        self.bufs["n_regions"] = (b, np.int32, 1)
        self.n_regions = b


def invert_regions_device(cfg, n_rows, n_cols, n_scenes, d_rrs, nodata, sec_sun, sec_view,
                          n_spatial, n, d_idx=None, d_hstart=None, outs=None, stream=None):
    """photic_invert_regions on device pointers (ints or None) and host nodata [4 n_scenes],
    sec_sun, sec_view [n_scenes]; returns the RegionOutputs."""
    if outs is None:
        outs = RegionOutputs(cfg, n, n_scenes, n_spatial)
    nod = np.ascontiguousarray(np.broadcast_to(np.asarray(nodata, np.float32), (n_scenes, 4)))
    ss = np.ascontiguousarray(np.broadcast_to(np.asarray(sec_sun, np.float64), (n_scenes,)))
    sv = np.ascontiguousarray(np.broadcast_to(np.asarray(sec_view, np.float64), (n_scenes,)))
    check(lib.photic_invert_regions(C.byref(cfg), n_rows, n_cols, n_scenes, d_rrs,
                                    nod.ctypes.data, ss.ctypes.data, sv.ctypes.data, n_spatial,
                                    n, d_idx, d_hstart, C.byref(outs.struct),
                                    outs.n_regions.value, stream))
    return outs


def invert_regions(cfg, rrs, n_rows, n_cols, nodata, sec_sun, sec_view, n_spatial=2, idx=None,
                   hstart=None, fields=None):
    """Invert host arrays: rrs [n_scenes][4][n_rows * n_cols] float32, nodata [n_scenes][4],
    sec_sun / sec_view [n_scenes]; idx [n] int64 pixel indices or None (every pixel, screened);
    hstart [n] or None.  Returns host outputs, with n_regions."""
    rrs = np.ascontiguousarray(rrs, dtype=np.float32)
    nd = rrs.shape[0]
    assert rrs.shape == (nd, 4, n_rows * n_cols)
    n = n_rows * n_cols if idx is None else len(idx)
    d_rrs = DeviceBuffer.from_array(rrs)
    d_idx = DeviceBuffer.from_array(np.asarray(idx, np.int64)) if idx is not None else None
    d_h = DeviceBuffer.from_array(np.asarray(hstart, np.float32)) if hstart is not None else None
    outs = RegionOutputs(cfg, n, nd, n_spatial, fields)
    invert_regions_device(cfg, n_rows, n_cols, nd, d_rrs.value, nodata, sec_sun, sec_view,
                          n_spatial, n, d_idx.value if d_idx is not None else None,
                          d_h.value if d_h is not None else None, outs)
    check(lib.photic_synchronize())
    return outs.fetch()


def invert_regions_bands(cfg, rrs_rows, wavelengths, n_rows, n_cols, nodata, sec_sun, sec_view,
                         n_spatial=2, idx=None, hstart=None, fields=None):
    """photic_invert_regions_bands on host arrays: wavelengths[s] the band list of scene s;
    rrs_rows [sum len(wavelengths[s])][n_rows * n_cols] float32, the scenes' band grids one
    after another; nodata one value per band grid (or a scalar); sec_sun / sec_view
    [n_scenes]; idx [n] int64 pixel indices or None; hstart [n] or None."""
    rrs = np.ascontiguousarray(rrs_rows, dtype=np.float32)
    nbs = np.asarray([len(w) for w in wavelengths], np.int32)
    nd = len(nbs)
    rows = int(nbs.sum())
    assert rrs.shape == (rows, n_rows * n_cols)
    wl = np.ascontiguousarray(np.concatenate([np.asarray(w, np.float64) for w in wavelengths]))
    n = n_rows * n_cols if idx is None else len(idx)
    d_rrs = DeviceBuffer.from_array(rrs)
    d_idx = DeviceBuffer.from_array(np.asarray(idx, np.int64)) if idx is not None else None
    d_h = DeviceBuffer.from_array(np.asarray(hstart, np.float32)) if hstart is not None else None
    outs = RegionOutputs(cfg, n, nd, n_spatial, fields)
    nod = np.ascontiguousarray(np.broadcast_to(np.asarray(nodata, np.float32), (rows,)))
    ss = np.ascontiguousarray(np.broadcast_to(np.asarray(sec_sun, np.float64), (nd,)))
    sv = np.ascontiguousarray(np.broadcast_to(np.asarray(sec_view, np.float64), (nd,)))
    check(lib.photic_invert_regions_bands(C.byref(cfg), n_rows, n_cols, nd, nbs.ctypes.data,
                                          wl.ctypes.data, d_rrs.value, nod.ctypes.data,
                                          ss.ctypes.data, sv.ctypes.data, n_spatial, n,
                                          d_idx.value if d_idx is not None else None,
                                          d_h.value if d_h is not None else None,
                                          C.byref(outs.struct), outs.n_regions.value, None))
    check(lib.photic_synchronize())
    return outs.fetch()


def gen_field(cfg, seed, n_rows, n_cols, theta=35.0, sigma=2e-4):
    """The region workload's coherent scene on the GPU: (rrs [nb][n], truth [6][n])."""
    nb = cfg.n_bands
    n = n_rows * n_cols
    d_rrs, d_tr = DeviceBuffer(4 * nb * n), DeviceBuffer(4 * 6 * n)
    check(lib.photic_gen_field(C.byref(cfg), seed, n_rows, n_cols, theta, sigma, d_rrs.value,
                               d_tr.value, None))
    check(lib.photic_synchronize())
    return d_rrs.to_array(np.float32, (nb, n)), d_tr.to_array(np.float32, (6, n))


def invert_stack_stream(cfg, n_rows, n_cols, rrs, sec_sun, valid, tile_rows=2048, net=None,
                        fields=None, progress=None):
    """photic_invert_stack_stream on HOST arrays of an n_rows x n_cols scene: rrs [D][4][n]
    float32, sec_sun [D][n], valid [D][n] uint8 (n = n_rows * n_cols), streamed through the
    GPU in tiles of tile_rows rows, MLP-seeded when net is given.  Returns host outputs as
    invert_stack; progress(tiles_done, n_tiles) is called as tiles finish."""
    nd = rrs.shape[0]
    n = n_rows * n_cols
    assert rrs.shape == (nd, 4, n) and rrs.dtype == np.float32 and rrs.flags.c_contiguous
    assert sec_sun.shape == (nd, n) and sec_sun.dtype == np.float32 and sec_sun.flags.c_contiguous
    assert valid.shape == (nd, n) and valid.dtype == np.uint8 and valid.flags.c_contiguous
    struct = Outputs()
    res = {}
    for name, (dt, rows) in _OUT_FIELDS.items():
        if name in _SINGLE_DATE_ONLY:
            continue
        if fields is not None and name not in fields:
            continue
        rows = 2 + 3 * nd if rows is None else rows
        a = np.empty((rows, n) if rows > 1 else (n,), dt)
        res[name] = a
        setattr(struct, name, a.ctypes.data)
    def _cb(done, tot, user):
        if progress is not None:
            progress(done, tot)
    cb = PROGRESS_FN(_cb)
    check(lib.photic_invert_stack_stream(C.byref(cfg), net.h if net is not None else None,
                                         n_rows, n_cols, nd, rrs.ctypes.data,
                                         sec_sun.ctypes.data, valid.ctypes.data, tile_rows,
                                         C.byref(struct), cb, None))
    return res


def gen_stack(cfg, seed, n, n_dates=8, pix0=0, sigma=2e-4, p_cloud=0.3, k_min=3, truth=True):
    """Synthetic stack scene on the GPU: (rrs [D][4][n], sec_sun [D][n], valid [D][n],
    truth [2 + 4D][n] or None)."""
    nb = cfg.n_bands
    d_rrs = DeviceBuffer(4 * n_dates * nb * n)
    d_ss = DeviceBuffer(4 * n_dates * n)
    d_v = DeviceBuffer(n_dates * n)
    d_tr = DeviceBuffer(4 * (2 + 4 * n_dates) * n) if truth else None
    check(lib.photic_gen_stack(C.byref(cfg), seed, pix0, n, n_dates, sigma, p_cloud, k_min,
                               d_rrs.value, d_ss.value, d_v.value,
                               d_tr.value if d_tr else None, None))
    check(lib.photic_synchronize())
    return (d_rrs.to_array(np.float32, (n_dates, nb, n)),
            d_ss.to_array(np.float32, (n_dates, n)), d_v.to_array(np.uint8, (n_dates, n)),
            d_tr.to_array(np.float32, (2 + 4 * n_dates, n)) if truth else None)


def disc_count(n_rows, n_cols, frac, r_min, r_max):
    """Discs of radius U(r_min, r_max) that mask about `frac` of an n_rows x n_cols scene."""
    mean_r2 = (r_min * r_min + r_min * r_max + r_max * r_max) / 3.0
    return int(np.ceil(-np.log1p(-frac) * n_rows * n_cols / (np.pi * mean_r2)))


def nodata_discs_device(seed, n_rows, n_cols, n_discs, r_min, r_max, n_bands, d_rrs,
                        nodata=-9999.0, stream=None):
    """photic_gen_nodata_discs on a device Rrs array [n_bands][n_rows * n_cols]."""
    check(lib.photic_gen_nodata_discs(seed, n_rows, n_cols, n_discs, r_min, r_max, n_bands,
                                      nodata, d_rrs, stream))


def gen_scene(cfg, seed, n, pix0=0, sigma=2e-4, deep_frac=0.0, truth=True):
    """Synthetic scene on the GPU: returns (rrs [nb][n], sec_sun [n], truth [6][n] or None)."""
    nb = cfg.n_bands
    d_rrs = DeviceBuffer(4 * nb * n)
    d_ss = DeviceBuffer(4 * n)
    d_tr = DeviceBuffer(4 * 6 * n) if truth else None
    check(lib.photic_gen_scene(C.byref(cfg), seed, pix0, n, sigma, deep_frac, d_rrs.value,
                               d_ss.value, d_tr.value if d_tr else None, None))
    check(lib.photic_synchronize())
    return (d_rrs.to_array(np.float32, (nb, n)), d_ss.to_array(np.float32, (n,)),
            d_tr.to_array(np.float32, (6, n)) if truth else None)


def forward(cfg, params, sec_sun, Y, precision=FP64):
    """Forward model on the GPU: params [n][5] (H, P, G, X, B) -> Rrs [n][n_bands] (float64)."""
    params = np.ascontiguousarray(params, np.float64)
    n = params.shape[0]
    bufs = [DeviceBuffer.from_array(params), DeviceBuffer.from_array(np.asarray(sec_sun, np.float64)),
            DeviceBuffer.from_array(np.asarray(Y, np.float64)), DeviceBuffer(8 * n * cfg.n_bands)]
    check(lib.photic_forward(C.byref(cfg), n, bufs[0].value, bufs[1].value, bufs[2].value,
                             bufs[3].value, precision, None))
    check(lib.photic_synchronize())
    return bufs[3].to_array(np.float64, (n, cfg.n_bands))


def objective(cfg, x, rrs, sec_sun, precision=FP64):
    """Misfit on the GPU: x [n][5] (MIX layout), rrs [n][n_bands] float32, sec_sun [n]."""
    x = np.ascontiguousarray(x, np.float64)
    n = x.shape[0]
    bx = DeviceBuffer.from_array(x)
    br = DeviceBuffer.from_array(np.ascontiguousarray(rrs, np.float32))
    bs = DeviceBuffer.from_array(np.asarray(sec_sun, np.float32))
    bo, be = DeviceBuffer(8 * n), DeviceBuffer(8 * n)
    check(lib.photic_objective(C.byref(cfg), n, bx.value, br.value, bs.value, bo.value, be.value,
                               precision, None))
    check(lib.photic_synchronize())
    return bo.to_array(np.float64, (n,)), be.to_array(np.float64, (n,))


def band_coefficients(cfg):
    """(coefficients [12][n_bands]: a0, a1, a_w, b_bw, 8 bottoms; a_w(640)) for cfg."""
    out = np.zeros((12, cfg.n_bands))
    aw640 = C.c_double()
    check(lib.photic_band_coefficients(C.byref(cfg), out.ctypes.data, C.byref(aw640)))
    return out, aw640.value


def kernel_info(cfg):
    v, s, b, n = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    check(lib.photic_kernel_info(C.byref(cfg), C.byref(v), C.byref(s), C.byref(b), C.byref(n)))
    return {"vgprs": v.value, "scratch_bytes": s.value, "blocks_per_cu": b.value, "n_cu": n.value}


def const_bytes(cfg, n_dates=1, has_seed=False):
    """The launch constants (PhConst) for cfg, as bytes -- used by the host test build."""
    buf = C.create_string_buffer(lib.ph_const_size())
    check(lib.ph_build_const(C.byref(cfg), n_dates, int(has_seed), buf))
    return buf


class Mlp:
    """The warm-start network on the device (photic_mlp_* of include/photic.h)."""

    def __init__(self, init_seed=1):
        self.h = C.c_void_p()
        check(lib.photic_mlp_create(C.byref(self.h), init_seed))
        self.n_params = lib.photic_mlp_n_params()

    def close(self):
        if self.h is not None and self.h.value:
            lib.photic_mlp_destroy(self.h)
        self.h = None

    def set_seed_skip(self, h_max):
        """Seeds whose predicted depth is below h_max are NaN: no seed start there."""
        check(lib.photic_mlp_set_seed_skip(self.h, h_max))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_weights(self, w):
        w = np.ascontiguousarray(w, np.float32)
        assert w.size == self.n_params
        check(lib.photic_mlp_set_weights(self.h, w.ctypes.data))

    def get_weights(self):
        w = np.empty(self.n_params, np.float32)
        check(lib.photic_mlp_get_weights(self.h, w.ctypes.data))
        return w

    def gradient(self, x, t):
        """Loss and gradient over host samples x, t [n][5] (n a multiple of 16)."""
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        bx, bt = DeviceBuffer.from_array(x), DeviceBuffer.from_array(t)
        g = np.empty(self.n_params, np.float32)
        loss = C.c_float()
        check(lib.photic_mlp_gradient(self.h, bx.value, bt.value, x.shape[0], g.ctypes.data,
                                      C.byref(loss)))
        return loss.value, g

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32)
        bx, by = DeviceBuffer.from_array(x), DeviceBuffer(x.nbytes)
        check(lib.photic_mlp_forward(self.h, bx.value, x.shape[0], by.value, None))
        check(lib.photic_synchronize())
        return by.to_array(np.float32, x.shape)

    def train_device(self, d_x, d_t, n, batch=8192, epochs=1, lr=1e-3, lr_final=1e-4,
                     stream=None):
        loss = C.c_float()
        check(lib.photic_mlp_train(self.h, d_x, d_t, n, batch, epochs, lr, lr_final,
                                   C.byref(loss), stream))
        return loss.value

    def seed_device(self, cfg, n, d_rrs, d_sec_sun, d_seed, stream=None):
        check(lib.photic_mlp_seed(self.h, C.byref(cfg), n, d_rrs, d_sec_sun, d_seed, stream))

    def seed_stack_device(self, cfg, n, n_dates, d_rrs, d_sec_sun, d_valid, d_seed,
                          stream=None):
        """Stack seeds [2 + 3 D][n] on the device (photic_mlp_seed_stack)."""
        check(lib.photic_mlp_seed_stack(self.h, C.byref(cfg), n, n_dates, d_rrs, d_sec_sun,
                                        d_valid, d_seed, stream))

    def seed(self, cfg, rrs, sec_sun):
        """Seeds [5][n] for host arrays rrs [4][n], sec_sun [n]."""
        rrs = np.ascontiguousarray(rrs, np.float32)
        n = rrs.shape[1]
        br, bs = DeviceBuffer.from_array(rrs), DeviceBuffer.from_array(
            np.ascontiguousarray(sec_sun, np.float32))
        bo = DeviceBuffer(4 * 5 * n)
        self.seed_device(cfg, n, br.value, bs.value, bo.value)
        check(lib.photic_synchronize())
        return bo.to_array(np.float32, (5, n))

    def seed_stack(self, cfg, rrs, sec_sun, valid):
        """Stack seeds [2 + 3 D][n] for host arrays rrs [D][4][n], sec_sun [D][n], valid [D][n]."""
        rrs = np.ascontiguousarray(rrs, np.float32)
        nd, n = rrs.shape[0], rrs.shape[2]
        br = DeviceBuffer.from_array(rrs)
        bs = DeviceBuffer.from_array(np.ascontiguousarray(sec_sun, np.float32))
        bv = DeviceBuffer.from_array(np.ascontiguousarray(valid, np.uint8))
        bo = DeviceBuffer(4 * (2 + 3 * nd) * n)
        self.seed_stack_device(cfg, n, nd, br.value, bs.value, bv.value, bo.value)
        check(lib.photic_synchronize())
        return bo.to_array(np.float32, (2 + 3 * nd, n))


def trainset_device(cfg, seed, n, pix0=0, sigma=2e-4):
    """Normalised training pairs on the device: (DeviceBuffer x [n][5], DeviceBuffer t [n][5])."""
    dx, dt = DeviceBuffer(4 * 5 * n), DeviceBuffer(4 * 5 * n)
    check(lib.photic_mlp_trainset(C.byref(cfg), seed, pix0, n, sigma, dx.value, dt.value, None))
    check(lib.photic_synchronize())
    return dx, dt
