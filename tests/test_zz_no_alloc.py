"""photic_invert_noalloc, the entry point that allocates nothing, called through the C ABI:
its outputs must equal photic_invert_h's on the same inputs bit for bit, on two calls in a
row (the second with a smaller input, so the shared scratch array keeps the first call's
contents when the second starts), plain and with a per-pixel depth start and a seed."""
import ctypes as C

import numpy as np
import pytest

import _oracle as O
import photic_amd as P

pytestmark = pytest.mark.gpu

_F = ("depth", "rrs_error", "model_error", "albedo", "k_min", "iod", "b_pct", "bottom_type",
      "x", "status", "n_evals", "n_iter", "n_starts")


def _noalloc():
    f = P.lib.photic_invert_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_void_p, C.c_long] + [C.c_void_p] * 5 + [C.c_void_p, C.c_void_p]
    return f


def _run(fn, cfg, rrs, ss, seed=None, hstart=None):
    n = rrs.shape[1]
    d_rrs = P.DeviceBuffer.from_array(np.ascontiguousarray(rrs, np.float32))
    d_ss = P.DeviceBuffer.from_array(np.ascontiguousarray(ss, np.float32))
    d_seed = None if seed is None else P.DeviceBuffer.from_array(
        np.ascontiguousarray(seed, np.float32))
    d_h = None if hstart is None else P.DeviceBuffer.from_array(
        np.ascontiguousarray(hstart, np.float32))
    outs = P.DeviceOutputs(cfg, n)
    rc = fn(C.addressof(cfg), n, d_rrs.value, d_ss.value, None,
            None if d_seed is None else d_seed.value, None if d_h is None else d_h.value,
            C.addressof(outs.struct), None)
    assert rc == P.OK, P.last_error()
    P.check(P.lib.photic_synchronize())
    return outs.fetch()


def _same(a, b):
    for k in _F:
        assert np.array_equal(a[k], b[k]), k


def test_noalloc_matches_invert_h():
    ref = P.lib.photic_invert_h
    ref.restype = C.c_int
    ref.argtypes = _noalloc().argtypes
    cfg = P.cfg_default(reqmin=1e-10, konvge=100, kcount=2000, check_watchdog=1)
    rrs, ss, truth, _, _ = O.gen(151, 4096)
    first = _run(_noalloc(), cfg, rrs, ss)
    _same(first, _run(ref, cfg, rrs, ss))
    assert (first["status"] == 0).mean() > 0.9
    # second call, smaller, with a seed and one depth start per pixel
    m = 1000
    rrs2, ss2, truth2, _, _ = O.gen(152, m)
    seed = np.stack([truth2[0], truth2[4], 100 * truth2[1], 100 * truth2[2],
                     100 * truth2[3]]).astype(np.float32) * np.float32(1.1)
    hs = np.full(m, 10.0, np.float32)
    second = _run(_noalloc(), cfg, rrs2, ss2, seed=seed, hstart=hs)
    _same(second, _run(ref, cfg, rrs2, ss2, seed=seed, hstart=hs))
    assert np.all(second["n_starts"][second["status"] >= 0] >= 1)
