"""The headline workload at its full size: the 8192 x 8192 single-date scene (67.1 M pixels,
Philox seed 4, noise 2e-4), warm-started by the network and inverted whole on the GPU with the
seed-start abort at 5, as the benchmark runs it.  Pixels are independent, so every pixel's
result must be the one the host build of the engine gives for that pixel alone with the same
seed vector: checked bit for bit (status, parameters, misfits, evaluation, iteration and
start counts) on 16 blocks of 64 pixels spread over the scene.  The host engine is itself
pinned to the FP64 oracle (test_engine_host.py)."""
import ctypes

import numpy as np
import pytest

import _oracle as O
import photic_amd as P

pytestmark = pytest.mark.gpu

SIDE = 8192
BLOCK, N_BLOCKS = 64, 16


def _rows(ptr, itemsize, n, rows, q0, m):
    """Rows [rows][n] of a device array, columns q0 .. q0 + m - 1, as a host array."""
    out = np.empty((rows, m), dtype={4: np.float32, 8: np.float64}[itemsize])
    for k in range(rows):
        P.check(P.lib.photic_memcpy_dtoh(out[k].ctypes.data, ptr + itemsize * (k * n + q0),
                                         itemsize * m))
    return out


def test_headline_scene_full_size_matches_host_engine():
    n = SIDE * SIDE
    cfg = P.cfg_default(reqmin=1e-10, konvge=100, kcount=2000, check_watchdog=1,
                        seed_abort_misfit=5.0)
    d_rrs, d_ss = P.DeviceBuffer(16 * n), P.DeviceBuffer(4 * n)
    P.check(P.lib.photic_gen_scene(ctypes.byref(cfg), 4, 0, n, 2e-4, 0.0, d_rrs.value,
                                   d_ss.value, None, None))
    n_tr = 1 << 22
    dx, dt = P.trainset_device(cfg, 101, n_tr)
    net = P.Mlp(init_seed=1)
    net.train_device(dx.value, dt.value, n_tr, batch=8192, epochs=1, lr=2e-3, lr_final=1e-4)
    P.check(P.lib.photic_synchronize())
    del dx, dt
    d_seed = P.DeviceBuffer(20 * n)
    net.seed_device(cfg, n, d_rrs.value, d_ss.value, d_seed.value)
    P.check(P.lib.photic_synchronize())
    net.close()
    fields = ("x", "status", "n_evals", "n_iter", "n_starts", "rrs_error", "model_error",
              "depth")
    outs = P.DeviceOutputs(cfg, n, fields=fields)
    P.invert_device(cfg, n, d_rrs.value, d_ss.value, d_seed.value, outs)
    P.check(P.lib.photic_synchronize())
    print("inverted %d pixels" % n, flush=True)

    rng = np.random.default_rng(4)
    starts = np.sort(rng.choice(n // BLOCK, N_BLOCKS, replace=False)) * BLOCK
    n_abort = 0
    for q0 in starts:
        q0 = int(q0)
        rrs = _rows(d_rrs.value, 4, n, 4, q0, BLOCK)
        ss = _rows(d_ss.value, 4, n, 1, q0, BLOCK)[0]
        seed = _rows(d_seed.value, 4, n, 5, q0, BLOCK)
        h = O.hostsim_invert(P.const_bytes(cfg, has_seed=True), 0, rrs, ss, seed=seed,
                             math="cr")
        for k in fields:
            buf, dtype, rows = outs.bufs[k]
            rows = outs.np if rows is None else rows
            g = _rows(buf.value, np.dtype(dtype).itemsize, n, rows, q0, BLOCK)
            g = g if k == "x" else g[0]
            if np.dtype(dtype).kind == "i":
                g = g.view(np.int32)
            assert np.array_equal(g, h[k]), (q0, k)
        n_abort += int((h["n_starts"] > 1).sum())
    print("blocks %d x %d pixels bit-identical to the host engine; %d fell back to the depth "
          "starts" % (N_BLOCKS, BLOCK, n_abort))
