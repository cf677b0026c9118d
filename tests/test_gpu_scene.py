"""The single-date path at scene scale: a 4000 x 3000 scene (12 M pixels) with 15 % optically
deep pixels (H ~ U(30, 60) m) and about 10 % no-data pixels in seeded random discs, inverted
whole on the GPU (plain method, reqmin 1e-10, kcount 2000, konvge 100) and checked pixel by
pixel against the FP64 oracle on a random sample of its pixels (status identical; at least
99 % bit-identical; H within 1 cm and P, G, X, B within 1e-4 relative wherever both converge,
or the same misfit to 1e-6 relative where it is flat, except on at most 0.2 % of them, each of
which must be shown to come from the host C library's rounding alone) and against the host
build of the engine with the same math (bit for bit)."""
import ctypes

import numpy as np
import pytest

import _oracle as O
import photic_amd as P

pytestmark = pytest.mark.gpu

ROWS, COLS = 3000, 4000
R_MIN, R_MAX = 20.0, 60.0


def test_scene_scale_parity():
    n = ROWS * COLS
    cfg = P.cfg_default(reqmin=1e-10, kcount=2000, konvge=100, check_watchdog=1)
    d_rrs, d_ss = P.DeviceBuffer(16 * n), P.DeviceBuffer(4 * n)
    P.check(P.lib.photic_gen_scene(ctypes.byref(cfg), 2, 0, n, 2e-4, 0.15, d_rrs.value,
                                   d_ss.value, None, None))
    n_discs = P.disc_count(ROWS, COLS, 0.10, R_MIN, R_MAX)
    P.nodata_discs_device(2, ROWS, COLS, n_discs, R_MIN, R_MAX, 4, d_rrs.value)
    P.check(P.lib.photic_synchronize())
    outs = P.DeviceOutputs(cfg, n)
    P.invert_device(cfg, n, d_rrs.value, d_ss.value, None, outs)
    P.check(P.lib.photic_synchronize())
    g = outs.fetch()
    # the same outputs copied into page-locked host arrays
    pin = outs.pinned()
    gp = outs.fetch(into=pin)
    for k in g:
        assert gp[k] is pin[k]
        assert np.array_equal(gp[k].view(np.uint8), g[k].view(np.uint8)), k
    del gp, pin
    rrs = d_rrs.to_array(np.float32, (4, n))
    ss = d_ss.to_array(np.float32, (n,))
    masked = np.all(rrs == -9999.0, axis=0)
    print("discs %d, masked %.4f, no-data %.4f" % (n_discs, masked.mean(), (g["status"] < 0).mean()))
    assert 0.07 < masked.mean() < 0.13
    assert np.all(g["status"][masked] == -1)

    rng = np.random.default_rng(2)
    idx = np.sort(rng.choice(n, 6000, replace=False))
    assert np.array_equal(O.disc_mask(2, ROWS, COLS, n_discs, R_MIN, R_MAX, idx), masked[idx])
    o = O.invert(rrs[:, idx], ss[idx], reqmin=1e-10, konvge=100, kcount=2000)
    gs = {k: (v[..., idx] if v.ndim == 2 else v[idx]) for k, v in g.items()}
    valid = o["status"] >= 0
    assert np.array_equal(gs["status"], o["status"])
    both = (gs["status"] == 0) & (o["status"] == 0)
    dH = np.abs(gs["x"][0] - o["x"][0])
    rel = np.abs(gs["x"][1:] - o["x"][1:]) / np.maximum(np.abs(o["x"][1:]), 1e-300)
    within = (dH < 0.01) & np.all(rel < 1e-4, axis=0)
    flat = np.abs(gs["rrs_error"] - o["rrs_error"]) <= 1e-6 * np.abs(o["rrs_error"])
    exact = np.all(gs["x"] == o["x"], axis=0)
    print("sample %d: valid %d, both converged %d, within %.4f, flat %d, bit-identical %.4f" %
          (len(idx), valid.sum(), both.sum(), within[both].mean(), (~within & flat)[both].sum(),
           exact[valid].mean()))
    assert exact[valid].mean() > 0.99

    # Pixels outside the bar: each must be explained by the host C library's rounding alone
    # -- the host build of the engine with glibc's exp/log/pow ends exactly where the oracle
    # ends, and with the GPU's correctly rounded functions exactly where the GPU ends -- and
    # they must stay rare (glibc misrounds about 1 call in 10^3).
    bad = np.nonzero(both & ~(within | flat))[0]
    for j in bad:
        print("outside the bar: pixel %d, H %.4f vs %.4f, E_Rrs %.6g vs %.6g" % (
            idx[j], gs["x"][0, j], o["x"][0, j], gs["rrs_error"][j], o["rrs_error"][j]))
    assert len(bad) <= 2e-3 * both.sum()
    if len(bad):
        sel = idx[bad]
        hl = O.hostsim_invert(P.const_bytes(cfg), 0, rrs[:, sel], ss[sel], math="libm")
        hc = O.hostsim_invert(P.const_bytes(cfg), 0, rrs[:, sel], ss[sel], math="cr")
        assert np.array_equal(hl["x"], o["x"][:, bad])
        assert np.array_equal(hc["x"], gs["x"][:, bad])

    m = 1024
    h = O.hostsim_invert(P.const_bytes(cfg), 0, rrs[:, idx[:m]], ss[idx[:m]], math="cr")
    assert np.array_equal(gs["status"][:m], h["status"])
    assert np.array_equal(gs["x"][:, :m], h["x"])
