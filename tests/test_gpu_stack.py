"""Multi-date stacks on the GPU (photic_invert_stack): the kernel against the host build of
the same engine with the same correctly rounded math (bit for bit on every pixel) and against
the FP64 oracle (H within 1 cm and the other parameters within 1e-4 relative wherever both
converge, or the same misfit where it is flat), the bucketing by valid-date count, no-data
pixels, seeds, and the GPU stack generator against the oracle's."""
import numpy as np
import pytest

import _oracle as O
import photic_amd as P

pytestmark = pytest.mark.gpu


def _cfg(**kw):
    return P.cfg_default(reqmin=1e-10, konvge=100, kcount=2000, check_watchdog=1, **kw)


@pytest.mark.parametrize("n_dates,k_min", [(8, 3), (4, 2)])
def test_stack_bit_identical_to_host_engine(n_dates, k_min):
    rrs, ss, valid, truth = O.gen_stack(31 + n_dates, 384, n_dates=n_dates, k_min=k_min)
    g = P.invert_stack(_cfg(), rrs, ss, valid)
    h = O.hostsim_invert_stack(P.const_bytes(_cfg()), rrs, ss, valid, math="cr")
    print("dates %d: k histogram %s" % (n_dates, np.bincount(valid.sum(0))))
    assert np.array_equal(g["status"], h["status"])
    assert np.array_equal(g["x"], h["x"])
    for k in ("n_iter", "n_evals", "n_starts", "rrs_error", "model_error", "depth"):
        assert np.array_equal(g[k], h[k]), k


def test_stack_against_oracle():
    rrs, ss, valid, truth = O.gen_stack(41, 256, n_dates=6)
    g = P.invert_stack(_cfg(), rrs, ss, valid)
    o = O.invert_stack(rrs, ss, valid)
    assert np.array_equal(g["status"], o["status"])
    both = (o["status"] == 0) & (g["status"] == 0)
    dH = np.abs(g["x"][0] - o["x"][0])
    rel = np.abs(g["x"][1:] - o["x"][1:]) / np.maximum(np.abs(o["x"][1:]), 1e-300)
    within = (dH < 0.01) & np.all((rel < 1e-4) | (o["x"][1:] == 0), axis=0)
    same_fit = np.abs(g["rrs_error"] - o["rrs_error"]) <= 1e-6 * np.abs(o["rrs_error"])
    exact = np.all(g["x"] == o["x"], axis=0)
    print("stacks: %d converged, within %.4f, bit-identical %.4f" %
          (both.sum(), within[both].mean(), exact[o["status"] >= 0].mean()))
    assert np.all((within | same_fit)[both])
    for k in ("k_min", "iod", "albedo", "bottom_type"):
        sel = (o["status"] >= 0) & exact
        assert np.allclose(g[k][sel], o[k][sel], rtol=1e-6, atol=1e-7), k


def test_stack_seeded_and_nodata():
    nd = 4
    rrs, ss, valid, truth = O.gen_stack(43, 192, n_dates=nd, k_min=2)
    valid[:, :5] = 0                    # no valid date
    rrs[1, 2, 7] = -9999.0              # no-data value on a valid date
    valid[1, 7] = 1
    seed = np.zeros((2 + 3 * nd, rrs.shape[2]), np.float32)
    seed[0], seed[1] = truth[0] * 0.9, truth[1]
    for d in range(nd):
        seed[2 + 3 * d:5 + 3 * d] = 100 * truth[2 + 4 * d:5 + 4 * d]
    g = P.invert_stack(_cfg(), rrs, ss, valid, seed=seed)
    h = O.hostsim_invert_stack(P.const_bytes(_cfg(), has_seed=True), rrs, ss, valid, seed=seed,
                               math="cr")
    assert np.all(g["status"][:5] == -1) and g["status"][7] == -1
    assert np.all(g["b_pct"][:, :5] == -9999.0)
    assert np.array_equal(g["status"], h["status"])
    assert np.array_equal(g["x"], h["x"])
    assert np.array_equal(g["n_starts"], h["n_starts"])


def test_stack_generator_matches_oracle():
    cfg = _cfg()
    rrs, ss, valid, truth = P.gen_stack(cfg, 17, 2048, n_dates=8)
    orrs, oss, ovalid, otruth = O.gen_stack(17, 2048, n_dates=8)
    assert np.array_equal(valid[ovalid == valid], ovalid[ovalid == valid])
    assert (valid == ovalid).mean() > 0.999
    ok = (valid == 1) & (ovalid == 1)
    d = np.abs(rrs.astype(np.float64) - orrs)[:, :, :][np.broadcast_to(ok[:, None, :], rrs.shape)]
    ulp = np.spacing(np.abs(orrs[np.broadcast_to(ok[:, None, :], rrs.shape)]))
    assert np.all(d <= ulp)
    assert np.allclose(ss, oss, rtol=2e-7)
    assert np.allclose(truth, otruth.astype(np.float32), rtol=2e-7)


def test_stack_empty_and_unsupported():
    out = P.invert_stack(_cfg(), np.zeros((3, 4, 0), np.float32), np.zeros((3, 0), np.float32),
                         np.zeros((3, 0), np.uint8))
    assert all(v.size == 0 for v in out.values())
    z = (np.zeros((2, 4, 8), np.float32), np.ones((2, 8), np.float32), np.ones((2, 8), np.uint8))
    with pytest.raises(P.PhoticError):   # seeds are made for the mix layout
        P.invert_stack(_cfg(mode=P.MODE_NATIVE, n_bottoms=2), *z, seed=np.zeros((7, 8), np.float32))
    with pytest.raises(P.PhoticError):   # 4 native bottoms are not compiled
        P.invert_stack(_cfg(mode=P.MODE_NATIVE, n_bottoms=4), *z)
    with pytest.raises(P.PhoticError):   # more than 8 dates
        P.invert_stack(_cfg(), np.zeros((9, 4, 8), np.float32), np.ones((9, 8), np.float32),
                       np.ones((9, 8), np.uint8))


def test_stack_mlp_seeded():
    """MLP-seeded stacks (the seeds of photic_mlp_seed_stack): bit-identical to the host
    engine given the same seeds, and what the trained network gains over plain stacks."""
    wl = (443.0, 482.0, 561.0, 655.0)
    cfg = P.cfg_default(wl, reqmin=1e-10, konvge=100, kcount=2000, check_watchdog=1)
    nd, n = 4, 8192
    rrs, ss, valid, truth = P.gen_stack(cfg, 5, n, n_dates=nd, k_min=2)
    dx, dt = P.trainset_device(cfg, 101, 1 << 21)
    net = P.Mlp(init_seed=1)
    net.train_device(dx.value, dt.value, 1 << 21, batch=8192, epochs=2, lr=2e-3, lr_final=1e-4)
    seed = net.seed_stack(cfg, rrs, ss, valid)
    g = P.invert_stack(cfg, rrs, ss, valid, seed=seed)
    p = P.invert_stack(cfg, rrs, ss, valid)
    m = 384
    h = O.hostsim_invert_stack(P.const_bytes(cfg, has_seed=True), rrs[:, :, :m], ss[:, :m],
                               valid[:, :m], seed=seed[:, :m], math="cr")
    assert np.array_equal(g["status"][:m], h["status"])
    assert np.array_equal(g["x"][:, :m], h["x"])
    assert np.array_equal(g["n_evals"][:m], h["n_evals"])
    v = g["status"] >= 0
    se, pe = g["n_evals"][v].mean(), p["n_evals"][v].mean()
    sm, pm = np.median(g["rrs_error"][v]), np.median(p["rrs_error"][v])
    print("stacks: seeded %.0f evals/stack, median E %.3g; plain %.0f, %.3g" % (se, sm, pe, pm))
    assert se < 0.8 * pe
    assert sm <= pm


@pytest.mark.parametrize("seeded", [False, True])
def test_stack_stream_matches_whole_scene(seeded):
    """photic_invert_stack_stream (host scene, 2-slot pinned staging, copy and compute
    streams, a ragged last tile) gives every pixel exactly what photic_invert_stack gives on
    the whole scene at once, and reports each tile as it lands."""
    wl = (443.0, 482.0, 561.0, 655.0)
    cfg = P.cfg_default(wl, reqmin=1e-10, konvge=100, kcount=2000)
    nd, n_rows, n_cols = 4, 37, 101
    n = n_rows * n_cols
    rrs, ss, valid, truth = P.gen_stack(cfg, 5, n, n_dates=nd, k_min=2)
    valid[:, 5:9] = 0
    net = None
    seed = None
    if seeded:
        dx, dt = P.trainset_device(cfg, 101, 1 << 18)
        net = P.Mlp(init_seed=1)
        net.train_device(dx.value, dt.value, 1 << 18, batch=8192, epochs=1, lr=2e-3,
                         lr_final=1e-4)
        seed = net.seed_stack(cfg, rrs, ss, valid)
    whole = P.invert_stack(cfg, rrs, ss, valid, seed=seed)
    seen = []
    st = P.invert_stack_stream(cfg, n_rows, n_cols, np.ascontiguousarray(rrs),
                               np.ascontiguousarray(ss), np.ascontiguousarray(valid),
                               tile_rows=8, net=net, progress=lambda d, t: seen.append((d, t)))
    assert seen == [(k, 5) for k in range(1, 6)]
    for k in whole:
        assert np.array_equal(st[k], whole[k]), k
    one = P.invert_stack_stream(cfg, n_rows, n_cols, np.ascontiguousarray(rrs),
                                np.ascontiguousarray(ss), np.ascontiguousarray(valid),
                                tile_rows=1000, net=net, fields=("depth", "status"))
    assert np.array_equal(one["depth"], whole["depth"])
    assert np.array_equal(one["status"], whole["status"])


@pytest.mark.parametrize("nbot", [0, 1, 2, 3])
def test_one_date_stack_is_the_single_date_inversion(nbot):
    """A stack of one date is the single-date problem, in the mix layout and with 1-3 native
    bottoms: the group engine (8-lane groups) and the single-date kernel (one pixel per lane)
    give identical results."""
    wl = (443.0, 482.0, 561.0, 655.0)
    kw = dict(mode=P.MODE_NATIVE, n_bottoms=nbot) if nbot else {}
    cfg = P.cfg_default(wl, reqmin=1e-10, konvge=100, kcount=2000, **kw)
    rrs, ss, truth, nodata, _ = O.gen(23 + nbot, 1024)
    g1 = P.invert(cfg, rrs, ss)
    gs = P.invert_stack(cfg, rrs[None], ss[None], np.ones((1, rrs.shape[1]), np.uint8))
    for k in ("status", "x", "n_evals", "n_iter", "n_starts", "depth", "rrs_error",
              "model_error", "albedo", "k_min", "iod", "b_pct", "bottom_type"):
        assert np.array_equal(gs[k], g1[k]), k


@pytest.mark.parametrize("nbot,own_h", [(2, False), (3, True), (1, False)])
def test_native_stack_against_oracle(nbot, own_h):
    """Native-layout stacks (the reference's own vector over the valid dates, per-date view
    angles, optionally one depth start per pixel) against the oracle's native multi-date
    inversion with the reference's Nelder-Mead settings: status identical, every converged
    stack within the bar (or with the same misfit where it is flat, or shown to differ by
    the host C library's rounding alone on at most 0.5 % of them), at least 98 %
    bit-identical."""
    nd = 4
    rrs, ss, valid, truth = O.gen_stack(61 + nbot, 384, n_dates=nd, k_min=2)
    rng = np.random.default_rng(nbot)
    sv = (1.0 / np.cos(np.radians(rng.uniform(0.0, 20.0, (nd, 1)))) * np.ones((1, rrs.shape[2]))).astype(np.float32)
    h = rng.uniform(1.0, 20.0, rrs.shape[2]).astype(np.float32) if own_h else None
    cfg = P.cfg_default(mode=P.MODE_NATIVE, n_bottoms=nbot, reqmin=1e-2, konvge=100,
                        kcount=5000, check_watchdog=1)
    g = P.invert_stack(cfg, rrs, ss, valid, sec_view=sv, hstart=h)
    o = O.invert_stack_native(rrs, ss, valid, n_bottoms=nbot, sec_view=sv, hstart_px=h)
    assert np.array_equal(g["status"], o["status"])
    v = o["status"] >= 0
    both = (g["status"] == 0) & (o["status"] == 0)
    dH = np.abs(g["x"][0] - o["x"][0])
    rel = np.abs(g["x"][1:] - o["x"][1:]) / np.maximum(np.abs(o["x"][1:]), 1e-300)
    within = (dH < 0.01) & np.all((rel < 1e-4) | (o["x"][1:] == 0), axis=0)
    flat = np.abs(g["rrs_error"] - o["rrs_error"]) <= 1e-6 * np.abs(o["rrs_error"])
    exact = np.all(g["x"] == o["x"], axis=0)
    bad = both & ~(within | flat)
    print("native %d bottoms%s: %d valid, %d converged, within %.4f, bit-identical %.4f, "
          "outside %d" % (nbot, " (own depth starts)" if own_h else "", v.sum(), both.sum(),
                          within[both].mean(), exact[v].mean(), bad.sum()))
    assert bad.sum() <= 0.005 * both.sum()
    assert exact[v].mean() > 0.98
    sel = v & exact
    for k in ("depth", "rrs_error", "k_min", "iod", "albedo", "bottom_type"):
        assert np.allclose(g[k][sel], o[k][sel], rtol=1e-6, atol=1e-7), k
    assert np.allclose(g["b_pct"][:, sel], o["b_pct"][:, sel], rtol=1e-6, atol=1e-6)


def _vs_oracle(g, o, label):
    """Status identical; every stack converged on both sides within the bar (H 1 cm, the other
    parameters 1e-4 relative) or with the same misfit where it is flat; the fraction
    bit-identical returned."""
    assert np.array_equal(g["status"], o["status"])
    v = o["status"] >= 0
    both = (o["status"] == 0) & (g["status"] == 0)
    dH = np.abs(g["x"][0] - o["x"][0])
    rel = np.abs(g["x"][1:] - o["x"][1:]) / np.maximum(np.abs(o["x"][1:]), 1e-300)
    within = (dH < 0.01) & np.all((rel < 1e-4) | (o["x"][1:] == 0), axis=0)
    flat = np.abs(g["rrs_error"] - o["rrs_error"]) <= 1e-6 * np.abs(o["rrs_error"])
    exact = np.all(g["x"] == o["x"], axis=0)
    print("%s: %d valid, %d converged, within %.4f, bit-identical %.4f" % (
        label, v.sum(), both.sum(), within[both].mean(), exact[v].mean()))
    assert np.all((within | flat)[both])
    return exact[v].mean()


def test_stack_eight_dates_against_oracle():
    """The 8-date stack workload (2048 x 2048 scene seed 3, Bernoulli(0.3) clouds, 3-8 valid
    dates, so 2 + 3k = 11..26 parameters) on a sample of its stacks, kernel against the FP64
    oracle directly: every date count from 3 to 8 present."""
    rrs, ss, valid, truth = O.gen_stack(3, 2048, n_dates=8, p_cloud=0.3, k_min=3)
    k = valid.sum(0)
    assert set(range(3, 9)) <= set(np.unique(k).tolist())
    g = P.invert_stack(_cfg(), rrs, ss, valid)
    o = O.invert_stack(rrs, ss, valid)
    exact = _vs_oracle(g, o, "8 dates, k %s" % np.bincount(k)[3:].tolist())
    assert exact >= 0.98
    for kk in (7, 8):   # the largest problems, N = 23 and 26
        sel = (k == kk) & (o["status"] >= 0)
        assert sel.sum() > 0 and np.all(g["status"][sel] == o["status"][sel])


def test_stack_mlp_seeded_stream_against_oracle():
    """The streamed MLP-seeded 4-date workload (scene seed 5, 2-4 valid dates, 2048-row tiles
    in the bench, here 8-row tiles over a 24 x 32 scene) against the FP64 oracle given the
    network's stack seeds: the stream's seeds are the whole scene's, so every stack must be
    what the oracle's seeded stack inversion gives for it."""
    wl = (443.0, 482.0, 561.0, 655.0)
    cfg = P.cfg_default(wl, reqmin=1e-10, konvge=100, kcount=2000)
    nd, n_rows, n_cols = 4, 24, 32
    n = n_rows * n_cols
    rrs, ss, valid, truth = P.gen_stack(cfg, 5, n, n_dates=nd, k_min=2)
    dx, dt = P.trainset_device(cfg, 101, 1 << 18)
    net = P.Mlp(init_seed=1)
    net.train_device(dx.value, dt.value, 1 << 18, batch=8192, epochs=1, lr=2e-3, lr_final=1e-4)
    seed = net.seed_stack(cfg, rrs, ss, valid)
    st = P.invert_stack_stream(cfg, n_rows, n_cols, np.ascontiguousarray(rrs),
                               np.ascontiguousarray(ss), np.ascontiguousarray(valid),
                               tile_rows=8, net=net)
    o = O.invert_stack(rrs, ss, valid, seed=seed)
    exact = _vs_oracle(st, o, "seeded stream, 4 dates")
    assert exact >= 0.98
    assert np.all(st["n_starts"][o["status"] >= 0] >= 1)
