"""Benchmark: pixels/sec of photic's per-pixel inversion on one MI355X.

Workload: the 8192 x 8192 synthetic single-date scene (4 bands at 443/482/561/655 nm,
Philox-seeded parameters drawn from the plain-method distributions, sun zenith 20-50 deg,
Gaussian noise 2e-4 1/sr), Nelder-Mead over the five parameters (H, B, P, G, X) with
reqmin 1e-10, konvge 100, kcount 2000 and the reference's depth starts with early exit.
A step is one inversion of the whole scene with inputs already resident in HBM.  With the
warm-start network available the headline is the MLP-seeded inversion (network inference +
seeded Nelder-Mead + plain fallback); beside it, on the same scene, the plain GPU Nelder-Mead
with the reference's early exit, and the plain method at equal median misfit (its early exit
lowered step by step until its median misfit is at or below the seeded one, on 2^22 pixels).

Prints one JSON line.  Besides the contract's fields it carries the roofline of the dominant
kernel (algorithmic FLOPs k(46 n_b + 15) + 6 = 205 per objective evaluation over the kernel's
duration from HIP events on its stream; HBM bytes per launch from the profiler counters in
profiles/) and the CPU baseline (the FP64 oracle, OpenMP over pixels, timed on a bounded
sample of the same scene on this host's cores).

Two further workloads, each with its own JSON line (not the headline):
  --workload stacks  2048 x 2048 pixels seen on 8 dates, 3-8 valid per pixel (Bernoulli(0.3)
                     clouds): plain stack inversion (2 + 3k parameters), inputs resident in HBM,
                     launch shape from the stack autotuner.
  --workload stream  12662 x 16077 pixels on 4 dates, 2-4 valid: MLP-seeded stack inversion of
                     a host-resident scene streamed through HBM in 2048-row tiles (pinned
                     double-buffered copies on a side stream); wall time includes the copies.
  --workload regions 512 x 512 pixels of a spatially coherent scene (smooth parameter fields,
                     one sun zenith), each pixel inverted with its 3 x 3 neighbourhood: the
                     reference's default n_spatial = 2 with 3 bottoms and its own Nelder-Mead
                     settings (reqmin 1e-2, konvge 100, kcount 5000), inputs resident in HBM.
  --workload ordered the headline scene through the ordered inversion (the reference's
                     spectral-index order, look-up table and hot start recast as runs with
                     leaders), beside the MLP-seeded and the plain inversion of the same scene.
  --workload dropin  4096 x 4096 pixels of the same kind of coherent scene through the drop-in
                     samodel() with host row-pointer grids in and out, as the reference's caller
                     hands them over (one scene, n_spatial 1, 2 bottoms, the reference's
                     settings): wall time includes the host gather, both PCIe copies and the
                     scatter into the output grids; the device-resident rate of the same
                     inversion is reported beside it.
"""
import argparse
import ctypes as C
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "pixels/sec inverted, MLP-seeded vs plain NM at equal misfit, 1 MI355X"
FLOP_PER_EVAL = 1 * (46 * 4 + 15) + 6  # 205
PEAK_FP64_TFLOPS = 78.6
WL = (443.0, 482.0, 561.0, 655.0)
# Retrieved depths beyond this are runaways: where exp(-K H) underflows the misfit is flat in
# H and the simplex may expand H without bound (the generating H is at most 25 m).  Reported
# as a fraction, with the RMSE over the other pixels beside the plain RMSE.
RUNAWAY_M = 100.0


def _events():
    import photic_amd as P
    e0, e1 = C.c_void_p(), C.c_void_p()
    P.check(P.lib.photic_event_create(C.byref(e0)))
    P.check(P.lib.photic_event_create(C.byref(e1)))
    return e0, e1


def timed_invert(P, cfg, n, d_rrs, d_ss, d_seed, outs, stream):
    """One photic_invert on `stream`, bracketed by HIP events on that stream; returns ms."""
    e0, e1 = _events()
    P.check(P.lib.photic_event_record(e0, stream))
    P.invert_device(cfg, n, d_rrs, d_ss, d_seed, outs, stream=stream)
    P.check(P.lib.photic_event_record(e1, stream))
    ms = C.c_float()
    P.check(P.lib.photic_event_elapsed_ms(e0, e1, C.byref(ms)))
    P.lib.photic_event_destroy(e0)
    P.lib.photic_event_destroy(e1)
    return float(ms.value)


SKIP_FIT_SCENE = 14  # Philox key of the scene the seed-skip depth is fitted on (the timed one is 4)


def fit_seed_skip(P, net, cfg, stream, fixed):
    """The depth below which the network's seed is not tried, learned from labelled forward-
    model pixels of a scene other than the timed one: both forms are run on 2^20 pixels of it,
    which gives each pixel's evaluations with and without its seed start, and the depth chosen
    is the one whose skipped pixels (network depth below it) save the most evaluations in
    total.  Shallow clear water cannot meet the early exit under the K penalty, so it runs
    every depth start whatever its seed; there the seed start is pure cost.  fixed > 0: that
    depth, no fit.  Sets the network's skip depth and returns what was fitted."""
    if fixed > 0.0:
        net.set_seed_skip(fixed)
        return {"depth_m": fixed, "fitted": False}
    n_c = 1 << 20
    d_r, d_s, d_t = P.DeviceBuffer(16 * n_c), P.DeviceBuffer(4 * n_c), P.DeviceBuffer(24 * n_c)
    P.check(P.lib.photic_gen_scene(C.byref(cfg), SKIP_FIT_SCENE, 0, n_c, 2e-4, 0.0, d_r.value,
                                   d_s.value, d_t.value, stream))
    net.set_seed_skip(0.0)
    d_seed = P.DeviceBuffer(20 * n_c)
    net.seed_device(cfg, n_c, d_r.value, d_s.value, d_seed.value, stream=stream)
    ev = {}
    for kind in ("seeded", "plain"):
        outs = P.DeviceOutputs(cfg, n_c, fields=("status", "n_evals"))
        P.invert_device(cfg, n_c, d_r.value, d_s.value,
                        d_seed.value if kind == "seeded" else None, outs, stream=stream)
        # the launches run on a non-blocking stream: wait for them before the copies back
        P.check(P.lib.photic_synchronize())
        ev[kind] = outs.fetch()
    h_pred = d_seed.to_array(np.float32, (5, n_c))[0]
    v = (ev["seeded"]["status"] >= 0) & (ev["plain"]["status"] >= 0)
    gain = (ev["seeded"]["n_evals"].astype(np.int64) - ev["plain"]["n_evals"])[v]
    hv = h_pred[v]
    grid = np.arange(1.0, 8.01, 0.25)
    saved = [int(gain[hv < h].sum()) for h in grid]
    best = int(np.argmax(saved))
    h = float(grid[best]) if saved[best] > 0 else 0.0
    net.set_seed_skip(h)
    return {"depth_m": h, "fitted": True, "fit_scene_seed": SKIP_FIT_SCENE, "fit_pixels": n_c,
            "evals_saved_per_pixel_on_fit_scene": saved[best] / max(int(v.sum()), 1) if h else 0.0,
            "skipped_frac_on_fit_scene": float((hv < h).mean()) if h else 0.0}


def progress(msg):
    """One line on stderr per finished phase (the JSON line alone goes to stdout)."""
    print("bench: " + msg, file=sys.stderr, flush=True)


def host_cpus():
    """The host's CPU facts and the thread count the CPU baseline uses: one thread per core
    this process may run on (its affinity mask), capped by OMP_NUM_THREADS where the
    environment sets it (the CPU share that a one-GPU job is given)."""
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    threads = aff or 1
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and int(omp) > 0:
        threads = min(threads, int(omp))
    info = {"nproc": os.cpu_count(), "affinity_cpus": aff, "omp_num_threads": omp or None,
            "threads_used": threads, "model": None, "threads_per_core": None}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for ln in out.splitlines():
            k, _, v = ln.partition(":")
            if k.strip() == "Model name":
                info["model"] = v.strip()
            elif k.strip() == "Thread(s) per core":
                info["threads_per_core"] = int(v.strip())
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    return threads, info


def cpu_baseline(rrs_h, ss_h, n_sample, seconds_hint):
    """The FP64 oracle (OpenMP over pixels, the reference's dynamic schedule) on the first
    n_sample pixels of the same scene; returns px/s and the threads used."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_bindings as O
    threads, _ = host_cpus()
    used = O.set_threads(threads)
    t = time.perf_counter()
    O.invert(rrs_h[:, :n_sample], ss_h[:n_sample], reqmin=1e-10, konvge=100, kcount=2000)
    dt = time.perf_counter() - t
    return n_sample / dt, used, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--side", type=int, default=8192, help="scene is side x side pixels")
    ap.add_argument("--precision", choices=("fp64", "fp32"), default="fp64")
    ap.add_argument("--cpu-sample", type=int, default=524288)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run one plain device copy of BYTES (the known-traffic launch that "
                         "calibrates the profiler's HBM byte counters) and exit")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--plain-only", action="store_true", help="skip the warm-start network")
    ap.add_argument("--no-equal-misfit", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shape")
    ap.add_argument("--equal-misfit-pixels", type=int, default=1 << 22)
    ap.add_argument("--train-samples", type=int, default=1 << 24)
    ap.add_argument("--train-epochs", type=int, default=2)
    ap.add_argument("--seed-skip", type=float, default=-1.0,
                    help="seeded runs: no seed start where the network predicts a depth below "
                         "this; -1 (default): fitted on a separate labelled scene; 0: none")
    ap.add_argument("--seed-abort", type=float, default=5.0,
                    help="seeded runs: give a seed start up while its objective is above this "
                         "(the early exit stops below 2.5) at its first convergence check "
                         "(0: never)")
    ap.add_argument("--workload", choices=("scene", "stacks", "stream", "regions", "dropin",
                                           "ordered"),
                    default="scene")
    ap.add_argument("--dropin-side", type=int, default=4096)
    ap.add_argument("--ordered-cpu-sample", type=int, default=1 << 20)
    ap.add_argument("--dropin-bottoms", type=int, default=2)
    ap.add_argument("--dropin-spatial", type=int, default=1,
                    help="the drop-in's n_spatial (the reference's default is 2)")
    ap.add_argument("--region-side", type=int, default=512)
    ap.add_argument("--region-spatial", type=int, default=2)
    ap.add_argument("--region-bottoms", type=int, default=3)
    ap.add_argument("--region-cpu-sample", type=int, default=4096)
    ap.add_argument("--stack-side", type=int, default=2048)
    ap.add_argument("--stack-cpu-sample", type=int, default=32768)
    ap.add_argument("--stream-cpu-sample", type=int, default=65536)
    ap.add_argument("--stream-rows", type=int, default=12662)
    ap.add_argument("--stream-cols", type=int, default=16077)
    ap.add_argument("--stream-tile-rows", type=int, default=2048)
    ap.add_argument("--stream-plain-rows", type=int, default=0,
                    help="also stream the plain method over the first ROWS rows")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the inversion has no cross-GPU step (pixels are "
              "independent, so a scene shards by row tiles over GPUs with nothing exchanged "
              "but the summary counts). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import photic_amd as P
    if args.workload == "stacks":
        return bench_stacks(args, P)
    if args.workload == "stream":
        return bench_stream(args, P)
    if args.workload == "regions":
        return bench_regions(args, P)
    if args.workload == "dropin":
        return bench_dropin(args, P)
    if args.workload == "ordered":
        return bench_ordered(args, P)
    if args.copy_calibration:
        nbytes = args.copy_calibration
        src, dst = P.DeviceBuffer(nbytes), P.DeviceBuffer(nbytes)
        P.check(P.lib.photic_memset_zero(src.value, nbytes))
        P.check(P.lib.photic_device_copy(dst.value, src.value, nbytes, None))
        P.check(P.lib.photic_synchronize())
        print(json.dumps({"copy_bytes": nbytes}))
        return
    prec = P.FP64 if args.precision == "fp64" else P.FP32
    n = args.side * args.side
    cfg = P.cfg_default(WL, reqmin=1e-10, konvge=100, kcount=2000, precision=prec,
                        watchdog_s=900.0, seed_abort_misfit=args.seed_abort)

    stream = C.c_void_p()
    P.check(P.lib.photic_stream_create(C.byref(stream)))

    # ---- the scene, generated on the GPU (Philox keyed by pixel index) ----
    d_rrs = P.DeviceBuffer(4 * 4 * n)
    d_ss = P.DeviceBuffer(4 * n)
    d_truth = P.DeviceBuffer(4 * 6 * n)
    P.check(P.lib.photic_gen_scene(C.byref(cfg), 4, 0, n, 2e-4, 0.0, d_rrs.value, d_ss.value,
                                   d_truth.value, stream))
    P.check(P.lib.photic_synchronize())
    truth_H = d_truth.to_array(np.float32, (6, n))[0]
    del d_truth
    fields = ("depth", "rrs_error", "model_error", "status", "n_evals", "n_starts")

    progress("scene of %d pixels generated" % n)
    # ---- the warm-start network, trained once on forward-model samples (not timed) ----
    net = None
    train_info = None
    if not args.plain_only:
        t = time.perf_counter()
        n_tr = args.train_samples
        dx, dt = P.trainset_device(cfg, 101, n_tr)
        net = P.Mlp(init_seed=1)
        loss = net.train_device(dx.value, dt.value, n_tr, batch=8192, epochs=args.train_epochs,
                                lr=2e-3, lr_final=1e-4, stream=stream)
        P.check(P.lib.photic_synchronize())
        train_info = {"samples": n_tr, "epochs": args.train_epochs, "final_loss": loss,
                      "seconds": time.perf_counter() - t}
        del dx, dt
        if args.seed_skip != 0.0:
            train_info["seed_skip"] = fit_seed_skip(P, net, cfg, stream, args.seed_skip)
    d_seed = P.DeviceBuffer(4 * 5 * n) if net is not None else None

    def scene_slice(n_sub):
        """The first n_sub pixels as their own planar arrays ([4][n_sub], [n_sub])."""
        if n_sub == n:
            return d_rrs, d_ss
        r = P.DeviceBuffer(4 * 4 * n_sub)
        sv = P.DeviceBuffer(4 * n_sub)
        for b in range(4):
            P.check(P.lib.photic_device_copy(r.value + 4 * b * n_sub, d_rrs.value + 4 * b * n,
                                             4 * n_sub, stream))
        P.check(P.lib.photic_device_copy(sv.value, d_ss.value, 4 * n_sub, stream))
        P.check(P.lib.photic_synchronize())
        return r, sv

    host_out = {}  # page-locked output arrays per pixel count, reused by every step

    def run(kind, cfg_run, n_run):
        """One step of `kind` over the first n_run pixels -- network inference when seeded, the
        inversion, and the outputs copied back to page-locked host arrays; returns (wall ms,
        kernel ms, outputs)."""
        r, sv = scene_slice(n_run)
        outs = P.DeviceOutputs(cfg_run, n_run, fields=fields)
        if n_run not in host_out:
            host_out[n_run] = outs.pinned()
        P.check(P.lib.photic_synchronize())
        t0 = time.perf_counter()
        if kind == "seeded":
            net.seed_device(cfg_run, n_run, r.value, sv.value, d_seed.value, stream=stream)
        k_ms = timed_invert(P, cfg_run, n_run, r.value, sv.value,
                            d_seed.value if kind == "seeded" else None, outs, stream)
        res = outs.fetch(into=host_out[n_run])
        wall = 1e3 * (time.perf_counter() - t0)
        # the next step reuses the page-locked arrays: the caller gets its own copies
        return wall, k_ms, {k: v.copy() for k, v in res.items()}

    def summary(ms, res, n_run):
        v = res["status"] >= 0
        dH = res["depth"][v].astype(np.float64) - truth_H[:n_run][v]
        fit = res["rrs_error"][v] < 1.0  # pixels the model fits to 1% (E_Rrs)
        return {
            "pixels_per_s": n_run / (ms / 1e3),
            "ms": ms,
            "evals_per_pixel": float(res["n_evals"][v].mean()),
            "starts_per_pixel": float(res["n_starts"][v].mean()),
            "converged_frac": float((res["status"][v] == 0).mean()),
            "median_rrs_error_pct": float(np.median(res["rrs_error"][v])),
            "p90_rrs_error_pct": float(np.percentile(res["rrs_error"][v], 90)),
            "depth_rmse_m": float(np.sqrt(np.mean(dH * dH))),
            "depth_median_abs_error_m": float(np.median(np.abs(dH))),
            "depth_rmse_fitted_m": float(np.sqrt(np.mean(dH[fit] ** 2))),
            "fitted_frac": float(fit.mean()),
            "depth_runaway_frac": float((res["depth"][v] > RUNAWAY_M).mean()),
            "depth_rmse_bounded_m": float(np.sqrt(np.mean(dH[res["depth"][v] <= RUNAWAY_M] ** 2))),
            "valid_pixels": int(v.sum()),
        }

    if train_info:
        progress("network trained, seed-skip depth %s" % train_info.get("seed_skip"))
    # ---- launch shapes: the autotuner times every candidate on the first 2^21 pixels and
    # keeps the fastest for the seeded and for the plain kernel (untimed) ----
    head = "plain" if net is None else "seeded"
    tuning = {}
    if not args.no_autotune:
        n_t = min(n, 1 << 21)
        r_t, s_t = scene_slice(n_t)
        if net is not None:
            d_st = P.DeviceBuffer(4 * 5 * n_t)
            net.seed_device(cfg, n_t, r_t.value, s_t.value, d_st.value, stream=stream)
            P.check(P.lib.photic_synchronize())
            tuning["seeded"] = P.autotune(cfg, n_t, r_t.value, s_t.value, d_st.value, stream)
            del d_st
        tuning["plain"] = P.autotune(cfg, n_t, r_t.value, s_t.value, None, stream)

    progress("launch shapes chosen")
    # ---- warm-up: W whole-scene steps of the headline method, untimed ----
    for _ in range(args.warmup):
        run(head, cfg, n)

    # ---- timed: K whole-scene steps ----
    walls, kms = [], []
    res = None
    for _ in range(args.steps):
        w_ms, k_ms, res = run(head, cfg, n)
        walls.append(w_ms)
        kms.append(k_ms)
    ms_per_step = float(np.mean(walls))
    head_sum = summary(ms_per_step, res, n)
    evals = float(res["n_evals"][res["status"] >= 0].astype(np.int64).sum())
    k_ms = float(np.mean(kms))
    px_s = n / (ms_per_step / 1e3)

    progress("%d timed steps, %.1f ms each" % (args.steps, ms_per_step))
    # ---- the plain GPU Nelder-Mead on the same scene, beside it ----
    comparisons = {}
    if net is not None:
        w_ms, _, pres = run("plain", cfg, n)
        comparisons["plain"] = summary(w_ms, pres, n)
        # The plain method as the arbiter: how many pixels end with a seeded misfit above the
        # plain method's.
        # A seeded pixel falls back to the reference's depth starts whenever its seed start
        # stays at or above the early exit, so it can end worse than the plain method only
        # where both end below the early exit (the plain method accepts any start below it).
        both = (res["status"] >= 0) & (pres["status"] >= 0)
        worse = both & (res["model_error"] > pres["model_error"])
        vs_plain = {"pixels": int(both.sum()),
                 "seeded_worse_than_plain_frac": float(worse.sum() / max(both.sum(), 1)),
                 "worse_with_seeded_at_or_above_early_exit": int(
                     (worse & (res["model_error"] >= 2.5)).sum()),
                 "median_misfit_of_worse_seeded": float(np.median(res["rrs_error"][worse]))
                 if worse.any() else None,
                 "median_misfit_of_worse_plain": float(np.median(pres["rrs_error"][worse]))
                 if worse.any() else None}
        comparisons["seeded_vs_plain"] = vs_plain
        if not args.no_equal_misfit:
            # The plain method at equal misfit: the reference's depth starts with the early
            # exit lowered until the plain median misfit is at or below the seeded one, chosen
            # on the first n_eq pixels of the scene (early exit 1e-300 = every start run, best
            # kept) and then timed on the whole scene.
            n_eq = min(n, args.equal_misfit_pixels)
            v_eq = res["status"][:n_eq] >= 0
            seeded_med = float(np.median(res["rrs_error"][:n_eq][v_eq]))
            sweep = []
            for ee in (1.0, 0.1, 0.01, 1e-3, 1e-4, 1e-300):
                cfg_eq = P.cfg_default(WL, reqmin=1e-10, konvge=100, kcount=2000,
                                       precision=prec, watchdog_s=900.0, early_exit=ee)
                w_ms, _, eres = run("plain", cfg_eq, n_eq)
                s_eq = summary(w_ms, eres, n_eq)
                s_eq["early_exit"] = ee
                sweep.append(s_eq)
                if s_eq["median_rrs_error_pct"] <= seeded_med:
                    break
            # the chosen early exit, timed on the whole scene like the seeded step
            ee_eq = sweep[-1]["early_exit"]
            cfg_eq = P.cfg_default(WL, reqmin=1e-10, konvge=100, kcount=2000, precision=prec,
                                   watchdog_s=900.0, early_exit=ee_eq)
            w_ms, _, eres = run("plain", cfg_eq, n)
            full = summary(w_ms, eres, n)
            both_eq = (res["status"] >= 0) & (eres["status"] >= 0)
            vs_plain["seeded_worse_than_plain_equal_misfit_frac"] = float(
                (both_eq & (res["model_error"] > eres["model_error"])).sum() /
                max(both_eq.sum(), 1))
            v_all = res["status"] >= 0
            comparisons["plain_equal_misfit"] = dict(
                full, early_exit=ee_eq, pixels=n,
                seeded_median_rrs_error_pct=float(np.median(res["rrs_error"][v_all])),
                chosen_on_pixels=n_eq, seeded_median_rrs_error_pct_on_those=seeded_med)
            progress("plain method at equal misfit: early exit %g" % ee_eq)
            comparisons["plain_equal_misfit_sweep"] = [
                {k: e[k] for k in ("early_exit", "pixels_per_s", "median_rrs_error_pct",
                                   "evals_per_pixel")} for e in sweep]
    achieved = evals * FLOP_PER_EVAL / (k_ms / 1e3) / 1e12
    # HBM bytes per launch: the per-pixel FETCH_SIZE / WRITE_SIZE of the same kernel from
    # rocprofv3 --pmc, each calibrated by a copy kernel of known size (profiles/)
    traffic = None
    tj = os.path.join(ROOT, "profiles", "invert_fp64_traffic.json")
    if prec == P.FP64 and os.path.exists(tj):
        with open(tj) as f:
            t = json.load(f).get(head)
        if t:
            traffic = (t["fetch_bytes_per_pixel"] + t["write_bytes_per_pixel"]) * n

    line = {
        "metric": METRIC,
        "value": px_s,
        "unit": "pixels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == P.FP64 else "f64+f32-transcendentals",
        "data": "synthetic (Philox-seeded scene and training samples, no real imagery)",
        "config": {
            "workload": "%s inversion, %dx%d single-date scene, 4 bands, 5 params (H, B, P, G, X), "
                        "reqmin 1e-10 kcount 2000 konvge 100" % (
                            "MLP-seeded Nelder-Mead with plain fallback" if head == "seeded"
                            else "plain GPU Nelder-Mead", args.side, args.side),
            "pixels": n, "seed": 4, "noise_sigma": 2e-4,
            "seed_abort_misfit": args.seed_abort if head == "seeded" else None,
            "parallelism": "1 GPU, persistent work-queue kernel, autotuned launch shape",
        },
        head: head_sum,
        "roofline": {
            "bound": "valu",
            "achieved": achieved,
            "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP64_TFLOPS,
            "traffic": traffic,
            "kernel": "ph_invert_kernel<%s, 4, 0, %s>" % (
                "double" if prec == P.FP64 else "float", "true" if head == "seeded" else "false"),
            "kernel_ms": k_ms,
            "flop_per_eval": FLOP_PER_EVAL,
            "evals_per_launch": evals,
        },
        "cpu_baseline": None,
        "launch": tuning or "default",
    }
    if train_info:
        line["mlp_training"] = train_info
    line.update(comparisons)
    if "plain" in comparisons:
        line["seeded_over_plain"] = px_s / comparisons["plain"]["pixels_per_s"]
    if "plain_equal_misfit" in comparisons:
        line["seeded_over_plain_equal_misfit"] = (
            px_s / comparisons["plain_equal_misfit"]["pixels_per_s"])

    if not args.no_cpu:
        ns = min(args.cpu_sample, n)
        rrs_h = d_rrs.to_array(np.float32, (4, n))[:, :ns].copy()
        ss_h = d_ss.to_array(np.float32, (n,))[:ns].copy()
        progress("CPU baseline on %d pixels" % ns)
        cpx, threads, dt = cpu_baseline(rrs_h, ss_h, ns, 20)
        line["cpu_baseline"] = {
            "value": cpx, "unit": "pixels/s", "cores": threads, "kind": "port",
            "sample": "first %d pixels of the same scene, FP64 oracle, OpenMP dynamic "
                      "schedule, %.1f s" % (ns, dt),
        }
        line["cpu_baseline"]["host"] = host_cpus()[1]
        line["gpu_over_cpu"] = px_s / cpx
        if "plain" in line:
            line["plain_gpu_over_cpu"] = line["plain"]["pixels_per_s"] / cpx

    if net is not None:
        net.close()
    P.lib.photic_stream_destroy(stream)
    print(json.dumps(line))



def _stack_flops(valid, n_evals):
    """Algorithmic FLOPs of a stack inversion: evaluations x (k (46 n_b + 15) + 6) per stack,
    k its valid dates, n_b = 4 bands."""
    k = valid.astype(np.int64).sum(0)
    return float((n_evals.astype(np.int64) * (k * (46 * 4 + 15) + 6)).sum())


def _stack_cpu(rrs_h, ss_h, valid_h, n_sample, seed=None):
    """The FP64 oracle's stack inversion (OpenMP over stacks) on the first n_sample stacks."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_bindings as O
    threads, _ = host_cpus()
    used = O.set_threads(threads)
    t = time.perf_counter()
    O.invert_stack(np.ascontiguousarray(rrs_h[:, :, :n_sample]),
                   np.ascontiguousarray(ss_h[:, :n_sample]),
                   np.ascontiguousarray(valid_h[:, :n_sample]))
    dt = time.perf_counter() - t
    return n_sample / dt, used, dt


def bench_stacks(args, P):
    """Plain stack inversion of a side x side x 8-date scene resident in HBM."""
    nd, side = 8, args.stack_side
    n = side * side
    cfg = P.cfg_default(WL, reqmin=1e-10, konvge=100, kcount=2000, watchdog_s=900.0)
    stream = C.c_void_p()
    P.check(P.lib.photic_stream_create(C.byref(stream)))
    d_rrs, d_ss, d_v = P.DeviceBuffer(16 * nd * n), P.DeviceBuffer(4 * nd * n), P.DeviceBuffer(nd * n)
    d_tr = P.DeviceBuffer(4 * (2 + 4 * nd) * n)
    P.check(P.lib.photic_gen_stack(C.byref(cfg), 3, 0, n, nd, 2e-4, 0.3, 3, d_rrs.value,
                                   d_ss.value, d_v.value, d_tr.value, stream))
    P.check(P.lib.photic_synchronize())
    truth_H = d_tr.to_array(np.float32, (2 + 4 * nd, n))[0]
    del d_tr
    valid_h = d_v.to_array(np.uint8, (nd, n))

    # autotuned launch shape on the first n_t stacks (untimed)
    n_t = min(n, 1 << 20)
    sub = [P.DeviceBuffer(16 * nd * n_t), P.DeviceBuffer(4 * nd * n_t), P.DeviceBuffer(nd * n_t)]
    for q in range(4 * nd):
        P.check(P.lib.photic_device_copy(sub[0].value + 4 * q * n_t, d_rrs.value + 4 * q * n,
                                         4 * n_t, stream))
    for d in range(nd):
        P.check(P.lib.photic_device_copy(sub[1].value + 4 * d * n_t, d_ss.value + 4 * d * n,
                                         4 * n_t, stream))
        P.check(P.lib.photic_device_copy(sub[2].value + d * n_t, d_v.value + d * n, n_t, stream))
    P.check(P.lib.photic_synchronize())
    tuning = P.autotune_stack(cfg, n_t, nd, sub[0].value, sub[1].value, sub[2].value, None, stream)
    del sub

    fields = ("depth", "rrs_error", "status", "n_evals", "n_starts")

    def step():
        outs = P.StackOutputs(cfg, n, nd, fields=fields)
        P.check(P.lib.photic_synchronize())
        e0, e1 = _events()
        t0 = time.perf_counter()
        P.check(P.lib.photic_event_record(e0, stream))
        P.invert_stack_device(cfg, n, nd, d_rrs.value, d_ss.value, d_v.value, None, outs,
                              stream=stream)
        P.check(P.lib.photic_event_record(e1, stream))
        P.check(P.lib.photic_synchronize())
        wall = 1e3 * (time.perf_counter() - t0)
        ms = C.c_float()
        P.check(P.lib.photic_event_elapsed_ms(e0, e1, C.byref(ms)))
        P.lib.photic_event_destroy(e0)
        P.lib.photic_event_destroy(e1)
        return wall, float(ms.value), outs.fetch()

    for _ in range(args.warmup):
        step()
    walls, kms, res = [], [], None
    for _ in range(args.steps):
        w, k_ms, res = step()
        walls.append(w)
        kms.append(k_ms)
    ms_step, k_ms = float(np.mean(walls)), float(np.mean(kms))
    v = res["status"] >= 0
    dH = res["depth"][v].astype(np.float64) - truth_H[v]
    flops = _stack_flops(valid_h, res["n_evals"])
    achieved = flops / (k_ms / 1e3) / 1e12
    line = {
        "metric": "pixels/sec inverted, plain multi-date stack Nelder-Mead, 1 MI355X",
        "value": n / (ms_step / 1e3), "unit": "pixels/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (Philox-seeded stack scene, no real imagery)",
        "config": {"workload": "plain stack inversion, %dx%d pixels x 8 dates, 3-8 valid "
                               "(Bernoulli(0.3) clouds), 2+3k params, reqmin 1e-10 kcount 2000 "
                               "konvge 100" % (side, side),
                   "pixels": n, "seed": 3, "noise_sigma": 2e-4,
                   "parallelism": "1 GPU, one persistent work-queue launch per valid-date "
                                  "count, 8/16/32-lane groups per stack"},
        "stacks": {"evals_per_stack": float(res["n_evals"][v].mean()),
                   "starts_per_stack": float(res["n_starts"][v].mean()),
                   "median_rrs_error_pct": float(np.median(res["rrs_error"][v])),
                   "depth_rmse_m": float(np.sqrt(np.mean(dH * dH))),
                   "depth_median_abs_error_m": float(np.median(np.abs(dH))),
                   "depth_runaway_frac": float((res["depth"][v] > RUNAWAY_M).mean()),
                   "depth_rmse_bounded_m": float(np.sqrt(np.mean(
                       dH[res["depth"][v] <= RUNAWAY_M] ** 2))),
                   "valid_date_histogram": np.bincount(valid_h.sum(0), minlength=nd + 1).tolist()},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP64_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS, "traffic": None,
                     "kernel": "ph_stack_kernel<k, false>, k = 1..8 (one launch each)",
                     "kernel_ms": k_ms,
                     "flop_per_eval": "k (46 x 4 + 15) + 6 per stack of k valid dates",
                     "flop_per_launch": flops},
        "cpu_baseline": None, "launch": tuning,
    }
    if not args.no_cpu:
        ns = min(args.stack_cpu_sample, n)
        cpx, threads, dt = _stack_cpu(d_rrs.to_array(np.float32, (nd, 4, n))[:, :, :ns],
                                      d_ss.to_array(np.float32, (nd, n))[:, :ns],
                                      valid_h[:, :ns], ns)
        line["cpu_baseline"] = {"value": cpx, "unit": "pixels/s", "cores": threads,
                                "kind": "port",
                                "sample": "first %d stacks of the same scene, FP64 oracle, "
                                          "OpenMP dynamic schedule, %.1f s" % (ns, dt)}
        line["cpu_baseline"]["host"] = host_cpus()[1]
        line["gpu_over_cpu"] = line["value"] / cpx
    P.lib.photic_stream_destroy(stream)
    print(json.dumps(line))


def _region_flops(n_regions, n_evals, nd=1):
    """Algorithmic FLOPs of a region inversion: evaluations x (R nd 46 n_b + 15 nd + 6) per
    pixel of R regions, n_b = 4 bands (the band evaluations and the per-scene terms; the
    continuity penalties' O(R n_bottoms) comparisons are not counted, like the simplex's
    bookkeeping)."""
    R = n_regions.astype(np.int64)
    return float((n_evals.astype(np.int64) * (R * nd * 46 * 4 + 15 * nd + 6)).sum())


def region_kernel_name(nbot, n_max):
    """The instance the library launches for a one-scene problem of n_max parameters: blocks in
    LDS up to 64, simplexes in device memory at three or two waves per SIMD up to 256 (two or four
    coordinate slots; a one-scene problem's other arrays fit in LDS, the last template
    argument 1), then six slots.  PHOTIC_REGION_DEVICE_BLOCKS, the library's placement switch,
    is followed: 0 keeps blocks of up to 128 parameters in LDS, 1 puts every simplex in device
    memory, 2 every whole block."""
    force = os.environ.get("PHOTIC_REGION_DEVICE_BLOCKS", "")
    force = int(force) if force.lstrip("-").isdigit() else -1
    mc = (n_max + 63) // 64
    # one-slot blocks stay in LDS where 12 fit a CU: for one scene, up to about 40 parameters
    if mc == 1 and (force == 0 or (force not in (1, 2) and n_max <= 40)):
        return "ph_region_kernel_lds3<%d>" % nbot
    if mc == 2 and force == 0:
        return "ph_region_kernel<%d, 2, false, 0>" % nbot
    tl = 2 if force == 2 else 1
    if mc <= 4:
        return "ph_region_kernel_dm2<%d, %d, %d>" % (nbot, 2 if mc <= 2 else 4, tl)
    return "ph_region_kernel<%d, 6, true, %d>" % (nbot, tl)


def bench_regions(args, P):
    """Spatial regions: every pixel of a side x side coherent scene with its neighbourhood."""
    side, ns, nbot = args.region_side, args.region_spatial, args.region_bottoms
    n = side * side
    theta = 35.0
    sec_sun = 1.0 / np.cos(np.radians(theta))
    cfg = P.cfg_default(WL, mode=P.MODE_NATIVE, n_bottoms=nbot, reqmin=1e-2, konvge=100,
                        kcount=5000, watchdog_s=900.0)
    stream = C.c_void_p()
    P.check(P.lib.photic_stream_create(C.byref(stream)))
    d_rrs, d_tr = P.DeviceBuffer(16 * n), P.DeviceBuffer(24 * n)
    P.check(P.lib.photic_gen_field(C.byref(cfg), 1, side, side, theta, 2e-4, d_rrs.value,
                                   d_tr.value, stream))
    P.check(P.lib.photic_synchronize())
    truth_H = d_tr.to_array(np.float32, (6, n))[0]
    del d_tr
    fields = ("depth", "rrs_error", "status", "n_evals", "n_starts", "bottom_type")

    def step():
        outs = P.RegionOutputs(cfg, n, 1, ns, fields=fields)
        P.check(P.lib.photic_synchronize())
        e0, e1 = _events()
        t0 = time.perf_counter()
        P.check(P.lib.photic_event_record(e0, stream))
        P.invert_regions_device(cfg, side, side, 1, d_rrs.value, -9999.0, sec_sun, 1.0, ns, n,
                                None, None, outs, stream=stream)
        P.check(P.lib.photic_event_record(e1, stream))
        P.check(P.lib.photic_synchronize())
        wall = 1e3 * (time.perf_counter() - t0)
        ms = C.c_float()
        P.check(P.lib.photic_event_elapsed_ms(e0, e1, C.byref(ms)))
        P.lib.photic_event_destroy(e0)
        P.lib.photic_event_destroy(e1)
        return wall, float(ms.value), outs.fetch()

    for _ in range(args.warmup):
        step()
    walls, kms, res = [], [], None
    for _ in range(args.steps):
        w, k_ms, res = step()
        walls.append(w)
        kms.append(k_ms)
        print("regions step: %.1f ms" % w, file=sys.stderr, flush=True)
    ms_step, k_ms = float(np.mean(walls)), float(np.mean(kms))
    v = res["status"] >= 0
    dH = res["depth"][v].astype(np.float64) - truth_H[v]
    flops = _region_flops(res["n_regions"][v], res["n_evals"][v])
    achieved = flops / (k_ms / 1e3) / 1e12

    # the same scene with one region per pixel (n_spatial = 1), for what the neighbourhood buys
    one = P.DeviceOutputs(cfg, n, fields=("depth", "rrs_error", "status", "n_evals"))
    e0, e1 = _events()
    P.check(P.lib.photic_event_record(e0, stream))
    cfg1 = P.cfg_default(WL, mode=P.MODE_NATIVE, n_bottoms=nbot, reqmin=1e-2, konvge=100,
                         kcount=5000, watchdog_s=900.0, sec_sun=float(sec_sun))
    P.invert_device(cfg1, n, d_rrs.value, None, None, one, stream=stream)
    P.check(P.lib.photic_event_record(e1, stream))
    P.check(P.lib.photic_synchronize())
    ms1 = C.c_float()
    P.check(P.lib.photic_event_elapsed_ms(e0, e1, C.byref(ms1)))
    r1 = one.fetch()
    v1 = r1["status"] >= 0
    dH1 = r1["depth"][v1].astype(np.float64) - truth_H[v1]

    def quality(depth, err, dh, evals):
        return {"evals_per_pixel": float(evals.mean()),
                "median_rrs_error_pct": float(np.median(err)),
                "depth_median_abs_error_m": float(np.median(np.abs(dh))),
                "depth_runaway_frac": float((depth > RUNAWAY_M).mean()),
                "depth_rmse_bounded_m": float(np.sqrt(np.mean(dh[depth <= RUNAWAY_M] ** 2)))}

    line = {
        "metric": "pixels/sec inverted, spatial regions (n_spatial %d, %d bottoms), 1 MI355X"
                  % (ns, nbot),
        "value": n / (ms_step / 1e3), "unit": "pixels/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (Philox-seeded smooth parameter fields, no real imagery)",
        "config": {"workload": "region inversion, %dx%d pixels, one scene, %dx%d neighbourhood, "
                               "%d bottoms, reqmin 1e-2 konvge 100 kcount 5000"
                               % (side, side, 2 * ns - 1, 2 * ns - 1, nbot),
                   "pixels": n, "seed": 1, "noise_sigma": 2e-4, "theta_sun_deg": theta,
                   "parallelism": "1 GPU, one wavefront per pixel from a global counter"},
        "regions": dict(quality(res["depth"][v], res["rrs_error"][v], dH, res["n_evals"][v]),
                        region_histogram=np.bincount(res["n_regions"][v]).tolist()),
        "one_region": dict(quality(r1["depth"][v1], r1["rrs_error"][v1], dH1, r1["n_evals"][v1]),
                           pixels_per_s=n / (ms1.value / 1e3)),
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP64_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS, "traffic": None,
                     "kernel": region_kernel_name(nbot, (2 * ns - 1) ** 2 * (1 + 2 * nbot) + 3),
                     "kernel_ms": k_ms,
                     "flop_per_eval": "R (46 x 4) + 15 + 6 per pixel of R regions",
                     "flop_per_launch": flops},
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_bindings as O
        threads, _ = host_cpus()
        used = O.set_threads(threads)
        ns_cpu = min(args.region_cpu_sample, n)
        rrs_h = d_rrs.to_array(np.float32, (1, 4, n))
        t = time.perf_counter()
        O.invert_regions(rrs_h, side, side, np.arange(ns_cpu), -9999.0, [sec_sun], [1.0],
                         n_spatial=ns, n_bottoms=nbot, screen=True)
        dt = time.perf_counter() - t
        cpx = ns_cpu / dt
        line["cpu_baseline"] = {"value": cpx, "unit": "pixels/s", "cores": used, "kind": "port",
                                "sample": "first %d pixels of the same scene (with their "
                                          "neighbourhoods), FP64 oracle, OpenMP dynamic "
                                          "schedule, %.1f s" % (ns_cpu, dt)}
        line["cpu_baseline"]["host"] = host_cpus()[1]
        line["gpu_over_cpu"] = line["value"] / cpx
    P.lib.photic_stream_destroy(stream)
    print(json.dumps(line))


def bench_dropin(args, P):
    """The drop-in samodel() end to end: host grids in, host grids out."""
    from photic_amd.samodel_types import Geogrid, RowGrid, Scene
    side, nbot, ns = args.dropin_side, args.dropin_bottoms, args.dropin_spatial
    n = side * side
    theta = 35.0
    sec_sun = 1.0 / np.cos(np.radians(theta))
    cfg = P.cfg_default(WL, mode=P.MODE_NATIVE, n_bottoms=nbot, reqmin=1e-2, konvge=100,
                        kcount=5000, watchdog_s=900.0, sec_sun=float(sec_sun))
    d_rrs, d_tr = P.DeviceBuffer(16 * n), P.DeviceBuffer(24 * n)
    P.check(P.lib.photic_gen_field(C.byref(cfg), 1, side, side, theta, 2e-4, d_rrs.value,
                                   d_tr.value, None))
    P.check(P.lib.photic_synchronize())
    bands = d_rrs.to_array(np.float32, (4, side, side))
    truth_H = d_tr.to_array(np.float32, (6, n))[0].reshape(side, side)
    del d_tr
    grids = [RowGrid(bands[b]) for b in range(4)]
    gd = (Geogrid * 4)()
    for b in range(4):
        gd[b].nrows, gd[b].ncols, gd[b].nodata_value = side, side, -9999.0
        gd[b].array = grids[b].pp
    scenes = (Scene * 1)()
    sc = scenes[0]
    sc.n_bands, sc.nrows, sc.ncols = 4, side, side
    for b, wl in enumerate((443, 482, 561, 655)):
        sc.band_indexes[b] = b
        sc.wavelengths[b] = wl
    sc.theta_w, sc.theta_v = theta, 0.0
    idx = (C.c_int * 1)(0)
    names = ("depth", "sigma", "err", "alb", "sand", "grass", "coral", "kmin", "btype", "iod")
    P.lib.photic_samodel.restype = C.c_int

    def step():
        outs = {k: RowGrid(np.zeros((side, side), np.float32)) for k in names}
        t0 = time.perf_counter()
        rc = P.lib.photic_samodel(scenes, gd, idx, 1, 0, Geogrid(), 1, ns, nbot,
                                  *[outs[k].pp for k in names])
        wall = 1e3 * (time.perf_counter() - t0)
        P.check(rc)
        return wall, outs

    for _ in range(args.warmup):
        step()
    walls, outs = [], None
    for _ in range(args.steps):
        w, outs = step()
        walls.append(w)
        print("dropin step: %.1f ms" % w, file=sys.stderr, flush=True)
    ms_step = float(np.mean(walls))
    depth = -outs["depth"].array()  # samodel() writes depth negated (model/samodel.c:1486)
    err = outs["err"].array()
    v = err > 0.0
    dH = depth[v].astype(np.float64) - truth_H[v]

    # the same inversion with the scene already in HBM (photic_invert or, with regions,
    # photic_invert_regions; no copies)
    P.check(P.lib.photic_synchronize())
    t0 = time.perf_counter()
    if ns <= 1:
        dev = P.DeviceOutputs(cfg, n, fields=("depth", "status", "n_evals"))
        P.invert_device(cfg, n, d_rrs.value, None, None, dev)
    else:
        dev = P.invert_regions_device(cfg, side, side, 1, d_rrs.value, -9999.0, sec_sun, 1.0,
                                      ns, n)
    P.check(P.lib.photic_synchronize())
    dev_ms = 1e3 * (time.perf_counter() - t0)
    r = dev.fetch()
    same = float(np.mean(r["depth"].reshape(side, side)[v] == depth[v]))
    line = {
        "metric": "pixels/sec inverted through the drop-in samodel(), host grids in and out "
                  "(PCIe and host gather/scatter included), 1 MI355X",
        "value": n / (ms_step / 1e3), "unit": "pixels/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (Philox-seeded smooth parameter fields, no real imagery)",
        "config": {"workload": "samodel() drop-in, %dx%d pixels, one scene, n_spatial %d, "
                               "%d bottoms, reqmin 1e-2 konvge 100 kcount 5000, with the "
                               "depth-error pass, row-pointer float grids on the host"
                               % (side, side, ns, nbot),
                   "pixels": n, "seed": 1, "noise_sigma": 2e-4, "theta_sun_deg": theta,
                   "parallelism": "1 GPU, persistent work-queue kernel"},
        "modelled_frac": float(v.mean()),
        "median_rrs_error_pct": float(np.median(err[v])),
        "depth_median_abs_error_m": float(np.median(np.abs(dH))),
        "device_resident": {"pixels_per_s": n / (dev_ms / 1e3), "ms": dev_ms,
                            "evals_per_pixel": float(r["n_evals"][r["status"] >= 0].mean()),
                            "depth_identical_frac": same},
        "host_overhead_frac": 1.0 - dev_ms / ms_step,
        "cpu_baseline": None,
    }
    print(json.dumps(line))


def bench_stream(args, P):
    """MLP-seeded stack inversion of a host-resident rows x cols x 4-date scene, streamed."""
    nd, rows, cols = 4, args.stream_rows, args.stream_cols
    n = rows * cols
    cfg = P.cfg_default(WL, reqmin=1e-10, konvge=100, kcount=2000, watchdog_s=900.0)
    # the scene in host memory, generated on the GPU a tile at a time (untimed)
    t0 = time.perf_counter()
    rrs_h = np.empty((nd, 4, n), np.float32)
    ss_h = np.empty((nd, n), np.float32)
    v_h = np.empty((nd, n), np.uint8)
    H_h = np.empty(n, np.float32)
    tile = args.stream_tile_rows * cols
    g = [P.DeviceBuffer(16 * nd * tile), P.DeviceBuffer(4 * nd * tile), P.DeviceBuffer(nd * tile),
         P.DeviceBuffer(4 * (2 + 4 * nd) * tile)]
    for p0 in range(0, n, tile):
        nt = min(tile, n - p0)
        P.check(P.lib.photic_gen_stack(C.byref(cfg), 5, p0, nt, nd, 2e-4, 0.3, 2, g[0].value,
                                       g[1].value, g[2].value, g[3].value, None))
        for q in range(4 * nd):
            P.check(P.lib.photic_memcpy_dtoh(rrs_h[q // 4, q % 4, p0:].ctypes.data,
                                             g[0].value + 4 * q * nt, 4 * nt))
        for d in range(nd):
            P.check(P.lib.photic_memcpy_dtoh(ss_h[d, p0:].ctypes.data, g[1].value + 4 * d * nt,
                                             4 * nt))
            P.check(P.lib.photic_memcpy_dtoh(v_h[d, p0:].ctypes.data, g[2].value + d * nt, nt))
        P.check(P.lib.photic_memcpy_dtoh(H_h[p0:].ctypes.data, g[3].value, 4 * nt))
    del g
    print("scene %d x %d x %d dates generated in %.1f s" % (rows, cols, nd, time.perf_counter() - t0),
          file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    dx, dt = P.trainset_device(cfg, 101, args.train_samples)
    net = P.Mlp(init_seed=1)
    loss = net.train_device(dx.value, dt.value, args.train_samples, batch=8192,
                            epochs=args.train_epochs, lr=2e-3, lr_final=1e-4)
    P.check(P.lib.photic_synchronize())
    del dx, dt
    train_info = {"samples": args.train_samples, "epochs": args.train_epochs,
                  "final_loss": loss, "seconds": time.perf_counter() - t0}

    fields = ("depth", "rrs_error", "status", "n_evals", "n_starts")

    def report(d, tot):
        print("tile %d/%d done at %.1f s" % (d, tot, time.perf_counter() - t_run),
              file=sys.stderr, flush=True)

    def summary(res, rows_run, wall):
        m = rows_run * cols
        v = res["status"] >= 0
        dH = res["depth"][v].astype(np.float64) - H_h[:m][v]
        fit = res["rrs_error"][v] < 1.0
        return {"pixels_per_s": m / wall if wall else None, "seconds": wall,
                "evals_per_stack": float(res["n_evals"][v].mean()),
                "fallback_rate": float((res["n_starts"][v] > 1).mean()),
                "median_rrs_error_pct": float(np.median(res["rrs_error"][v])),
                "depth_rmse_m": float(np.sqrt(np.mean(dH * dH))),
                "depth_median_abs_error_m": float(np.median(np.abs(dH))),
                "depth_rmse_fitted_m": float(np.sqrt(np.mean(dH[fit] ** 2))),
                "fitted_frac": float(fit.mean()),
                "depth_runaway_frac": float((res["depth"][v] > RUNAWAY_M).mean()),
                "depth_rmse_bounded_m": float(np.sqrt(np.mean(
                    dH[res["depth"][v] <= RUNAWAY_M] ** 2))),
                "valid_pixels": int(v.sum())}

    t_run = time.perf_counter()
    res = P.invert_stack_stream(cfg, rows, cols, rrs_h, ss_h, v_h,
                                tile_rows=args.stream_tile_rows, net=net, fields=fields,
                                progress=report)
    wall = time.perf_counter() - t_run
    seeded = summary(res, rows, wall)
    line = {
        "metric": "pixels/sec inverted, MLP-seeded multi-date stacks streamed through HBM, "
                  "1 MI355X",
        "value": n / wall, "unit": "pixels/s", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": 1e3 * wall, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (Philox-seeded stack scene in host memory, no real imagery)",
        "config": {"workload": "MLP-seeded stack inversion with plain fallback, %dx%d pixels "
                               "x 4 dates, 2-4 valid, streamed in %d-row tiles (pinned "
                               "double-buffered copies on a side stream; wall time includes "
                               "them)" % (rows, cols, args.stream_tile_rows),
                   "pixels": n, "seed": 5, "noise_sigma": 2e-4,
                   "parallelism": "1 GPU, copy and compute streams"},
        "seeded": seeded, "mlp_training": train_info,
    }
    if args.stream_plain_rows > 0:
        pr = min(args.stream_plain_rows, rows)
        m = pr * cols
        t_run = time.perf_counter()
        pres = P.invert_stack_stream(cfg, pr, cols, np.ascontiguousarray(rrs_h[:, :, :m]),
                                     np.ascontiguousarray(ss_h[:, :m]),
                                     np.ascontiguousarray(v_h[:, :m]),
                                     tile_rows=args.stream_tile_rows, net=None, fields=fields,
                                     progress=report)
        line["plain_first_rows"] = dict(summary(pres, pr, time.perf_counter() - t_run), rows=pr)
        sres = {k: a[..., :m] for k, a in res.items()}
        line["seeded_first_rows"] = dict(summary(sres, pr, None), rows=pr)
    if not args.no_cpu:
        ns = min(args.stream_cpu_sample, n)
        cpx, threads, dt = _stack_cpu(rrs_h[:, :, :ns], ss_h[:, :ns], v_h[:, :ns], ns)
        line["cpu_baseline"] = {"value": cpx, "unit": "pixels/s", "cores": threads,
                                "kind": "port",
                                "sample": "first %d stacks of the same scene, FP64 oracle "
                                          "(plain), OpenMP dynamic schedule, %.1f s" % (ns, dt)}
        line["cpu_baseline"]["host"] = host_cpus()[1]
        line["gpu_over_cpu"] = line["value"] / cpx
    net.close()
    print(json.dumps(line))

def bench_ordered(args, P):
    """The ordered inversion (spectral-index order, look-up table, hot start) on the headline
    scene, beside the MLP-seeded and the plain inversion of the same pixels: px/s, median
    misfit, evaluations and roles."""
    n = args.side * args.side
    cfg = P.cfg_default(WL, reqmin=1e-10, konvge=100, kcount=2000, watchdog_s=900.0,
                        seed_abort_misfit=args.seed_abort)
    stream = C.c_void_p()
    P.check(P.lib.photic_stream_create(C.byref(stream)))
    d_rrs, d_ss, d_tr = P.DeviceBuffer(16 * n), P.DeviceBuffer(4 * n), P.DeviceBuffer(24 * n)
    P.check(P.lib.photic_gen_scene(C.byref(cfg), 4, 0, n, 2e-4, 0.0, d_rrs.value, d_ss.value,
                                   d_tr.value, stream))
    P.check(P.lib.photic_synchronize())
    truth_H = d_tr.to_array(np.float32, (6, n))[0]
    del d_tr
    fields = ("depth", "rrs_error", "status", "n_evals", "n_starts")
    d_role = P.DeviceBuffer(4 * n)

    def summary(ms, res):
        v = res["status"] >= 0
        dH = res["depth"][v].astype(np.float64) - truth_H[v]
        bounded = res["depth"][v] <= RUNAWAY_M
        return {"pixels_per_s": n / (ms / 1e3), "ms": ms,
                "evals_per_pixel": float(res["n_evals"][v].mean()),
                "median_rrs_error_pct": float(np.median(res["rrs_error"][v])),
                "p90_rrs_error_pct": float(np.percentile(res["rrs_error"][v], 90)),
                "depth_median_abs_error_m": float(np.median(np.abs(dH))),
                "depth_rmse_bounded_m": float(np.sqrt(np.mean(dH[bounded] ** 2))),
                "depth_runaway_frac": float((~bounded).mean())}

    def ordered(oc):
        outs = P.DeviceOutputs(cfg, n, fields=fields)
        P.check(P.lib.photic_synchronize())
        t0 = time.perf_counter()
        _, stats = P.invert_ordered_device(cfg, oc, n, d_rrs.value, d_ss.value, outs,
                                           d_role.value, stream)
        res = outs.fetch()
        ms = 1e3 * (time.perf_counter() - t0)
        role = d_role.to_array(np.int32, (n,))
        out = summary(ms, res)
        v = role >= 0
        out["roles"] = {name: float((role[v] == k).mean()) for k, name in
                        enumerate(("leader", "lut", "hot_start", "restarted"))}
        out["stats"] = stats
        return out

    def plain_or_seeded(d_seed):
        outs = P.DeviceOutputs(cfg, n, fields=fields)
        P.check(P.lib.photic_synchronize())
        t0 = time.perf_counter()
        if d_seed is not None:
            net.seed_device(cfg, n, d_rrs.value, d_ss.value, d_seed.value, stream=stream)
        P.invert_device(cfg, n, d_rrs.value, d_ss.value, d_seed.value if d_seed else None, outs,
                        stream=stream)
        P.check(P.lib.photic_synchronize())  # a non-blocking stream: wait before copying back
        res = outs.fetch()
        return summary(1e3 * (time.perf_counter() - t0), res)

    line = {"metric": "pixels/sec inverted, ordered inversion (look-up table, hot start) vs "
                      "MLP-seeded and plain, 1 MI355X",
            "unit": "pixels/s", "n_gpus": 1, "higher_is_better": True, "dtype": "f64",
            "data": "synthetic (Philox-seeded scene, no real imagery)",
            "config": {"workload": "ordered inversion of the %dx%d single-date scene (seed 4): "
                                   "spectral-index order, 64-pixel runs, 16-leader window, "
                                   "the reference's match and entry levels, no restart" %
                                   (args.side, args.side), "pixels": n}}
    for _ in range(args.warmup):
        ordered(P.order_cfg())
    ref = ordered(P.order_cfg())
    progress("ordered inversion, %.1f ms" % ref["ms"])
    line["ordered_lut"] = ref
    line["ordered_lut_noise_angle"] = ordered(P.order_cfg(r_sigma=[2e-4] * 4))
    line["ordered_lut_restart_5"] = ordered(P.order_cfg(restart_misfit=5.0))
    line["ordered_hot_start_only"] = ordered(P.order_cfg(use_lut=0))
    line["value"] = ref["pixels_per_s"]
    line["ms_per_step"] = ref["ms"]
    line["steps"], line["warmup"] = 1, args.warmup
    line["plain"] = plain_or_seeded(None)
    progress("ordered variants and plain done")
    dx, dt = P.trainset_device(cfg, 101, args.train_samples)
    net = P.Mlp(init_seed=1)
    net.train_device(dx.value, dt.value, args.train_samples, batch=8192,
                     epochs=args.train_epochs, lr=2e-3, lr_final=1e-4, stream=stream)
    P.check(P.lib.photic_synchronize())
    del dx, dt
    line["seeded"] = plain_or_seeded(P.DeviceBuffer(20 * n))
    line["cpu_baseline"] = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_bindings as O
        threads, _ = host_cpus()
        used = O.set_threads(threads)
        ns = min(args.ordered_cpu_sample, n)
        rrs_h = d_rrs.to_array(np.float32, (4, n))[:, :ns].copy()
        ss_h = d_ss.to_array(np.float32, (n,))[:ns].copy()
        t = time.perf_counter()
        O.invert_ordered(rrs_h, ss_h)
        dt_cpu = time.perf_counter() - t
        line["cpu_baseline"] = {"value": ns / dt_cpu, "unit": "pixels/s", "cores": used,
                                "kind": "port", "host": host_cpus()[1],
                                "sample": "first %d pixels of the same scene ordered among "
                                          "themselves, FP64 oracle, %.1f s" % (ns, dt_cpu)}
    net.close()
    P.lib.photic_stream_destroy(stream)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
