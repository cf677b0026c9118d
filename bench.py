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
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "pixels/sec inverted, MLP-seeded vs plain NM at equal misfit, 1 MI355X"
FLOP_PER_EVAL = 1 * (46 * 4 + 15) + 6  # 205
PEAK_FP64_TFLOPS = 78.6
WL = (443.0, 482.0, 561.0, 655.0)


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


def cpu_baseline(rrs_h, ss_h, n_sample, seconds_hint):
    """The FP64 oracle (OpenMP over pixels, the reference's dynamic schedule) on the first
    n_sample pixels of the same scene; returns px/s and the threads used."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_bindings as O
    threads = min(16, os.cpu_count() or 1)
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
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every box in this setup has one MI355X and the "
              "inversion has no cross-GPU step (pixels are independent; see DESIGN.md for how "
              "it shards). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import photic_amd as P
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
                        watchdog_s=900.0)

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
    fields = ("depth", "rrs_error", "status", "n_evals", "n_starts")

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

    def run(kind, cfg_run, n_run):
        """One step of `kind` over the first n_run pixels; returns (wall ms, kernel ms,
        outputs)."""
        r, sv = scene_slice(n_run)
        outs = P.DeviceOutputs(cfg_run, n_run, fields=fields)
        P.check(P.lib.photic_synchronize())
        t0 = time.perf_counter()
        if kind == "seeded":
            net.seed_device(cfg_run, n_run, r.value, sv.value, d_seed.value, stream=stream)
        k_ms = timed_invert(P, cfg_run, n_run, r.value, sv.value,
                            d_seed.value if kind == "seeded" else None, outs, stream)
        P.check(P.lib.photic_synchronize())
        return 1e3 * (time.perf_counter() - t0), k_ms, outs.fetch()

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
            "depth_rmse_m": float(np.sqrt(np.mean(dH * dH))),
            "depth_median_abs_error_m": float(np.median(np.abs(dH))),
            "depth_rmse_fitted_m": float(np.sqrt(np.mean(dH[fit] ** 2))),
            "fitted_frac": float(fit.mean()),
            "valid_pixels": int(v.sum()),
        }

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

    # ---- the plain GPU Nelder-Mead on the same scene, beside it ----
    comparisons = {}
    if net is not None:
        w_ms, _, pres = run("plain", cfg, n)
        comparisons["plain"] = summary(w_ms, pres, n)
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
            v_all = res["status"] >= 0
            comparisons["plain_equal_misfit"] = dict(
                full, early_exit=ee_eq, pixels=n,
                seeded_median_rrs_error_pct=float(np.median(res["rrs_error"][v_all])),
                chosen_on_pixels=n_eq, seeded_median_rrs_error_pct_on_those=seeded_med)
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
        cpx, threads, dt = cpu_baseline(rrs_h, ss_h, ns, 20)
        line["cpu_baseline"] = {
            "value": cpx, "unit": "pixels/s", "cores": threads, "kind": "port",
            "sample": "first %d pixels of the same scene, FP64 oracle, OpenMP dynamic "
                      "schedule, %.1f s" % (ns, dt),
        }
        line["gpu_over_cpu"] = px_s / cpx
        if "plain" in line:
            line["plain_gpu_over_cpu"] = line["plain"]["pixels_per_s"] / cpx

    if net is not None:
        net.close()
    P.lib.photic_stream_destroy(stream)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
