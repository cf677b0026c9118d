"""Benchmark: pixels/sec of photic's per-pixel inversion on one MI355X.

Workload: the 8192 x 8192 synthetic single-date scene (4 bands at 443/482/561/655 nm,
Philox-seeded parameters drawn from the plain-method distributions, sun zenith 20-50 deg,
Gaussian noise 2e-4 1/sr), Nelder-Mead over the five parameters (H, B, P, G, X) with
reqmin 1e-10, konvge 100, kcount 2000 and the reference's depth starts with early exit.
A step is one inversion of the whole scene with inputs already resident in HBM.  With the
warm-start network available the headline is the MLP-seeded inversion (network inference +
seeded Nelder-Mead + plain fallback); the plain GPU Nelder-Mead is timed on the same scene
beside it.

Prints one JSON line.  Besides the contract's fields it carries the roofline of the dominant
kernel (algorithmic FLOPs k(46 n_b + 15) + 6 = 205 per objective evaluation over the kernel's
duration from HIP events on its stream) and the CPU baseline (the FP64 oracle, OpenMP over
pixels, timed on a bounded sample of the same scene on this host's cores).
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
    ap.add_argument("--cpu-sample", type=int, default=131072)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every box in this setup has one MI355X and the "
              "inversion has no cross-GPU step (pixels are independent; see DESIGN.md for how "
              "it shards). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import photic_amd as P
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
    outs = P.DeviceOutputs(cfg, n, fields=("depth", "rrs_error", "status", "n_evals"))

    # ---- warm-up on a slice ----
    n_w = min(n, 1 << 20)
    outs_w = P.DeviceOutputs(cfg, n_w, fields=("depth", "status"))
    for _ in range(args.warmup):
        timed_invert(P, cfg, n_w, d_rrs.value, d_ss.value, None, outs_w, stream)

    # ---- timed: K whole-scene plain inversions ----
    P.check(P.lib.photic_synchronize())
    t0 = time.perf_counter()
    kernel_ms = []
    for _ in range(args.steps):
        kernel_ms.append(timed_invert(P, cfg, n, d_rrs.value, d_ss.value, None, outs, stream))
    P.check(P.lib.photic_synchronize())
    wall = time.perf_counter() - t0
    ms_per_step = 1e3 * wall / args.steps

    res = outs.fetch()
    valid = res["status"] >= 0
    evals = float(res["n_evals"][valid].astype(np.int64).sum())
    truth_h = d_truth.to_array(np.float32, (6, n))
    dH = res["depth"][valid].astype(np.float64) - truth_h[0][valid]
    px_s = n / (ms_per_step / 1e3)
    k_ms = float(np.mean(kernel_ms))
    achieved = evals * FLOP_PER_EVAL / (k_ms / 1e3) / 1e12
    # HBM bytes per launch: the per-pixel FETCH_SIZE / WRITE_SIZE of the same kernel from
    # rocprofv3 --pmc, each calibrated by a copy kernel of known size (profiles/)
    traffic = None
    tj = os.path.join(ROOT, "profiles", "invert_fp64_traffic.json")
    if prec == P.FP64 and os.path.exists(tj):
        with open(tj) as f:
            t = json.load(f)
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
        "data": "synthetic (Philox-seeded scene, no real imagery)",
        "config": {
            "workload": "plain GPU Nelder-Mead, %dx%d single-date scene, 4 bands, 5 params, "
                        "reqmin 1e-10 kcount 2000" % (args.side, args.side),
            "pixels": n, "seed": 4, "noise_sigma": 2e-4,
            "parallelism": "1 GPU, persistent work-queue kernel",
            "mlp_seeded": None,
        },
        "plain": {
            "pixels_per_s": px_s,
            "evals_per_pixel": evals / max(valid.sum(), 1),
            "converged_frac": float((res["status"][valid] == 0).mean()),
            "median_rrs_error_pct": float(np.median(res["rrs_error"][valid])),
            "depth_rmse_m": float(np.sqrt(np.mean(dH * dH))),
            "valid_pixels": int(valid.sum()),
        },
        "roofline": {
            "bound": "valu",
            "achieved": achieved,
            "peak": PEAK_FP64_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP64_TFLOPS,
            "traffic": traffic,
            "kernel": "ph_invert_kernel<double,4,0>" if prec == P.FP64 else "ph_invert_kernel<float,4,0>",
            "kernel_ms": k_ms,
            "flop_per_eval": FLOP_PER_EVAL,
            "evals_per_launch": evals,
        },
        "cpu_baseline": None,
    }

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

    P.lib.photic_stream_destroy(stream)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
