"""The launch-shape autotuner: it times every candidate shape on a sample, keeps the fastest
for the kernel instance, later calls use it, and the inversion's results are the same bit
for bit whatever shape runs it."""
import numpy as np
import pytest

import _oracle as O
import photic_amd as P

pytestmark = pytest.mark.gpu


def test_autotune_keeps_results_identical():
    rrs, ss, truth, nodata, _ = O.gen(51, 1 << 14)
    cfg = P.cfg_default(reqmin=1e-10, kcount=2000, check_watchdog=1)
    base = P.invert(cfg, rrs, ss)
    d_rrs = P.DeviceBuffer.from_array(rrs)
    d_ss = P.DeviceBuffer.from_array(ss)
    t = P.autotune(cfg, rrs.shape[1], d_rrs.value, d_ss.value)
    print("tuned:", t)
    assert t["block"] in (64, 128, 192, 256) and t["chunk"] in (16, 32, 64, 128, 256)
    assert t["blocks_per_cu"] >= 1 and t["n_tried"] >= 10 and t["ms"] > 0
    assert t["ms"] < 1e29
    tuned = P.invert(cfg, rrs, ss)
    for k in ("x", "status", "n_evals", "rrs_error"):
        assert np.array_equal(base[k], tuned[k]), k
    # an explicit shape still overrides the tuned one, with the same results
    forced = P.invert(P.cfg_default(reqmin=1e-10, kcount=2000, block=64, grid=7, chunk=16), rrs, ss)
    assert np.array_equal(base["x"], forced["x"])


def test_autotune_stack_keeps_results_identical():
    rrs, ss, valid, truth = O.gen_stack(52, 256, n_dates=4, k_min=2)
    cfg = P.cfg_default(reqmin=1e-10, kcount=2000, check_watchdog=1)
    base = P.invert_stack(cfg, rrs, ss, valid)
    d = [P.DeviceBuffer.from_array(a) for a in (rrs, ss, valid)]
    t = P.autotune_stack(cfg, rrs.shape[2], 4, d[0].value, d[1].value, d[2].value)
    print("tuned stacks:", t)
    tuned = P.invert_stack(cfg, rrs, ss, valid)
    assert np.array_equal(base["x"], tuned["x"])
    assert np.array_equal(base["n_evals"], tuned["n_evals"])
