"""The product's lane engine (photic_amd/csrc/ph_engine.h, the code the GPU kernel runs),
compiled for the host by tests/hostsim and run one lane at a time, against the FP64 oracle:
with the host's libm on both sides the two must agree bit for bit -- every parameter, every
evaluation count, every convergence flag.  This pins the state-machine recasting of the
reference's Nelder-Mead and multi-start, independent of GPU math-library rounding."""
import numpy as np
import pytest

import _oracle as O
import photic_amd as P


def _cfg(**kw):
    base = dict(reqmin=1e-10, konvge=100, kcount=2000)
    base.update(kw)
    return P.cfg_default(**base)


@pytest.mark.parametrize("mode,nbot", [(1, 0), (0, 1), (0, 2), (0, 3)])
def test_engine_matches_oracle(mode, nbot):
    rrs, ss, truth, nodata, _ = O.gen(1, 160)
    cfg = _cfg(mode=mode, n_bottoms=max(nbot, 1) if mode == 0 else 2)
    const = P.const_bytes(cfg)
    h = O.hostsim_invert(const, nbot, rrs, ss)
    o = O.invert(rrs, ss, mode=mode, n_bottoms=max(nbot, 1), reqmin=1e-10, konvge=100,
                 kcount=2000)
    valid = o["status"] >= 0
    assert np.array_equal(h["status"], o["status"])
    assert np.array_equal(h["x"][:, valid], o["x"][:, valid])
    assert np.array_equal(h["n_iter"][valid], o["n_iter"][valid])
    assert np.array_equal(h["n_starts"][valid], o["n_starts"][valid])
    assert np.array_equal(h["rrs_error"][valid], o["rrs_error"][valid])
    assert np.array_equal(h["model_error"][valid], o["model_error"][valid])
    # the oracle evaluates each depth start once before handing it to nelmin, as the
    # reference does (model/samodel.c:2365); the engine skips that discarded evaluation
    assert np.array_equal(h["n_evals"][valid] + o["n_starts"][valid], o["n_evals"][valid])


def test_engine_matches_oracle_reference_settings():
    # the reference's own settings: reqmin 1e-2, kcount 5000
    rrs, ss, _, _, _ = O.gen(2, 128)
    cfg = P.cfg_default()
    h = O.hostsim_invert(P.const_bytes(cfg), 0, rrs, ss)
    o = O.invert(rrs, ss, reqmin=1e-2, konvge=100, kcount=5000)
    assert np.array_equal(h["x"], o["x"])
    assert np.array_equal(h["n_iter"], o["n_iter"])


def test_engine_matches_oracle_seeded():
    rrs, ss, truth, nodata, _ = O.gen(3, 128)
    seed = np.stack([truth[0] * 1.2, truth[4], 100 * truth[1], 100 * truth[2], 100 * truth[3]])
    cfg = _cfg()
    const = P.const_bytes(cfg, has_seed=True)
    h = O.hostsim_invert(const, 0, rrs, ss, seed=seed.astype(np.float32))
    o = O.invert(rrs, ss, seed=seed.astype(np.float32))
    assert np.array_equal(h["x"], o["x"])
    assert np.array_equal(h["n_iter"], o["n_iter"])
    assert np.array_equal(h["n_starts"], o["n_starts"])


def test_engine_seed_abort_hands_over_to_the_depth_starts():
    """cfg.seed_abort_misfit: a seed start still above the level at its first convergence check
    is given up.  The depth starts then run as the reference's: where the plain oracle ends
    below the level, an aborted pixel's result is the oracle's plain result bit for bit, with
    the seed start counted as one more start; every other pixel runs as without the abort."""
    rrs, ss, truth, nodata, _ = O.gen(5, 256)
    seed = np.stack([truth[0] * 0.9, truth[4], 100 * truth[1], 100 * truth[2],
                     100 * truth[3]]).astype(np.float32)
    wrong = np.arange(seed.shape[1]) % 2 == 1  # every other pixel: its neighbour's seed
    seed[:, wrong] = np.roll(seed, 1, axis=1)[:, wrong]
    h = O.hostsim_invert(P.const_bytes(_cfg(seed_abort_misfit=5.0), has_seed=True), 0, rrs, ss,
                         seed=seed)
    full = O.hostsim_invert(P.const_bytes(_cfg(), has_seed=True), 0, rrs, ss, seed=seed)
    o = O.invert(rrs, ss, reqmin=1e-10, konvge=100, kcount=2000)
    valid = o["status"] >= 0
    same = np.all(h["x"] == full["x"], axis=0) & (h["n_evals"] == full["n_evals"])
    aborted = valid & ~same
    sel = aborted & (o["model_error"] < 5.0)
    print("aborted %d of %d (wrong seed %d), compared with the plain oracle %d"
          % (aborted.sum(), valid.sum(), (aborted & wrong).sum(), sel.sum()))
    assert sel.sum() >= 5
    for k in ("x", "status", "n_iter", "rrs_error", "model_error"):
        assert np.array_equal(h[k][..., sel], o[k][..., sel]), k
    assert np.array_equal(h["n_starts"][sel], o["n_starts"][sel] + 1)
    assert np.all(h["n_evals"][sel] > o["n_evals"][sel] - o["n_starts"][sel])
    # a level below the early exit (2.5) acts as the early exit: a given-up seed start can
    # never be the accepted result
    lo = O.hostsim_invert(P.const_bytes(_cfg(seed_abort_misfit=1.0), has_seed=True), 0, rrs, ss,
                          seed=seed)
    at = O.hostsim_invert(P.const_bytes(_cfg(seed_abort_misfit=2.5), has_seed=True), 0, rrs, ss,
                          seed=seed)
    for k in ("x", "status", "n_evals", "n_starts"):
        assert np.array_equal(lo[k], at[k]), k


def test_engine_small_kcount_and_restarts():
    # tiny kcount: runs end on the evaluation limit (ifault 2) and the restart branch
    rrs, ss, _, _, _ = O.gen(4, 96)
    for kcount, reqmin in ((60, 1e-10), (400, 1e-1), (3000, 1e-1)):
        cfg = _cfg(kcount=kcount, reqmin=reqmin, konvge=7)
        h = O.hostsim_invert(P.const_bytes(cfg), 0, rrs, ss)
        o = O.invert(rrs, ss, reqmin=reqmin, konvge=7, kcount=kcount)
        assert np.array_equal(h["x"], o["x"]), kcount
        assert np.array_equal(h["status"], o["status"]), kcount


@pytest.mark.parametrize("mode,nbot", [(1, 0), (0, 2)])
def test_engine_cr_math_against_oracle(mode, nbot):
    """The engine with the correctly rounded exp/log/pow the GPU uses (the host C library
    misrounds about 1 call in 10^3): every converged pixel within the parity bar -- H within
    1 cm, the other parameters within 1e-4 relative -- and nearly all bit-identical."""
    rrs, ss, truth, nodata, _ = O.gen(4, 600)
    cfg = _cfg(mode=mode, n_bottoms=max(nbot, 1))
    h = O.hostsim_invert(P.const_bytes(cfg), nbot, rrs, ss, math="cr")
    o = O.invert(rrs, ss, mode=mode, n_bottoms=max(nbot, 1), reqmin=1e-10, konvge=100,
                 kcount=2000)
    valid = o["status"] >= 0
    assert np.array_equal(h["status"], o["status"])
    both = valid & (o["status"] == 0)
    dH = np.abs(h["x"][0] - o["x"][0])
    rel = np.abs(h["x"][1:] - o["x"][1:]) / np.maximum(np.abs(o["x"][1:]), 1e-300)
    within = (dH < 0.01) & np.all(rel < 1e-4, axis=0)
    # a pixel outside the bar must be one whose misfit is flat along the coordinates that
    # differ (deep or turbid water: exp(-K H) below the double epsilon makes H and the bottom
    # invisible): both end points then have the same misfit
    same_fit = np.abs(h["rrs_error"] - o["rrs_error"]) <= 1e-6 * np.abs(o["rrs_error"])
    assert np.all((within | same_fit)[both])
    assert within[both].mean() > 0.99
    assert np.all(h["x"][:, valid] == o["x"][:, valid], axis=0).mean() > 0.99


def test_engine_nan_seed_means_no_seed_start():
    """A pixel whose seed depth is NaN has no seed start: it runs the depth starts alone and
    ends exactly where the plain inversion ends, in the host build of the engine and in the
    oracle alike; the other pixels keep their seed start."""
    rrs, ss, truth, nodata, _ = O.gen(26, 96)
    seed = np.stack([truth[0] * 1.1, truth[4], 100 * truth[1], 100 * truth[2],
                     100 * truth[3]]).astype(np.float32)
    skip = np.arange(seed.shape[1]) % 3 == 0
    seed[0, skip] = np.nan
    cfg = P.cfg_default(reqmin=1e-10, konvge=100, kcount=2000)
    h = O.hostsim_invert(P.const_bytes(cfg, has_seed=True), 0, rrs, ss, seed=seed)
    o = O.invert(rrs, ss, seed=seed)
    plain = O.invert(rrs, ss)
    v = o["status"] >= 0
    assert np.array_equal(h["x"], o["x"]) and np.array_equal(h["n_starts"], o["n_starts"])
    assert np.array_equal(o["x"][:, skip & v], plain["x"][:, skip & v])
    assert np.array_equal(o["n_starts"][skip & v], plain["n_starts"][skip & v])
    assert np.any(o["n_starts"][~skip & v] != plain["n_starts"][~skip & v])
