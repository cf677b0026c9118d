"""Multi-date stacks on the CPU: the per-lane stack engine (tests/hostsim/ph_stackengine.h, the
state machine of the product's ph_engine.h over 2 + 3k parameters, built for the host) against
the FP64 oracle bit for bit, and the oracle's multi-date problem against properties the
reference's formulation fixes.  The GPU's group engine is held to this host engine bit for bit
by tests/test_gpu_stack.py."""
import numpy as np
import pytest

import _oracle as O
import photic_amd as P


def _cfg():
    return P.cfg_default(reqmin=1e-10, konvge=100, kcount=2000)


def test_stack_engine_matches_oracle():
    rrs, ss, valid, truth = O.gen_stack(5, 96)
    h = O.hostsim_invert_stack(P.const_bytes(_cfg()), rrs, ss, valid)
    o = O.invert_stack(rrs, ss, valid)
    v = o["status"] >= 0
    assert np.array_equal(h["status"], o["status"])
    assert np.array_equal(h["x"], o["x"])
    for k in ("n_iter", "n_starts", "rrs_error", "model_error", "depth"):
        assert np.array_equal(h[k], o[k]), k
    # the oracle evaluates each depth start once before nelmin, as the reference does
    assert np.array_equal(h["n_evals"][v] + o["n_starts"][v], o["n_evals"][v])
    ks = valid.sum(0)
    assert len(set(ks[v])) >= 4  # several date counts exercised


def test_stack_engine_seeded_matches_oracle():
    rrs, ss, valid, truth = O.gen_stack(6, 48, n_dates=4, k_min=2)
    nd = 4
    seed = np.zeros((2 + 3 * nd, rrs.shape[2]), np.float32)
    seed[0] = truth[0] * 1.2
    seed[1] = truth[1]
    for d in range(nd):
        seed[2 + 3 * d:5 + 3 * d] = 100 * truth[2 + 4 * d:5 + 4 * d]
    h = O.hostsim_invert_stack(P.const_bytes(_cfg(), has_seed=True), rrs, ss, valid, seed=seed)
    o = O.invert_stack(rrs, ss, valid, seed=seed)
    assert np.array_equal(h["x"], o["x"])
    assert np.array_equal(h["n_iter"], o["n_iter"])
    assert np.array_equal(h["n_starts"], o["n_starts"])


def test_stack_cr_math_against_oracle():
    """With the correctly rounded exp/log/pow the GPU uses: within the parity bar on every
    converged pixel (or with the same misfit where the misfit is flat)."""
    rrs, ss, valid, truth = O.gen_stack(7, 96, n_dates=6)
    h = O.hostsim_invert_stack(P.const_bytes(_cfg()), rrs, ss, valid, math="cr")
    o = O.invert_stack(rrs, ss, valid)
    both = (o["status"] == 0) & (h["status"] == 0)
    dH = np.abs(h["x"][0] - o["x"][0])
    rel = np.abs(h["x"][1:] - o["x"][1:]) / np.maximum(np.abs(o["x"][1:]), 1e-300)
    within = (dH < 0.01) & np.all((rel < 1e-4) | (o["x"][1:] == 0), axis=0)
    same_fit = np.abs(h["rrs_error"] - o["rrs_error"]) <= 1e-6 * np.abs(o["rrs_error"])
    assert np.all((within | same_fit)[both])
    assert np.all(h["x"] == o["x"], axis=0)[o["status"] >= 0].mean() > 0.95


def test_stack_of_identical_dates_is_the_single_date_problem():
    """k copies of one date, with no K penalty in play (H >= 5, K_min <= 0.7): E_Rrs is a
    mean over (date, band), so the k-date misfit equals the single-date one at the same H, B,
    P, G, X (model/samodel.c:2545-2579)."""
    rrs1, ss1, _, _, _ = O.gen(9, 8)
    x1 = np.array([[7.0, 0.4, 3.0, 2.0, 1.0]])
    for k in (2, 3):
        xk = np.concatenate([[7.0, 0.4], np.tile([3.0, 2.0, 1.0], k)])
        for p in range(8):
            o1, e1 = O.objective(x1, rrs1[:, p:p + 1].T, ss1[p:p + 1])
            r = np.repeat(rrs1[:, p][None, :], k, axis=0)
            s = np.repeat(ss1[p:p + 1], k)
            ok, ek = O.objective_stack(xk, r, s)
            assert np.isclose(ek, e1[0], rtol=1e-13)
            assert np.isclose(ok, o1[0], rtol=1e-13)


@pytest.mark.parametrize("nbot", [1, 2, 3])
def test_native_stack_of_identical_dates_is_the_single_date_problem(nbot):
    """The native layout [H, B_k, q_k, (P, G, X) per date]: k copies of one date give the
    single-date misfit at the same parameters (no K penalty in play), and the same start point
    -- amplitudes 100 * 4 * mean Rrs(490) over the dates, weights 1, each date's P, G, X from
    its own Rrs -- as the single date (model/samodel.c:2249-2354, 2545-2579)."""
    rrs1, ss1, _, _, _ = O.gen(19 + nbot, 6)
    bq = [30.0 + 10 * j for j in range(nbot)] + [1.0 + 0.5 * j for j in range(nbot)]
    x1 = np.array([8.0] + bq + [3.0, 2.0, 1.0])
    for k in (2, 3):
        xk = np.concatenate([[8.0], bq, np.tile([3.0, 2.0, 1.0], k)])
        for p in range(6):
            r = np.repeat(rrs1[:, p][None, :], k, axis=0)
            o1, e1 = O.objective_stack(x1, rrs1[:, p][None, :], ss1[p:p + 1], n_bottoms=nbot)
            ok, ek = O.objective_stack(xk, r, np.repeat(ss1[p:p + 1], k), n_bottoms=nbot)
            assert np.isclose(ek, e1, rtol=1e-13) and np.isclose(ok, o1, rtol=1e-13)
            s1, st1 = O.start_step_stack(rrs1[:, p][None, :].astype(np.float64), 15.0, mode=0,
                                         n_bottoms=nbot)
            sk, stk = O.start_step_stack(r.astype(np.float64), 15.0, mode=0, n_bottoms=nbot)
            off = 1 + 2 * nbot
            assert np.allclose(sk[:off], s1[:off], rtol=1e-15, atol=0)
            assert np.allclose(stk[:off], st1[:off], rtol=1e-15, atol=0)
            assert np.array_equal(sk[off:], np.tile(s1[off:], k))
            assert np.array_equal(stk[off:], np.tile(st1[off:], k))


def test_stack_generator_draws():
    rrs, ss, valid, truth = O.gen_stack(11, 4000, n_dates=8, p_cloud=0.3, k_min=3)
    k = valid.sum(0)
    clear = (rrs >= 0).all(1)
    assert np.all(valid <= clear)            # a negative date is never valid
    assert np.all(k >= 1)
    frac_valid = valid.mean()
    assert 0.6 < frac_valid < 0.75           # Bernoulli(0.3) clouds, less negative dates
    theta = truth[5::4]
    assert np.allclose(ss, 1 / np.cos(np.radians(theta)), rtol=1e-6)
    assert np.all((truth[0] >= 0.5) & (truth[0] <= 25)) and np.all((truth[1] >= 0) & (truth[1] <= 1))
