"""The ordered inversion (photic_invert_ordered: the reference's spectral-index ordering,
look-up table and hot start, model/samodel.c:701-1109, recast as runs with leaders) on the GPU
against the oracle's sequential restatement (oracle/orc_order.c): the same role for every
pixel (leader, LUT match, hot start, hot start then restart), status identical, every
converged pixel within the bar (H 1 cm, the rest 1e-4 relative) or flat, most bit-identical."""
import numpy as np
import pytest

import _oracle as O
import photic_amd as P

pytestmark = pytest.mark.gpu


def _scene(seed, rows, cols):
    rrs = O.gen_field(seed, rows, cols)[0]
    rrs[:, 5] = -9999.0            # no data
    rrs[2, 77] = -0.001            # negative
    return np.ascontiguousarray(rrs)


@pytest.mark.parametrize("settings,noise,mode,restart", [
    (dict(reqmin=1e-10, konvge=100, kcount=2000), None, 1, 5.0),
    (dict(reqmin=1e-2, konvge=100, kcount=5000), 2e-4, 1, 0.0),
    (dict(reqmin=1e-2, konvge=100, kcount=5000), None, 0, 5.0),
])
def test_ordered_against_oracle(settings, noise, mode, restart):
    rrs = _scene(9 + mode, 48, 64)
    n = rrs.shape[1]
    ss = np.full(n, O.sec_deg(35.0), np.float32)
    kw = dict(mode=P.MODE_NATIVE, n_bottoms=2) if mode == 0 else {}
    cfg = P.cfg_default(check_watchdog=1, **settings, **kw)
    rs = [noise] * 4 if noise else [0.0] * 4
    oc = P.order_cfg(run=16, window=4, r_sigma=rs, restart_misfit=restart)
    g, stats = P.invert_ordered(cfg, oc, rrs, ss)
    o = O.invert_ordered(rrs, ss, mode=mode, n_bottoms=2, run=16, window=4, restart=restart,
                         noise=None if noise is None else [noise] * 4, **settings)
    print("roles %s, stats %s" % (np.bincount(o["role"] + 1).tolist(), stats))
    assert np.array_equal(g["role"], o["role"])
    assert np.array_equal(g["status"], o["status"])
    assert stats["n_valid"] == (o["role"] >= 0).sum()
    assert stats["n_lut"] == (o["role"] == 1).sum()
    assert stats["n_restarted"] == (o["role"] == 3).sum()
    v = o["status"] >= 0
    both = (o["status"] == 0) & (g["status"] == 0)
    dH = np.abs(g["x"][0] - o["x"][0])
    rel = np.abs(g["x"][1:] - o["x"][1:]) / np.maximum(np.abs(o["x"][1:]), 1e-300)
    within = (dH < 0.01) & np.all((rel < 1e-4) | (o["x"][1:] == 0), axis=0)
    flat = np.abs(g["rrs_error"] - o["rrs_error"]) <= 1e-6 * np.abs(o["rrs_error"])
    exact = np.all(g["x"] == o["x"], axis=0)
    print("within %.4f, bit-identical %.4f" % (within[both].mean(), exact[v].mean()))
    assert np.all((within | flat)[both])
    assert exact[v].mean() > 0.98
    same = v & exact
    # the oracle also counts the reference's evaluation of each depth start's start point,
    # which the kernel (like nelmin itself) does not repeat
    assert np.array_equal(g["n_evals"][same], o["n_evals"][same] - o["n_starts"][same])
    assert np.array_equal(g["n_starts"][same], o["n_starts"][same])
    assert np.array_equal(g["depth"][same], o["depth"][same])
    assert np.all(g["depth"][~v] == 0.0) and np.all(g["status"][~v] == -1)
    assert (o["role"] == 2).sum() > 0 and (o["role"] == 0).sum() > 0
    assert ((o["role"] == 3).sum() > 0) == (restart > 0)


def test_ordered_lut_matches_adopt_the_leader():
    """With a permissive match level many followers (here 40 %) adopt a leader's vector: its
    x is that leader's, its misfit its own at that vector (one evaluation), no Nelder-Mead
    run."""
    rrs = _scene(21, 32, 32)
    n = rrs.shape[1]
    cfg = P.cfg_default(reqmin=1e-10, konvge=100, kcount=2000, check_watchdog=1)
    oc = P.order_cfg(run=8, window=8, match_max=50.0, entry_max=1e4)
    g, stats = P.invert_ordered(cfg, oc, rrs)
    o = O.invert_ordered(rrs, None, run=8, window=8, match_max=50.0, entry_max=1e4)
    assert np.array_equal(g["role"], o["role"])
    lut = g["role"] == 1
    assert lut.sum() > 0.3 * (g["role"] > 0).sum()
    assert np.all(g["n_evals"][lut] == 1) and np.all(g["n_starts"][lut] == 0)
    leaders = {tuple(g["x"][:, p]) for p in np.nonzero(g["role"] == 0)[0]}
    assert all(tuple(g["x"][:, p]) in leaders for p in np.nonzero(lut)[0])
    same = np.all(g["x"] == o["x"], axis=0) & lut
    assert same.sum() > 0.98 * lut.sum()
    assert np.array_equal(g["rrs_error"][same], o["rrs_error"][same])


def test_ordered_empty_and_all_nodata():
    cfg = P.cfg_default(check_watchdog=1)
    g, stats = P.invert_ordered(cfg, P.order_cfg(), np.full((4, 300), -9999.0, np.float32))
    assert stats["n_valid"] == 0 and np.all(g["role"] == -1) and np.all(g["status"] == -1)
