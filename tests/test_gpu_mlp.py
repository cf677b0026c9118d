"""The warm-start network's BF16-MFMA kernels against the FP64 oracle (oracle/orc_mlp.c),
its training, and what it gains the inversion.

Tolerances: the GPU rounds weights, inputs and every layer's activations and back-propagated
gradients to BF16 (8-bit mantissa, relative spacing 2^-8 = 0.4 %) and accumulates in FP32, so
outputs are held to 2 % and the gradient to 5 % in norm (cosine > 0.995 per layer) of the
FP64 oracle evaluated on the same BF16-representable weights and inputs."""
import numpy as np
import pytest

import _oracle as O
import photic_amd as P

pytestmark = pytest.mark.gpu

NP = 34437
LAYERS = {"W1": (0, 640), "b1": (640, 768), "W2": (768, 17152), "b2": (17152, 17280),
          "W3": (17280, 33664), "b3": (33664, 33792), "W4": (33792, 34432), "b4": (34432, 34437)}


def _bf16(a):
    a = np.asarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32)


def _rand_net(seed, scale=1.0):
    rng = np.random.default_rng(seed)
    w = np.zeros(NP, np.float32)
    for name, (a, b) in LAYERS.items():
        fan = {"W1": 5, "W2": 128, "W3": 128, "W4": 128}.get(name, 1)
        w[a:b] = rng.normal(0, scale / np.sqrt(fan) if name[0] == "W" else 0.1, b - a)
    return _bf16(w)


def test_mlp_forward_matches_oracle():
    net = P.Mlp()
    w = _rand_net(1)
    net.set_weights(w)
    rng = np.random.default_rng(2)
    x = _bf16(rng.normal(0, 1, (4096, 5)))
    y = net.forward(x)
    yo = np.zeros((4096, 5))
    wd, xd = w.astype(np.float64), x.astype(np.float64)
    O.lib().orc_py_mlp_forward(wd.ctypes.data, 4096, xd.ctypes.data, yo.ctypes.data)
    rel = np.linalg.norm(y - yo) / np.linalg.norm(yo)
    print("forward rel err %.4g" % rel)
    assert rel < 0.02


def test_mlp_gradient_matches_oracle():
    net = P.Mlp()
    w = _rand_net(3)
    net.set_weights(w)
    rng = np.random.default_rng(4)
    n = 1024
    x = _bf16(rng.normal(0, 1, (n, 5)))
    t = rng.normal(0, 1, (n, 5)).astype(np.float32)
    loss, g = net.gradient(x, t)
    go = np.zeros(NP)
    wd, xd, td = w.astype(np.float64), x.astype(np.float64), t.astype(np.float64)
    lo = O.lib().orc_py_mlp_loss_grad(wd.ctypes.data, n, xd.ctypes.data, td.ctypes.data,
                                      go.ctypes.data)
    print("loss gpu %.6g oracle %.6g" % (loss, lo))
    assert abs(loss - lo) <= 0.02 * abs(lo)
    for name, (a, b) in LAYERS.items():
        ga, gb = g[a:b].astype(np.float64), go[a:b]
        cos = ga @ gb / (np.linalg.norm(ga) * np.linalg.norm(gb) + 1e-30)
        rel = np.linalg.norm(ga - gb) / (np.linalg.norm(gb) + 1e-30)
        print("%s: cos %.5f rel %.4f" % (name, cos, rel))
        assert cos > 0.995 and rel < 0.05, name
    assert np.linalg.norm(g - go) / np.linalg.norm(go) < 0.05


def test_mlp_training_reduces_loss():
    cfg = P.cfg_default()
    n = 1 << 18
    dx, dt = P.trainset_device(cfg, 101, n)
    net = P.Mlp(init_seed=5)
    x = dx.to_array(np.float32, (n, 5))
    t = dt.to_array(np.float32, (n, 5))
    l0, _ = net.gradient(x[:8192], t[:8192])
    l1 = net.train_device(dx.value, dt.value, n, batch=4096, epochs=4, lr=2e-3, lr_final=2e-4)
    l2, _ = net.gradient(x[:8192], t[:8192])
    print("loss before %.4f, last epoch %.4f, after %.4f" % (l0, l1, l2))
    assert l2 < 0.6 * l0
