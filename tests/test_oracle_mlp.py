"""The FP64 oracle of the warm-start network: its backward pass against central finite
differences of its own loss (the check that makes it a reference for the GPU kernels)."""
import ctypes as C

import numpy as np

import _oracle as O


def _loss_grad(w, x, t):
    L = O.lib()
    g = np.zeros_like(w)
    loss = L.orc_py_mlp_loss_grad(w.ctypes.data, x.shape[0], x.ctypes.data, t.ctypes.data,
                                  g.ctypes.data)
    return loss, g


def test_mlp_backward_matches_finite_differences():
    rng = np.random.default_rng(0)
    npar = O.lib().orc_py_mlp_nparam()
    assert npar == 34437  # 5*128+128 + 2*(128*128+128) + 128*5+5
    w = rng.normal(0, 0.2, npar)
    x = rng.normal(0, 1, (12, 5))
    t = rng.normal(0, 1, (12, 5))
    loss, g = _loss_grad(w, x, t)
    idx = np.concatenate([rng.integers(0, npar, 300), [0, 639, 640, 767, 768, npar - 1]])
    h = 1e-6
    for i in idx:
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        fd = (_loss_grad(wp, x, t)[0] - _loss_grad(wm, x, t)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])


def test_mlp_forward_linear_limit():
    # with every hidden weight zero the network is the output bias; with W1, W2, W3 zero and
    # biases b the hidden activations are gelu(b) -- closed forms
    npar = O.lib().orc_py_mlp_nparam()
    w = np.zeros(npar)
    w[-5:] = [0.1, -0.2, 0.3, -0.4, 0.5]
    x = np.ones((3, 5))
    y = np.zeros((3, 5))
    O.lib().orc_py_mlp_forward(w.ctypes.data, 3, x.ctypes.data, y.ctypes.data)
    assert np.allclose(y, [[0.1, -0.2, 0.3, -0.4, 0.5]] * 3)
