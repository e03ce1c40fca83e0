"""The cost model's CPU side: the fp64 oracle of the perceptron (oracle/
e13mlp_ref.c) against central finite differences, and the library's
deterministic initialisation.  No kernel is launched here."""
import ctypes

import numpy as np

from conftest import load_pkg, oracle_lib

NP_ = 6402


def oracle_mlp():
    L = oracle_lib()
    D = ctypes.POINTER(ctypes.c_double)
    L.eo_mlp_grad.argtypes = [D, D, D, ctypes.c_int, D]
    L.eo_mlp_grad.restype = ctypes.c_double
    L.eo_mlp_forward.argtypes = [D, D, ctypes.c_int, D]
    L.eo_mlp_forward.restype = None
    return L


def dptr(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def ref_grad(L, P, x, t):
    g = np.zeros(NP_)
    loss = L.eo_mlp_grad(dptr(P), dptr(x), dptr(t), x.shape[0], dptr(g))
    return loss, g


def test_oracle_gradient_matches_finite_differences():
    L = oracle_mlp()
    rng = np.random.default_rng(7)
    P = rng.normal(0, 0.3, NP_)
    x = rng.normal(0, 1, (9, 32))
    t = rng.normal(0, 1, (9, 2))
    loss, g = ref_grad(L, P, x, t)
    idx = np.concatenate([rng.choice(NP_, 500, replace=False), np.arange(2048, 2112), np.arange(6400, 6402)])
    h = 1e-5
    for i in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += h
        Pm[i] -= h
        fd = (ref_grad(L, Pp, x, t)[0] - ref_grad(L, Pm, x, t)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(g[i]), (i, fd, g[i])
    # the loss is the mean squared error of the forward pass
    y = np.zeros((9, 2))
    L.eo_mlp_forward(dptr(P), dptr(x), 9, dptr(y))
    assert abs(loss - np.sum((y - t) ** 2) / 9) < 1e-12


def test_oracle_forward_matches_closed_form():
    """With W1, W2 identity-like and W3 summing, the output is a closed form."""
    L = oracle_mlp()
    P = np.zeros(NP_)
    for i in range(32):
        P[i * 32 + i] = 1.0               # W1: first 32 units copy the input
    for k in range(64):
        P[2112 + k * 64 + k] = 1.0        # W2: identity
    P[6272:6272 + 64] = 1.0               # y0 = sum of h2
    P[6336:6336 + 64] = -1.0              # y1 = -sum of h2
    P[6400], P[6401] = 0.5, -0.25
    x = np.array([[(-1) ** j * (j + 1) / 10 for j in range(32)]])
    y = np.zeros((1, 2))
    L.eo_mlp_forward(dptr(P), dptr(x), 1, dptr(y))
    s = np.sum(np.maximum(x, 0))
    assert abs(y[0, 0] - (s + 0.5)) < 1e-12 and abs(y[0, 1] - (-s - 0.25)) < 1e-12


def test_mlp_init_is_deterministic_and_scaled():
    e13 = load_pkg()
    a, b = np.array(e13.mlp_init(5)), np.array(e13.mlp_init(5))
    assert np.array_equal(a, b) and not np.array_equal(a, np.array(e13.mlp_init(6)))
    assert np.all(a[2048:2112] == 0) and np.all(a[6208:6272] == 0) and np.all(a[6400:] == 0)
    assert np.max(np.abs(a[:2048])) <= np.sqrt(6 / 32) and np.max(np.abs(a[2112:6208])) <= np.sqrt(6 / 64)
    assert np.std(a[:2048]) > 0.2


def test_cost_targets():
    e13 = load_pkg()
    t = e13.cost_targets(0, 0)
    assert t == [0.0, 0.0]
    t = e13.cost_targets(1000, 10 ** 6)
    assert abs(t[0] - np.log1p(1000) / 16) < 1e-6 and abs(t[1] - np.log1p(1e6) / 24) < 1e-6
