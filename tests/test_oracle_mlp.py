"""The CPU reference of the learned warm start's network (oracle/oracle_mlp.cpp):
its backward pass against central finite differences of its own loss (with the
BF16 rounding switched off, so the loss is smooth in the parameters)."""
import ctypes as C
import os

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
_FP = C.POINTER(C.c_float)
_DP = C.POINTER(C.c_double)


def oracle_mlp():
    lib = C.CDLL(os.path.join(REPO, "oracle", "liboracle2d.so"))
    lib.orc_mlp_set_exact.argtypes = [C.c_int]
    lib.orc_mlp_forward.argtypes = [_FP, C.c_int64, _FP, _FP]
    lib.orc_mlp_grad.argtypes = [_FP, C.c_int64, _FP, _FP, _DP, _DP]
    return lib


def rand_params(rng, scale=0.3):
    return (rng.standard_normal(4998) * scale).astype(np.float32)


def loss_grad(lib, p, x, t):
    g = np.zeros(4998)
    L = C.c_double(0)
    lib.orc_mlp_grad(p.ctypes.data_as(_FP), x.shape[0], x.ctypes.data_as(_FP), t.ctypes.data_as(_FP),
                     C.byref(L), g.ctypes.data_as(_DP))
    return L.value, g


def test_oracle_mlp_gradient_vs_finite_differences():
    lib = oracle_mlp()
    lib.orc_mlp_set_exact(1)
    try:
        rng = np.random.default_rng(7)
        p = rand_params(rng)
        x = rng.standard_normal((32, 6)).astype(np.float32)
        t = rng.standard_normal((32, 6)).astype(np.float32)
        L0, g = loss_grad(lib, p, x, t)
        # one coordinate from every parameter block: W1, b1, W2, b2, W3, b3
        for k in (5, 100, 390, 420, 500, 3000, 4550, 4600, 4700, 4990, 4993, 4997):
            h = 2e-2 * max(1.0, abs(float(p[k])))
            pp, pm = p.copy(), p.copy()
            pp[k] += h
            pm[k] -= h
            fd = (loss_grad(lib, pp, x, t)[0] - loss_grad(lib, pm, x, t)[0]) / (float(pp[k]) - float(pm[k]))
            assert abs(fd - g[k]) <= 2e-3 * max(abs(fd), 1e-3), (k, fd, g[k])
    finally:
        lib.orc_mlp_set_exact(0)


def test_oracle_mlp_bf16_rounding_is_round_to_nearest_even():
    """With rounding on, a network whose weights are already BF16 values and an
    identity-like path gives outputs that differ from the exact network only
    by the rounding of the activations (sanity of the rounding switch)."""
    lib = oracle_mlp()
    rng = np.random.default_rng(3)
    p = rand_params(rng, 0.1)
    x = rng.standard_normal((8, 6)).astype(np.float32)
    y = np.zeros((8, 6), np.float32)
    ye = np.zeros((8, 6), np.float32)
    lib.orc_mlp_forward(p.ctypes.data_as(_FP), 8, x.ctypes.data_as(_FP), y.ctypes.data_as(_FP))
    lib.orc_mlp_set_exact(1)
    lib.orc_mlp_forward(p.ctypes.data_as(_FP), 8, x.ctypes.data_as(_FP), ye.ctypes.data_as(_FP))
    lib.orc_mlp_set_exact(0)
    d = np.max(np.abs(y - ye)) / np.max(np.abs(ye))
    assert 1e-5 < d < 5e-2, d
