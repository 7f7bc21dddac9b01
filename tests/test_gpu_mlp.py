"""GPU MLP kernels (BF16 MFMA products, FP32 accumulation) against the numpy oracle with the same
BF16 rounding: forward outputs, loss and parameter gradients; a short GPU training run."""
import ctypes as C

import numpy as np
import pytest

from gsstructuralanalysis_amd import _lib as L
from oracle import mlp

pytestmark = pytest.mark.gpu


def _net(widths=mlp.WIDTHS, seed=3):
    h = C.c_void_p()
    w = (C.c_int * 4)(*widths)
    L.check(L.load().ksh_mlp_create(w, seed, C.byref(h)))
    n = C.c_int64()
    L.check(L.load().ksh_mlp_num_params(h, C.byref(n)))
    return h, n.value


def _f(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


@pytest.mark.parametrize("batch", [1, 37, 256])
def test_forward_and_gradients_match_oracle(batch, ksh):
    lib = L.load()
    h, n = _net()
    try:
        rng = np.random.default_rng(batch)
        p = np.zeros(n, dtype=np.float32)
        L.check(lib.ksh_mlp_get_params(h, _f(p)))
        p[-(256 + 256 + 65):] = rng.standard_normal(256 + 256 + 65).astype(np.float32) * 0.1  # nonzero biases
        L.check(lib.ksh_mlp_set_params(h, _f(p)))
        x = rng.standard_normal((batch, 264)).astype(np.float32)
        t = rng.standard_normal((batch, 65)).astype(np.float32)
        y = np.zeros((batch, 65), dtype=np.float32)
        L.check(lib.ksh_mlp_forward(h, _f(x), batch, _f(y)))
        hs, _ = mlp.forward(p, x, bf=True)
        yr = hs[-1]
        assert np.abs(y - yr).max() <= 2e-3 * np.abs(yr).max()
        g = np.zeros(n, dtype=np.float32)
        loss = C.c_double()
        L.check(lib.ksh_mlp_gradients(h, _f(x), _f(t), batch, _f(g), C.byref(loss)))
        lr, gr = mlp.loss_and_grad(p, x, t, bf=True)
        assert abs(loss.value - lr) <= 1e-3 * lr
        # per layer block: BF16 operands of the backward products bound the agreement
        err = np.linalg.norm(g - gr) / np.linalg.norm(gr)
        assert err <= 1e-2, err
        cos = float(g @ gr / (np.linalg.norm(g) * np.linalg.norm(gr)))
        assert cos >= 0.9999
    finally:
        lib.ksh_mlp_destroy(h)


def test_training_reduces_the_loss(ksh):
    """A smooth synthetic map learned by GPU Adam: the loss falls by more than 10x."""
    lib = L.load()
    h, n = _net(seed=5)
    try:
        rng = np.random.default_rng(7)
        A = rng.standard_normal((264, 65)) / np.sqrt(264)
        x = rng.standard_normal((4096, 264)).astype(np.float32)
        t = np.tanh(x @ A).astype(np.float32)
        epochs = 30
        losses = np.zeros(epochs)
        L.check(lib.ksh_mlp_train(h, _f(x), _f(t), 4096, epochs, 256, 1e-3, 0.0, 11,
                                  losses.ctypes.data_as(C.POINTER(C.c_double))))
        assert np.all(np.isfinite(losses))
        assert losses[-1] < 0.1 * losses[0]
    finally:
        lib.ksh_mlp_destroy(h)
