"""Learned arc-length predictor: training data from the solver's own paths, training on the GPU by
the library's MLP kernels (ksh_mlp_*), and use through gsALMCrisfieldLearned (the `learned=`
option of ShellProblem.run_alm / run_alm_batch).

Features of a converged state (66, ksh_predictor_features): lambda, ||U||, 64 out-of-plane mode
coefficients.  A sample's input is the features of four consecutive converged states (264), its
target the deviation of the next state from the quadratic extrapolation 3 x_t - 3 x_{t-1} +
x_{t-2} of the last three, in the mode coefficients (64) and the load factor (1) -- the
correction gsALMCrisfieldLearned adds to that extrapolation.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L

FEAT, IN, OUT = 66, 264, 65
WIDTHS = (IN, 256, 256, OUT)


def dataset(paths):
    """(X, Y) from per-path feature arrays [steps x 66] of converged states."""
    xs, ys = [], []
    for f in paths:
        f = np.asarray(f, dtype=np.float64)
        for t in range(3, len(f) - 1):
            ext = 3.0 * f[t] - 3.0 * f[t - 1] + f[t - 2]
            xs.append(f[t - 3:t + 1].ravel())
            ys.append(np.concatenate([f[t + 1, 2:] - ext[2:], [f[t + 1, 0] - ext[0]]]))
    return np.array(xs).reshape(-1, IN), np.array(ys).reshape(-1, OUT)


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class LearnedPredictor:
    """The 264-256-256-65 GELU MLP on the GPU, with its input / output normalization."""

    def __init__(self, seed: int = 1):
        self.lib = L.load()
        self.handle = C.c_void_p()
        w = (C.c_int * 4)(*WIDTHS)
        L.check(self.lib.ksh_mlp_create(w, seed, C.byref(self.handle)))
        self.losses = None

    def __del__(self):
        try:
            if self.handle:
                self.lib.ksh_mlp_destroy(self.handle)
                self.handle = C.c_void_p()
        except Exception:
            pass

    def fit(self, X, Y, epochs=200, batch=256, lr=1e-3, weight_decay=0.0, seed=7):
        X = np.asarray(X, dtype=np.float64)
        Y = np.asarray(Y, dtype=np.float64)
        self.in_mean, self.out_mean = X.mean(0), Y.mean(0)
        self.in_std = np.maximum(X.std(0), 1e-12 + 1e-6 * np.abs(X).max(0))
        self.out_std = np.maximum(Y.std(0), 1e-12 + 1e-6 * np.abs(Y).max(0))
        L.check(self.lib.ksh_mlp_set_normalization(self.handle, _dp(self.in_mean), _dp(self.in_std),
                                                   _dp(self.out_mean), _dp(self.out_std)))
        Xn = np.ascontiguousarray((X - self.in_mean) / self.in_std, dtype=np.float32)
        Yn = np.ascontiguousarray((Y - self.out_mean) / self.out_std, dtype=np.float32)
        batch = min(batch, len(Xn))
        self.losses = np.zeros(epochs)
        L.check(self.lib.ksh_mlp_train(self.handle, _fp(Xn), _fp(Yn), len(Xn), epochs, batch, lr, weight_decay, seed,
                                       _dp(self.losses)))
        return self.losses

    def predict(self, X):
        X = np.ascontiguousarray(X, dtype=np.float64).reshape(-1, IN)
        Y = np.zeros((len(X), OUT))
        L.check(self.lib.ksh_mlp_predict(self.handle, _dp(X), len(X), _dp(Y)))
        return Y
