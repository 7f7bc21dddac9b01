"""Learned arc-length predictor: training data from the solver's own paths, training on the GPU by
the library's MLP kernels (ksh_mlp_*), and use by the path engine's learned predictor (the
`learned=` option of ShellProblem.run_alm / run_alm_batch).

Features of a converged state (66, ksh_predictor_features): lambda, ||U||, 64 out-of-plane mode
coefficients.  A sample's input is the features of four consecutive converged states (264), its
target the deviation of the next state from the quadratic extrapolation 3 x_t - 3 x_{t-1} +
x_{t-2} of the last three, in the mode coefficients (64) and the load factor (1) -- the
correction the predictor adds to that extrapolation.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L

FEAT, IN, OUT = 66, 264, 65
WIDTHS = (IN, 256, 256, OUT)


def dataset(paths, with_groups=False):
    """(X, Y) from per-path feature arrays [steps x 66] of converged states (and, with
    with_groups, the index of the path each sample comes from)."""
    xs, ys, gs = [], [], []
    for k, f in enumerate(paths):
        f = np.asarray(f, dtype=np.float64)
        for t in range(3, len(f) - 1):
            ext = 3.0 * f[t] - 3.0 * f[t - 1] + f[t - 2]
            xs.append(f[t - 3:t + 1].ravel())
            ys.append(np.concatenate([f[t + 1, 2:] - ext[2:], [f[t + 1, 0] - ext[0]]]))
            gs.append(k)
    X, Y = np.array(xs).reshape(-1, IN), np.array(ys).reshape(-1, OUT)
    return (X, Y, np.array(gs, dtype=np.int64)) if with_groups else (X, Y)


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

    def _normalise(self, X, Y):
        self.in_mean, self.out_mean = X.mean(0), Y.mean(0)
        self.in_std = np.maximum(X.std(0), 1e-12 + 1e-6 * np.abs(X).max(0))
        self.out_std = np.maximum(Y.std(0), 1e-12 + 1e-6 * np.abs(Y).max(0))
        self._set_normalisation(self.out_mean, self.out_std)

    def _set_normalisation(self, out_mean, out_std):
        L.check(self.lib.ksh_mlp_set_normalization(self.handle, _dp(self.in_mean), _dp(self.in_std),
                                                   _dp(np.ascontiguousarray(out_mean)),
                                                   _dp(np.ascontiguousarray(out_std))))

    def fit(self, X, Y, epochs=200, batch=256, lr=1e-3, weight_decay=0.0, seed=7, groups=None, val_frac=0.0,
            chunk=25):
        """Adam on the MSE of the standardised targets.  With `groups` (the path of each sample) and
        val_frac > 0, whole paths are held out for validation: training runs in chunks of `chunk`
        epochs, the parameters with the lowest validation error are kept, and each output block
        (the 64 mode coefficients, the load factor) is scaled by the least-squares factor in [0, 1]
        that best fits the held-out targets (0: the correction does not generalise, and the
        predictor falls back to the extrapolation it corrects).  The validation errors of the
        network and of no correction are kept in self.validation."""
        X = np.asarray(X, dtype=np.float64)
        Y = np.asarray(Y, dtype=np.float64)
        val = None
        if groups is not None and val_frac > 0:
            groups = np.asarray(groups)
            ug = np.unique(groups)
            rng = np.random.default_rng(seed)
            held = set(rng.permutation(ug)[:max(1, int(round(val_frac * len(ug))))].tolist())
            vm = np.array([gk in held for gk in groups])
            val = (X[vm], Y[vm])
            X, Y = X[~vm], Y[~vm]
        self._normalise(X, Y)
        Xn = np.ascontiguousarray((X - self.in_mean) / self.in_std, dtype=np.float32)
        Yn = np.ascontiguousarray((Y - self.out_mean) / self.out_std, dtype=np.float32)
        batch = min(batch, len(Xn))
        self.losses = np.zeros(epochs)
        if val is None:
            L.check(self.lib.ksh_mlp_train(self.handle, _fp(Xn), _fp(Yn), len(Xn), epochs, batch, lr, weight_decay,
                                           seed, _dp(self.losses)))
            return self.losses
        Xv, Yv = val
        zero = float(np.mean(((Yv - 0.0) / self.out_std) ** 2))
        best, best_params, best_epoch, done = np.inf, None, 0, 0
        npar = C.c_int64()
        L.check(self.lib.ksh_mlp_num_params(self.handle, C.byref(npar)))
        params = np.zeros(npar.value, dtype=np.float32)
        while done < epochs:
            k = min(chunk, epochs - done)
            part = np.zeros(k)
            L.check(self.lib.ksh_mlp_train(self.handle, _fp(Xn), _fp(Yn), len(Xn), k, batch, lr, weight_decay,
                                           seed + done, _dp(part)))
            self.losses[done:done + k] = part
            done += k
            err = float(np.mean(((self.predict(Xv) - Yv) / self.out_std) ** 2))
            if err < best:
                best, best_epoch = err, done
                L.check(self.lib.ksh_mlp_get_params(self.handle, _fp(params)))
                best_params = params.copy()
        L.check(self.lib.ksh_mlp_set_params(self.handle, _fp(best_params)))
        P = self.predict(Xv)
        alpha = np.ones(OUT)
        for blk in (slice(0, OUT - 1), slice(OUT - 1, OUT)):
            num, den = float(np.sum(P[:, blk] * Yv[:, blk])), float(np.sum(P[:, blk] ** 2))
            alpha[blk] = min(1.0, max(0.0, num / den)) if den > 0 else 0.0
        self._set_normalisation(alpha * self.out_mean, alpha * self.out_std)
        scaled = float(np.mean(((self.predict(Xv) - Yv) / self.out_std) ** 2))
        self.validation = {"paths_held_out": len(held), "samples": int(len(Xv)), "best_epoch": best_epoch,
                           "mse_network": best, "mse_scaled": scaled, "mse_no_correction": zero,
                           "scale_modes": float(alpha[0]), "scale_load": float(alpha[-1])}
        return self.losses

    def predict(self, X):
        X = np.ascontiguousarray(X, dtype=np.float64).reshape(-1, IN)
        Y = np.zeros((len(X), OUT))
        L.check(self.lib.ksh_mlp_predict(self.handle, _dp(X), len(X), _dp(Y)))
        return Y
