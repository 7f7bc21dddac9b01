"""CPU: the numpy MLP oracle's gradient against central finite differences of its FP64 loss."""
import numpy as np

from oracle import mlp


def test_oracle_gradient_matches_finite_differences():
    widths = (6, 5, 4, 3)
    rng = np.random.default_rng(0)
    n = sum(widths[i] * widths[i + 1] + widths[i + 1] for i in range(3))
    p = rng.standard_normal(n) * 0.5
    x = rng.standard_normal((7, widths[0]))
    t = rng.standard_normal((7, widths[-1]))
    _, g = mlp.loss_and_grad(p, x, t, widths)
    h = 1e-6
    fd = np.zeros(n)
    for i in range(n):
        e = np.zeros(n)
        e[i] = h
        fd[i] = (mlp.loss_and_grad(p + e, x, t, widths)[0] - mlp.loss_and_grad(p - e, x, t, widths)[0]) / (2 * h)
    assert np.abs(fd - g).max() <= 1e-7 * max(1.0, np.abs(g).max())


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.0e-3], dtype=np.float32)
    r = mlp.bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0          # tie -> even
    assert r[2] == np.float32(1.0 + 2 ** -7)     # above the tie -> up
    assert r[3] == -2.5
    assert abs(r[4] - 3.0e-3) <= 2 ** -8 * 3.0e-3       # half an ulp of an 8-bit significand
