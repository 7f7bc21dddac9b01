"""Test infrastructure only: CPU (numpy) restatement of the learned predictor's MLP, the oracle
the GPU kernels' forward and gradients are checked against.

Network: widths (264, 256, 256, 65), GELU (erf form) after the hidden layers, linear output,
loss = mean over batch x outputs of (y - t)^2.  Parameters packed as W0, W1, W2 (row-major
[in][out]) then b0, b1, b2, as the device library packs them.  With bf16=True every product
operand (activations and weights) is rounded to BF16 (round to nearest even) and products
accumulate in FP64, which is what the device computes up to FP32 accumulation order.

Pinned by: central finite differences of the FP64 loss (tests/test_mlp_oracle.py).
"""
import numpy as np
from scipy.special import erf

WIDTHS = (264, 256, 256, 65)


def bf16(x):
    """Round to BF16 (nearest even), returned as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def unpack(params, widths=WIDTHS):
    W, b, off = [], [], 0
    for i in range(len(widths) - 1):
        n = widths[i] * widths[i + 1]
        W.append(params[off:off + n].reshape(widths[i], widths[i + 1]))
        off += n
    for i in range(len(widths) - 1):
        b.append(params[off:off + widths[i + 1]])
        off += widths[i + 1]
    return W, b


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_prime(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def _mm(a, w, bf):
    if bf:
        return bf16(a).astype(np.float64) @ bf16(w).astype(np.float64)
    return a.astype(np.float64) @ w.astype(np.float64)


def forward(params, x, widths=WIDTHS, bf=False):
    W, b = unpack(np.asarray(params, dtype=np.float64), widths)
    h = [np.asarray(x, dtype=np.float64)]
    z = []
    for l in range(len(W)):
        zl = _mm(h[-1], W[l], bf) + b[l]
        z.append(zl)
        h.append(gelu(zl) if l + 1 < len(W) else zl)
    return h, z


def loss_and_grad(params, x, t, widths=WIDTHS, bf=False):
    params = np.asarray(params, dtype=np.float64)
    W, _ = unpack(params, widths)
    h, z = forward(params, x, widths, bf)
    y = h[-1]
    count = y.size
    loss = float(np.sum((y - t) ** 2) / count)
    dz = 2.0 * (y - t) / count
    gW, gb = [None] * len(W), [None] * len(W)
    for l in range(len(W) - 1, -1, -1):
        gW[l] = _mm(h[l].T, dz, bf)
        gb[l] = dz.sum(axis=0)
        if l > 0:
            dh = _mm(dz, W[l].T, bf)
            dz = dh * gelu_prime(z[l - 1])
    g = np.concatenate([w.ravel() for w in gW] + list(gb))
    return loss, g
