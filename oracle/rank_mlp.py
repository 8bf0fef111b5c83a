"""CPU oracle of the rank predictor MLP (test infrastructure only).

The network of rcs_amd/csrc/mlp.hip restated in numpy: standardised 8 inputs -> 64 GELU ->
64 GELU -> 1 output, loss = mean (y - t)^2. The forward pass uses the weights rounded to
BF16 (round to nearest even), as the GPU kernels do, and float64 arithmetic, so the GPU
(FP32 accumulation) agrees to FP32 rounding. Parameter layout: W1[64][8], b1[64],
W2[64][64], b2[64], W3[64], b3 (4801 floats). It has no counterpart in the reference.

Pinned by tests/test_oracle.py::test_rank_mlp_oracle_gradient_finite_differences (central
differences of the loss in float64) and by the GELU closed form at known points.
"""
import math

import numpy as np

IN, H = 8, 64
NP = H * IN + H + H * H + H + H + 1


def bf16_round(a):
    """float32 -> nearest-even BF16, returned as float32."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + H * IN].reshape(H, IN); o += H * IN
    b1 = p[o:o + H]; o += H
    W2 = p[o:o + H * H].reshape(H, H); o += H * H
    b2 = p[o:o + H]; o += H
    W3 = p[o:o + H]; o += H
    b3 = p[o]
    return W1, b1, W2, b2, W3, b3


_erf = np.vectorize(math.erf)


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z / math.sqrt(2.0)))


def gelu_d(z):
    return 0.5 * (1.0 + _erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def forward(params, mean, scale, X, bf16=True):
    p = bf16_round(params) if bf16 else np.asarray(params, dtype=np.float32)
    W1, b1, W2, b2, W3, b3 = unpack(p)
    xn = (np.asarray(X, np.float64) - mean) / scale
    z1 = xn @ W1.T + b1
    h1 = gelu(z1)
    z2 = h1 @ W2.T + b2
    h2 = gelu(z2)
    return h2 @ W3 + b3


def loss_and_grad(params, mean, scale, X, t, bf16=True):
    """loss = mean (y - t)^2 and its gradient (float64) with respect to the parameters,
    evaluated at the (BF16-rounded) weights."""
    p = bf16_round(params) if bf16 else np.asarray(params, dtype=np.float64)
    W1, b1, W2, b2, W3, b3 = unpack(p)
    n = len(X)
    xn = (np.asarray(X, np.float64) - mean) / scale
    z1 = xn @ W1.T + b1
    h1 = gelu(z1)
    z2 = h1 @ W2.T + b2
    h2 = gelu(z2)
    y = h2 @ W3 + b3
    err = y - np.asarray(t, np.float64)
    dy = 2.0 * err / n
    dz2 = (dy[:, None] * W3[None, :]) * gelu_d(z2)
    dz1 = (dz2 @ W2) * gelu_d(z1)
    g = np.concatenate([(dz1.T @ xn).ravel(), dz1.sum(0), (dz2.T @ h1).ravel(), dz2.sum(0), h2.T @ dy,
                        [dy.sum()]])
    return float(np.mean(err ** 2)), g
