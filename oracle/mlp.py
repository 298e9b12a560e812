"""CPU oracle of the Fourier-feature MLP surrogate (test infrastructure only).

The surrogate has no counterpart in the reference: it is this library's own
network, 3 -> 2F Fourier features -> 64 SiLU -> 64 SiLU -> 1, used to choose
which bricks of a grid are evaluated exactly.  This module restates its
forward pass and the gradient of the mean squared error in numpy, in two
variants:

  * ``exact``: float64 throughout.  Pinned against central finite
    differences (tests/test_oracle.py), which catch a dropped term, a wrong
    sign or a transposed operand in the backward formulas.
  * ``bf16``: the device's rounding points, i.e. BF16 (round to nearest even)
    for every matrix-core operand -- features, weights W1 and W2, the layer-1
    activations, and the back-propagated dz2 / dz1 -- with float64
    accumulation standing in for the FP32 accumulators.  The device results
    are compared with this variant within tolerances stated in the tests.

Parameter layout (floats): W1[64][96], b1[64], W2[64][64], b2[64], w3[64], b3.
Features of a point p in the cube [lo, lo + size]: x = (p - lo) * 2/size - 1,
phase_f = B_f . x (in revolutions), feature f = sin(2 pi phase_f),
feature F + f = cos(2 pi phase_f).
"""

import numpy as np

FREQ = 48
IN = 2 * FREQ
HIDDEN = 64
W1 = 0
B1 = W1 + HIDDEN * IN
W2 = B1 + HIDDEN
B2 = W2 + HIDDEN * HIDDEN
W3 = B2 + HIDDEN
B3 = W3 + HIDDEN
PARAMS = B3 + 1


def bf16(a):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def unpack(params):
    p = np.asarray(params, dtype=np.float64)
    return (p[W1:B1].reshape(HIDDEN, IN), p[B1:W2], p[W2:B2].reshape(HIDDEN, HIDDEN), p[B2:W3], p[W3:B3], p[B3])


def features(xyz, fourier, lo, size, rounded):
    x = (np.asarray(xyz, dtype=np.float64) - np.asarray(lo, dtype=np.float64)) * (2.0 / size) - 1.0
    ph = x @ np.asarray(fourier, dtype=np.float64).reshape(FREQ, 3).T
    f = np.concatenate([np.sin(2 * np.pi * ph), np.cos(2 * np.pi * ph)], axis=1)
    return bf16(f).astype(np.float64) if rounded else f


def _r(a, rounded):
    return bf16(a).astype(np.float64) if rounded else a


def sigmoid(z):
    return 1.0 / (1.0 + np.exp(-z))


def forward(params, fourier, lo, size, xyz, rounded=False):
    """Returns y and the intermediates (f, z1, a1, z2, a2)."""
    w1, b1, w2, b2, w3, b3 = unpack(params)
    f = features(xyz, fourier, lo, size, rounded)
    z1 = f @ _r(w1, rounded).T + b1
    a1 = _r(z1 * sigmoid(z1), rounded)
    z2 = a1 @ _r(w2, rounded).T + b2
    a2 = z2 * sigmoid(z2)
    y = a2 @ w3 + b3
    return y, (f, z1, a1, z2, a2)


def loss_and_grad(params, fourier, lo, size, xyz, target, clamp, rounded=False):
    """Mean squared error against clamp(target) and its parameter gradient."""
    w1, b1, w2, b2, w3, b3 = unpack(params)
    y, (f, z1, a1, z2, a2) = forward(params, fourier, lo, size, xyz, rounded)
    t = np.clip(np.asarray(target, dtype=np.float64), -clamp, clamp)
    n = y.shape[0]
    e = y - t
    loss = float(np.mean(e * e))
    dy = 2.0 * e / n
    s2 = sigmoid(z2)
    dz2 = (dy[:, None] * w3[None, :]) * (s2 * (1 + z2 * (1 - s2)))
    dz2r = _r(dz2, rounded)
    g = np.zeros(PARAMS)
    g[W3:B3] = a2.T @ dy
    g[B3] = dy.sum()
    g[B2:W3] = dz2.sum(axis=0)
    g[W2:B2] = (dz2r.T @ a1).reshape(-1)
    da1 = dz2r @ _r(w2, rounded)
    s1 = sigmoid(z1)
    dz1 = da1 * (s1 * (1 + z1 * (1 - s1)))
    g[B1:W2] = dz1.sum(axis=0)
    g[W1:B1] = (_r(dz1, rounded).T @ f).reshape(-1)
    return loss, g


def loss_only(params, fourier, lo, size, xyz, target, clamp):
    y, _ = forward(params, fourier, lo, size, xyz, False)
    t = np.clip(np.asarray(target, dtype=np.float64), -clamp, clamp)
    return float(np.mean((y - t) ** 2))
