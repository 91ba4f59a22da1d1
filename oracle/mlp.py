"""CPU oracle of the MLP continuation regressor (test infrastructure only:
imported by tests/ and nothing on the product path).

The network is 3-64-64-1 with tanh-form GELU. Parameters are one float32
vector of 4,481 entries:

    W1 [64][3] | b1 [64] | W2 [64 out][64 in] | b2 [64] | W3 [64] | b3

The GPU computes with bf16 copies of W1, W2, W3, bf16 hidden activations
and bf16 layer-2 gradients as MFMA operands, accumulating in FP32;
`forward_backward(..., bf16=True)` rounds at exactly those points.
`bf16=False` is the plain float64 network, which the tests pin with
central finite differences.

Adam follows Kingma & Ba (2015), Algorithm 1, with beta1 = 0.9,
beta2 = 0.999, eps = 1e-8 and bias correction; moments restart at zero for
every exercise date.

There is no reference counterpart: kiyosi's LSMC regression is the
quadratic least-squares fit (engines/monte_carlo/american.rs:1693-1749),
which this regressor replaces.
"""

import math

import numpy as np

H = 64
W1, B1, W2, B2, W3, B3, COUNT = 0, 192, 256, 4352, 4416, 4480, 4481
GELU_K0 = math.sqrt(2.0 / math.pi)
GELU_K1 = 0.044715


def bf16(a):
    """Round float32 values to the nearest bf16 (ties to even), returned as float32."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def gelu(x):
    """0.5 x (1 + tanh(k0 (x + k1 x^3))) and its derivative."""
    t = np.tanh(GELU_K0 * (x + GELU_K1 * x ** 3))
    g = 0.5 * x * (1.0 + t)
    d = 0.5 * (1.0 + t) + 0.5 * x * (1.0 - t * t) * GELU_K0 * (1.0 + 3.0 * GELU_K1 * x * x)
    return g, d


def unpack(p):
    p = np.asarray(p)
    return (p[W1:B1].reshape(H, 3), p[B1:W2], p[W2:B2].reshape(H, H), p[B2:W3], p[W3:B3], p[B3])


def forward(p, x, bf16_ops=True):
    w1, b1, w2, b2, w3, b3 = unpack(np.asarray(p, dtype=np.float64))
    rnd = (lambda a: bf16(a).astype(np.float64)) if bf16_ops else (lambda a: a)
    w1r, w2r, w3r = rnd(w1), rnd(w2), rnd(w3)
    z1 = x[:, :3] @ w1r.T + b1
    h1, d1 = gelu(z1)
    h1 = rnd(h1)
    z2 = h1 @ w2r.T + b2
    h2, d2 = gelu(z2)
    h2 = rnd(h2)
    out = h2 @ w3r + b3
    return out, (w1r, w2r, w3r, z1, h1, d1, z2, h2, d2)


def forward_backward(p, samples, n_batch=None, bf16_ops=True):
    """Gradient of sum((out - y)^2) / n_batch over `samples` rows
    (u0, u1, u2, y); returns (gradient vector, loss sum / n_batch)."""
    x = np.asarray(samples, dtype=np.float64)
    n_batch = len(x) if n_batch is None else n_batch
    out, (w1r, w2r, w3r, z1, h1, d1, z2, h2, d2) = forward(p, x, bf16_ops)
    rnd = (lambda a: bf16(a).astype(np.float64)) if bf16_ops else (lambda a: a)
    r = out - x[:, 3]
    g = 2.0 * r / n_batch
    grad = np.zeros(COUNT)
    grad[W3:B3] = g @ h2
    grad[B3] = g.sum()
    dz2 = g[:, None] * w3r[None, :] * d2
    grad[B2:W3] = dz2.sum(0)
    dz2 = rnd(dz2)
    grad[W2:B2] = (dz2.T @ h1).ravel()
    dz1 = (dz2 @ w2r) * d1
    grad[B1:W2] = dz1.sum(0)
    grad[W1:B1] = (dz1.T @ x[:, :3]).ravel()
    return grad, float((r * r).sum() / n_batch)


def adam(p, m, v, grad, step, lr):
    """One Adam update in float32 (step counts from 1)."""
    p, m, v = (np.asarray(a, dtype=np.float32).copy() for a in (p, m, v))
    g = np.asarray(grad, dtype=np.float32)
    m = np.float32(0.9) * m + np.float32(0.1) * g
    v = np.float32(0.999) * v + np.float32(0.001) * g * g
    c1 = np.float32(1.0 / (1.0 - 0.9 ** step))
    c2 = np.float32(1.0 / (1.0 - 0.999 ** step))
    p = p - np.float32(lr) * (m * c1) / (np.sqrt(v * c2) + np.float32(1e-8))
    return p.astype(np.float32), m, v


def train(p, samples, batch, steps, lr):
    """`steps` Adam steps over consecutive batches of `samples` (wrapping),
    starting from zero moments. Gradients use the bf16 operand rounding."""
    p = np.asarray(p, dtype=np.float32).copy()
    m = np.zeros(COUNT, np.float32)
    v = np.zeros(COUNT, np.float32)
    n = len(samples)
    begin = 0
    for t in range(1, steps + 1):
        rows = samples[begin:begin + batch]
        grad, _ = forward_backward(p, rows, len(rows))
        p, m, v = adam(p, m, v, grad, t, lr)
        begin = begin + batch if begin + batch < n else 0
    return p, m, v
