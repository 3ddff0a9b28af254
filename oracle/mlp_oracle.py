"""CPU oracle (TEST INFRASTRUCTURE ONLY) of the face-flux surrogate network.

The reference has no learned component; this module defines the network the
GPU kernels implement and restates its forward pass, backward pass and Adam
update in numpy so the GPU kernels can be checked against it:

  MLP 16 -> 128 -> 128 -> 64 -> 8, SiLU between layers, y = x W + b.

The GPU keeps FP32 master weights, computes with BF16 weights and BF16
activations and accumulates in FP32.  ``rounding=True`` emulates exactly
where the GPU rounds to BF16 (inputs, weights, hidden activations, output
gradient and the pre-activation gradients that feed the next GEMM); the
products and sums themselves are formed here in FP64, so the GPU differs
from this oracle only by FP32 accumulation order.  ``rounding=False`` is the
plain FP64 network, pinned by finite differences in the tests.

Only tests/ and the benchmark's checks import this module.
"""
from __future__ import annotations

import numpy as np

DIMS = (16, 128, 128, 64, 8)


def n_params(dims=DIMS):
    return sum(dims[i] * dims[i + 1] + dims[i + 1] for i in range(len(dims) - 1))


def param_shapes(dims=DIMS):
    out = []
    for i in range(len(dims) - 1):
        out.append(("W%d" % (i + 1), (dims[i], dims[i + 1])))
        out.append(("b%d" % (i + 1), (dims[i + 1],)))
    return out


def unflatten(flat, dims=DIMS):
    """Flat FP32 parameter vector (W1, b1, W2, b2, ... with W row-major
    [fan_in][fan_out]) -> dict of arrays (views)."""
    p, o = {}, 0
    for name, shp in param_shapes(dims):
        n = int(np.prod(shp))
        p[name] = flat[o:o + n].reshape(shp)
        o += n
    assert o == flat.size
    return p


def flatten(p, dims=DIMS):
    return np.concatenate([np.asarray(p[n], np.float64).ravel()
                           for n, _ in param_shapes(dims)])


def bf16(x):
    """Round to the nearest BF16 value (ties to even), returned as float64."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    r = ((u + np.uint64(0x7FFF) + ((u >> np.uint64(16)) & np.uint64(1)))
         >> np.uint64(16)) << np.uint64(16)
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(flat, X, rounding=True, keep=False):
    """Y = MLP(X).  With keep=True also returns the per-layer inputs A_l and
    pre-activations Z_l needed by the backward pass."""
    p = unflatten(np.asarray(flat, np.float64))
    r = bf16 if rounding else (lambda a: np.asarray(a, np.float64))
    nl = len(DIMS) - 1
    A = r(X)
    As, Zs = [A], []
    for l in range(1, nl + 1):
        Z = A @ r(p["W%d" % l]) + p["b%d" % l]
        Zs.append(Z)
        if l < nl:
            A = r(silu(Z))
            As.append(A)
    Y = Zs[-1]
    if keep:
        return Y, As, Zs
    return Y


def loss_and_grad(flat, X, T, rounding=True):
    """Mean-square loss over all outputs and its gradient with respect to the
    flat parameter vector (same layout as flat)."""
    Y, As, Zs = forward(flat, X, rounding, keep=True)
    p = unflatten(np.asarray(flat, np.float64))
    r = bf16 if rounding else (lambda a: np.asarray(a, np.float64))
    B = X.shape[0]
    nl = len(DIMS) - 1
    diff = Y - T
    loss = float(np.mean(diff * diff))
    dZ = r(2.0 * diff / (B * DIMS[-1]))
    g = {}
    for l in range(nl, 0, -1):
        g["W%d" % l] = As[l - 1].T @ dZ
        g["b%d" % l] = dZ.sum(axis=0)
        if l > 1:
            dA = dZ @ r(p["W%d" % l]).T
            dZ = r(dA * dsilu(Zs[l - 2]))
    return loss, flatten(g)


def init_params(seed=0, dims=DIMS):
    """Uniform(-1/sqrt(fan_in), 1/sqrt(fan_in)) weights, zero biases."""
    rng = np.random.default_rng(seed)
    p = {}
    for i in range(len(dims) - 1):
        lim = 1.0 / np.sqrt(dims[i])
        p["W%d" % (i + 1)] = rng.uniform(-lim, lim, (dims[i], dims[i + 1]))
        p["b%d" % (i + 1)] = np.zeros(dims[i + 1])
    return flatten(p).astype(np.float32)


def adam_step(w, m, v, g, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015) of FP32 master weights, in place;
    t is the 1-based step count."""
    m[:] = b1 * m + (1 - b1) * g
    v[:] = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    w[:] = w - lr * mh / (np.sqrt(vh) + eps)


def cosine_lr(lr0, step, total):
    return lr0 * 0.5 * (1.0 + np.cos(np.pi * min(step, total) / total))
