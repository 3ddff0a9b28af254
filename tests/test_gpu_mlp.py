"""The surrogate network's GPU kernels (matrix-core forward, backward, Adam)
against the CPU oracle of the same network with the same BF16 rounding
points.  The oracle forms products and sums in FP64, the GPU accumulates in
FP32, so results agree to FP32 accumulation error except where an
intermediate lands on the other side of a BF16 rounding boundary."""
import numpy as np
import pytest

import mlp_oracle as mo
from batsrus_amd.core import Mlp

pytestmark = pytest.mark.gpu


def _params(seed):
    rng = np.random.default_rng(seed)
    w = mo.init_params(seed=seed).astype(np.float64)
    w += rng.normal(0, 0.02, w.size)            # non-zero biases
    return w.astype(np.float32)


def _bf16_ulp(x):
    """spacing of the BF16 values around x (8 significant bits)"""
    e = np.floor(np.log2(np.maximum(np.abs(x), 1e-38)))
    return 2.0 ** (e - 7)


def _forward_bound(w, X):
    """Per row and output, a bound on |GPU - oracle| of the forward pass.
    Each layer's pre-activation differs by the FP32 accumulation error
    (at most n * 2^-24 * sum|a w| for n terms of exact BF16 products) plus
    what earlier differences carry through |W|; the activation silu(z),
    rounded to BF16, can then differ by one BF16 spacing, and only where
    the oracle's value lies within that difference (plus the fast sigmoid's
    3e-7 relative error) of a BF16 rounding midpoint."""
    Y, As, Zs = mo.forward(w, X, rounding=True, keep=True)
    p = mo.unflatten(np.asarray(w, np.float64))
    u = 2.0 ** -24
    dA = np.zeros_like(As[0])
    nl = len(mo.DIMS) - 1
    for l in range(1, nl + 1):
        Wa = np.abs(mo.bf16(p["W%d" % l]))
        n = Wa.shape[0] + 1
        dZ = n * u * (np.abs(As[l - 1]) @ Wa + np.abs(p["b%d" % l])) + dA @ Wa
        if l == nl:
            return Y, dZ
        z = Zs[l - 1]
        s = mo.silu(z)
        delta = 1.1 * dZ + 3e-7 * (np.abs(s) + np.abs(z))
        ulp = _bf16_ulp(s)
        # distance of s to the nearest midpoint between BF16 neighbours
        k = np.floor(np.abs(s) / ulp)
        dist = np.abs(np.abs(s) - (k + 0.5) * ulp)
        dist = np.minimum(dist, np.abs(np.abs(s) - (k - 0.5) * ulp))
        flip = dist <= delta
        dA = np.where(flip, ulp, 0.0)


@pytest.mark.parametrize("rows", [1, 16, 1003])
def test_forward_matches_oracle(rows):
    """Every row within the bound that FP32 accumulation and BF16 rounding
    allow (_forward_bound); most rows carry no BF16 flip at all."""
    print(f"case mlp forward rows={rows}", flush=True)
    rng = np.random.default_rng(rows)
    w = _params(1)
    X = rng.normal(0, 1, (rows, 16)).astype(np.float32)
    with Mlp(w) as m:
        Y = m.forward(X)
    Yo, bound = _forward_bound(w, X)
    err = np.abs(Y - Yo)
    scale = np.abs(Yo).max()
    ratio = err / (bound + 1e-7 * scale)
    print(f"  max err {err.max():.2e}, max err/bound {ratio.max():.3f}, "
          f"median bound {np.median(bound):.1e}")
    assert np.all(ratio <= 1.0)
    assert np.median(err) <= 1e-4 * scale


def test_gradient_matches_oracle():
    print("case mlp gradient", flush=True)
    rng = np.random.default_rng(5)
    w = _params(2)
    X = rng.normal(0, 1, (700, 16)).astype(np.float32)
    T = rng.normal(0, 1, (700, 8)).astype(np.float32)
    with Mlp(w) as m:
        loss, g = m.loss_grad(X, T)
    lo, go = mo.loss_and_grad(w, X, T, rounding=True)
    assert abs(loss - lo) <= 1e-4 * lo
    # per parameter block: relative L2 error and direction
    for name, shp in mo.param_shapes():
        o = 0
        for n2, s2 in mo.param_shapes():
            if n2 == name:
                break
            o += int(np.prod(s2))
        n = int(np.prod(shp))
        a, b = g[o:o + n].astype(np.float64), go[o:o + n]
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        cos = a @ b / (np.linalg.norm(a) * np.linalg.norm(b))
        print(f"  {name}: rel L2 {rel:.2e} cos {cos:.6f}")
        assert rel < 2e-2 and cos > 0.9995, name


def test_adam_step_matches_oracle():
    """One training step on the whole set (no hold-out): the new weights are
    the oracle's Adam update of the old ones with the GPU's gradient."""
    print("case mlp adam", flush=True)
    rng = np.random.default_rng(6)
    w0 = _params(3)
    X = rng.normal(0, 1, (512, 16)).astype(np.float32)
    T = rng.normal(0, 1, (512, 8)).astype(np.float32)
    with Mlp(w0) as m:
        m.set_data(X, T)
        m.train(1, batch=512, lr=1e-3, hold=0.0, seed=1)
        w1, g = m.params()
    wo = w0.astype(np.float64).copy()
    mom, vel = np.zeros_like(wo), np.zeros_like(wo)
    mo.adam_step(wo, mom, vel, g.astype(np.float64), 1, 1e-3)
    assert np.abs(w1 - wo).max() <= 1e-6
    # and the gradient of that step is the oracle's gradient
    _, go = mo.loss_and_grad(w0, X, T, rounding=True)
    assert np.linalg.norm(g - go) < 2e-2 * np.linalg.norm(go)


def test_training_learns_a_smooth_map():
    """Training on a smooth synthetic map reduces the held-out error far
    below the untrained network's, deterministically for a fixed seed."""
    print("case mlp training", flush=True)
    rng = np.random.default_rng(7)
    n = 40000
    X = rng.uniform(-1, 1, (n, 16)).astype(np.float32)
    A = rng.normal(0, 0.5, (16, 8))
    T = (np.tanh(X @ A) + 0.1 * (X[:, :8] ** 2)).astype(np.float32)
    res = []
    for _ in range(2):
        with Mlp(mo.init_params(seed=9)) as m:
            m.set_data(X, T)
            _, e0 = m.train(0, batch=4096, hold=0.1, seed=3)
            loss, e1 = m.train(40, batch=1024, lr=3e-3, hold=0.1, seed=3)
            res.append((e0[0], e1[0], m.params()[0]))
    print(f"  held-out rel L2: untrained {res[0][0]:.3f} trained {res[0][1]:.4f}")
    # (a CPU torch run of the same network and schedule reaches 0.029)
    assert res[0][1] < 0.1 * res[0][0] and res[0][1] < 0.06
    assert np.array_equal(res[0][2], res[1][2])   # deterministic


def test_graph_replayed_training_matches_eager_steps():
    """Training replays one captured step per minibatch, with the learning
    rate and Adam's bias corrections read on the device.  Five full-batch
    steps (one per epoch, so the host uploads a permutation and reads the
    loss between replays) equal five eager steps: the library's gradient
    of the current weights and the oracle's Adam update with the cosine
    schedule."""
    print("case mlp graph training", flush=True)
    rng = np.random.default_rng(11)
    w0 = _params(5)
    X = rng.normal(0, 1, (512, 16)).astype(np.float32)
    T = rng.normal(0, 1, (512, 8)).astype(np.float32)
    K, lr0 = 5, 2e-3
    with Mlp(w0) as m:
        m.set_data(X, T)
        m.train(K, batch=512, lr=lr0, hold=0.0, seed=2)
        wg = m.params()[0]
    w = w0.astype(np.float64).copy()
    mom, vel = np.zeros_like(w), np.zeros_like(w)
    with Mlp(w0) as m:
        for q in range(K):
            m.set_params(w.astype(np.float32))
            _, g = m.loss_grad(X, T)
            mo.adam_step(w, mom, vel, g.astype(np.float64), q + 1,
                         mo.cosine_lr(lr0, q, K))
    d = np.abs(wg - w)
    bad = float(np.mean(d > 1e-5))
    print(f"  |graph - eager|: median {np.median(d):.2e}, max {d.max():.2e}, "
          f"fraction above 1e-5 {bad:.1e} (step size {lr0:.0e})")
    # the minibatch rows are gathered in another order, so the FP32 gradient
    # sums differ at rounding level; Adam's normalised step amplifies that
    # only for components whose gradient is itself at rounding level.  A
    # wrong learning rate, bias correction or step count moves every
    # parameter by a fraction of the step size.
    assert np.median(d) <= 1e-7 and bad < 1e-3


def test_graph_training_equals_kernel_by_kernel_training():
    """The replayed step graph and the same kernels launched one by one give
    the same bits over several epochs of minibatches (host work between the
    epochs' replays)."""
    print("case mlp graph vs eager training", flush=True)
    rng = np.random.default_rng(12)
    X = rng.uniform(-1, 1, (6000, 16)).astype(np.float32)
    T = np.tanh(X[:, :8] + 0.5 * X[:, 8:]).astype(np.float32)
    out = []
    for graph in (True, False):
        with Mlp(mo.init_params(seed=13)) as m:
            m.set_data(X, T)
            loss, err = m.train(3, batch=1024, lr=2e-3, hold=0.1, seed=4,
                                graph=graph)
            out.append((m.params()[0], loss, err))
    assert np.array_equal(out[0][0], out[1][0])
    assert out[0][1] == out[1][1] and np.array_equal(out[0][2], out[1][2])
