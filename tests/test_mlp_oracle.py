"""Pins the CPU oracle of the surrogate network: finite differences of the
unrounded FP64 network, the BF16 rounding against torch's bfloat16 cast, the
Adam update against its closed form on a quadratic."""
import numpy as np
import pytest
import torch

import mlp_oracle as mo


def test_param_count():
    assert mo.n_params() == 27464


def test_bf16_matches_torch():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(20000) * 10.0 ** rng.integers(-6, 6, 20000),
                        [0.0, -0.0, 1.0, 1.00390625, 1.0078125, 65504.0]])
    x = x.astype(np.float32)
    t = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(mo.bf16(x), t)


def test_gradient_finite_differences():
    """Analytic backward of the FP64 network equals central differences."""
    rng = np.random.default_rng(1)
    w = mo.init_params(seed=3).astype(np.float64)
    w += rng.normal(0, 0.05, w.size)        # non-zero biases too
    X = rng.normal(0, 1, (7, 16))
    T = rng.normal(0, 1, (7, 8))
    loss, g = mo.loss_and_grad(w, X, T, rounding=False)
    idx = rng.choice(w.size, 60, replace=False)
    h = 1e-6
    for i in idx:
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        fd = (mo.loss_and_grad(wp, X, T, rounding=False)[0] -
              mo.loss_and_grad(wm, X, T, rounding=False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, i


def test_rounded_network_close_to_fp64():
    """BF16 rounding changes outputs and gradients only at the BF16 level."""
    rng = np.random.default_rng(2)
    w = mo.init_params(seed=4).astype(np.float64)
    X = rng.normal(0, 1, (64, 16))
    T = rng.normal(0, 1, (64, 8))
    y0 = mo.forward(w, X, rounding=False)
    y1 = mo.forward(w, X, rounding=True)
    assert np.abs(y1 - y0).max() < 3e-2 * np.abs(y0).max()
    _, g0 = mo.loss_and_grad(w, X, T, rounding=False)
    _, g1 = mo.loss_and_grad(w, X, T, rounding=True)
    assert np.linalg.norm(g1 - g0) < 5e-2 * np.linalg.norm(g0)


def test_adam_closed_form():
    """First Adam step moves every weight by lr*sign(g) (bias-corrected
    moments m/(1-b1) = g, v/(1-b2) = g^2)."""
    w = np.array([1.0, -2.0, 3.0])
    g = np.array([0.5, -4.0, 1e-3])
    m, v = np.zeros(3), np.zeros(3)
    mo.adam_step(w, m, v, g, 1, 1e-2, eps=0.0)
    assert np.allclose(w, [1.0 - 1e-2, -2.0 + 1e-2, 3.0 - 1e-2])


def test_cosine_lr():
    assert mo.cosine_lr(1e-3, 0, 100) == 1e-3
    assert abs(mo.cosine_lr(1e-3, 50, 100) - 5e-4) < 1e-12
    assert mo.cosine_lr(1e-3, 100, 100) == 0.0
