"""The oracle's proposer network (forward, gradient, AdamW) against central
finite differences and closed forms; a planted mistake (tanh' dropped in
the first layer's backward) must fail the finite-difference check."""
import numpy as np
import pytest

import oracle

NP = oracle.MLP_NPARAMS


def rand_params(rng, scale=0.3):
    return rng.normal(0, scale, NP)


def data(rng, n):
    X = rng.normal(0, 1, (n, 12))
    T = rng.normal(0, 1, (n, 2))
    W = rng.uniform(0, 1, (n, 2))
    return X, T, W


def fd_check(lib_loss_grad, p, X, T, W, idx, h=1e-6):
    _, g = lib_loss_grad(p, X, T, W)
    errs = []
    for e in idx:
        pp, pm = p.copy(), p.copy()
        pp[e] += h
        pm[e] -= h
        fd = (lib_loss_grad(pp, X, T, W)[0] - lib_loss_grad(pm, X, T, W)[0]) / (2 * h)
        errs.append(abs(fd - g[e]) / max(1e-8, abs(fd) + abs(g[e])))
    return max(errs)


def param_sample(rng):
    # indices from every block: W1, b1, W2, b2, W3, b3
    blocks = [(0, 768), (768, 832), (832, 4928), (4928, 4992), (4992, 5120), (5120, 5122)]
    return [int(rng.integers(a, b)) for a, b in blocks for _ in range(6)]


def test_gradient_matches_finite_differences(orc):
    rng = np.random.default_rng(1)
    p = rand_params(rng)
    X, T, W = data(rng, 7)
    assert fd_check(orc.mlp_loss_grad, p, X, T, W, param_sample(rng)) < 1e-6


def test_gradient_unweighted_equals_unit_weights(orc):
    rng = np.random.default_rng(2)
    p = rand_params(rng)
    X, T, _ = data(rng, 5)
    l1, g1 = orc.mlp_loss_grad(p, X, T, None)
    l2, g2 = orc.mlp_loss_grad(p, X, T, np.ones((5, 2)))
    assert l1 == l2 and np.array_equal(g1, g2)


def test_closed_forms(orc):
    rng = np.random.default_rng(3)
    X, T, W = data(rng, 9)
    p = np.zeros(NP)
    p[5120:5122] = [0.5, -1.5]
    # all hidden units are tanh(0) = 0: y = b3 for every input
    Y = orc.mlp_forward(p, X)
    assert np.array_equal(Y, np.tile([0.5, -1.5], (9, 1)))
    loss, g = orc.mlp_loss_grad(p, X, T, W)
    e = np.array([0.5, -1.5]) - T
    assert loss == pytest.approx(np.sum(W * e * e) / 9, rel=1e-14)
    assert g[5120:5122] == pytest.approx(2 * np.sum(W * e, axis=0) / 9, rel=1e-13)
    # dL/dW3 = dy * h2 = 0, and the hidden layers get no signal through W3 = 0
    assert not np.any(g[:5120])


def test_forward_by_numpy(orc):
    rng = np.random.default_rng(4)
    p = rand_params(rng)
    X = rng.normal(0, 1, (6, 12))
    W1, b1 = p[:768].reshape(64, 12), p[768:832]
    W2, b2 = p[832:4928].reshape(64, 64), p[4928:4992]
    W3, b3 = p[4992:5120].reshape(2, 64), p[5120:]
    ref = np.tanh(np.tanh(X @ W1.T + b1) @ W2.T + b2) @ W3.T + b3
    assert np.allclose(orc.mlp_forward(p, X), ref, rtol=1e-13, atol=1e-13)


def test_adam_first_steps(orc):
    rng = np.random.default_rng(5)
    p = rng.normal(0, 1, NP)
    g = rng.normal(0, 1, NP)
    m, v = np.zeros(NP), np.zeros(NP)
    q = p.copy()
    orc.mlp_adam(q, g, m, v, 1e-2, 0.9, 0.999, 1e-8, 0.0, 1)
    # step 1: mhat = g, vhat = g^2, so the update is lr * g / (|g| + eps)
    assert np.allclose(q, p - 1e-2 * g / (np.abs(g) + 1e-8), rtol=1e-12, atol=1e-15)
    # decoupled weight decay with zero gradient shrinks by (1 - lr * wd)
    m, v = np.zeros(NP), np.zeros(NP)
    q = p.copy()
    orc.mlp_adam(q, np.zeros(NP), m, v, 1e-2, 0.9, 0.999, 1e-8, 0.1, 1)
    assert np.allclose(q, p * (1 - 1e-3), rtol=1e-14)


def test_adam_minimizes_quadratic(orc):
    # AdamW on loss = sum w (y - t)^2 with a learnable b3 only reaches the
    # weighted mean of the targets
    rng = np.random.default_rng(6)
    X, T, W = data(rng, 16)
    p = np.zeros(NP)
    m, v = np.zeros(NP), np.zeros(NP)
    for step in range(1, 3001):
        _, g = orc.mlp_loss_grad(p, X, T, W)
        g[:5120] = 0
        orc.mlp_adam(p, g, m, v, 1e-2, 0.9, 0.999, 1e-8, 0.0, step)
    assert p[5120:] == pytest.approx(np.sum(W * T, axis=0) / np.sum(W, axis=0), abs=1e-4)


def test_planted_backward_is_caught():
    bad = oracle.planted(4)
    rng = np.random.default_rng(7)
    p = rand_params(rng, 0.6)
    X, T, W = data(rng, 7)
    with oracle.use(bad):
        err = fd_check(oracle.mlp_loss_grad, p, X, T, W, list(range(0, 768, 37)))
    assert err > 1e-2
