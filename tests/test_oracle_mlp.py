"""Pinning the binary64 proposal-network oracle (oracle/vr_mlp.c): central
finite differences for every parameter block, softmax normalisation, and
Adam's closed-form first step (Kingma & Ba 2015)."""
import numpy as np
import pytest

import oracle as O


def _params(rng):
    n = O.mlp_num_params()
    assert n == 9 * 64 + 64 + 64 * 64 + 64 + 256 * 64 + 256 == 21440
    return rng.normal(scale=0.3, size=n)


def _batch(rng, B):
    x = np.zeros((B, 9))
    x[:, :3] = rng.random((B, 3))
    x[np.arange(B), 3 + rng.integers(0, 6, B)] = 1.0
    t = rng.integers(0, 256, B).astype(np.uint32)
    w = rng.random(B) + 0.5
    return x, t, w


def test_softmax_rows_sum_to_one():
    rng = np.random.default_rng(0)
    P = _params(rng)
    x, _, _ = _batch(rng, 16)
    z, q = O.mlp_forward(P, x)
    assert np.allclose(q.sum(axis=1), 1.0, atol=1e-12)
    assert np.allclose(q, np.exp(z - z.max(1, keepdims=True)) / np.exp(z - z.max(1, keepdims=True)).sum(1, keepdims=True))


@pytest.mark.parametrize("block", ["W1", "b1", "W2", "b2", "W3", "b3"])
def test_gradient_matches_central_differences(block):
    rng = np.random.default_rng(1)
    P = _params(rng)
    x, t, w = _batch(rng, 12)
    loss, g = O.mlp_loss_grad(P, x, t, w)
    offs = {"W1": (0, 576), "b1": (576, 64), "W2": (640, 4096), "b2": (4736, 64), "W3": (4800, 16384),
            "b3": (21184, 256)}[block]
    idx = rng.choice(np.arange(offs[0], offs[0] + offs[1]), size=12, replace=False)
    h = 1e-6
    for i in idx:
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += h
        Pm[i] -= h
        num = (O.mlp_loss_grad(Pp, x, t, w)[0] - O.mlp_loss_grad(Pm, x, t, w)[0]) / (2 * h)
        assert g[i] == pytest.approx(num, rel=1e-5, abs=1e-9), (block, i)


def test_adam_first_step_closed_form():
    """After one step from zero moments, m_hat = g and v_hat = g^2, so each
    parameter moves by -lr * g / (|g| + eps)."""
    rng = np.random.default_rng(2)
    n = 1000
    P = rng.normal(size=n)
    g = rng.normal(size=n)
    m = np.zeros(n)
    v = np.zeros(n)
    P0 = P.copy()
    O.adam_step(P, g, m, v, 1, lr=1e-3)
    assert np.allclose(P - P0, -1e-3 * g / (np.abs(g) + 1e-8), rtol=1e-9, atol=1e-15)
    assert np.allclose(m, 0.1 * g) and np.allclose(v, 0.001 * g * g)


def test_training_decreases_the_loss():
    """Adam on a fixed batch lowers the weighted cross-entropy."""
    rng = np.random.default_rng(3)
    P = _params(rng) * 0.3
    x, t, w = _batch(rng, 64)
    t = (np.floor(x[:, 0] * 4).astype(np.uint32) * 64 + np.argmax(x[:, 3:], 1).astype(np.uint32))  # learnable target
    m, v = np.zeros_like(P), np.zeros_like(P)
    first = O.mlp_loss_grad(P, x, t, w)[0]
    for step in range(1, 61):
        loss, g = O.mlp_loss_grad(P, x, t, w)
        O.adam_step(P, g, m, v, step, lr=1e-2)
    assert O.mlp_loss_grad(P, x, t, w)[0] < 0.5 * first
