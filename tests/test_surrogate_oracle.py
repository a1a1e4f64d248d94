"""Pins of the MLP oracle (oracle/mlp_ref.py) on the CPU: bfloat16 rounding
against known bit patterns, backpropagation against central finite
differences of the loss, and a one-layer network against its closed-form
gradient."""
import numpy as np

from oracle import mlp_ref


def test_bf16_round_known_values():
    x = np.array([1.0, 1.00390625, 1.005859375, 1.0078125, -2.5, 3.140625, 1e-40, 65504.0])
    r = mlp_ref.bf16_round(x)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: rounds to even (1.0)
    assert r[0] == 1.0 and r[1] == 1.0
    # 1 + 1.5 * 2^-8 rounds up to 1 + 2^-7
    assert r[2] == 1.0078125 and r[3] == 1.0078125
    assert r[4] == -2.5 and r[5] == 3.140625
    # relative error bound 2^-8 (8 significant bits) for normal numbers
    y = np.random.default_rng(3).normal(size=1000) * 10.0
    assert np.all(np.abs(mlp_ref.bf16_round(y) - y) <= 2.0 ** -8 * np.abs(y) * (1 + 1e-12))


def test_gradient_against_finite_differences():
    rng = np.random.default_rng(11)
    dims = [3, 7, 5, 2]
    P = sum(dims[i + 1] * dims[i] + dims[i + 1] for i in range(3))
    p = rng.normal(size=P) * 0.7
    x = rng.normal(size=(9, 3))
    t = rng.normal(size=(9, 2))
    loss, g = mlp_ref.loss_and_grad(p, dims, x, t)
    h = 1e-6
    fd = np.zeros(P)
    for i in range(P):
        e = np.zeros(P)
        e[i] = h
        fd[i] = (mlp_ref.loss_and_grad(p + e, dims, x, t)[0] - mlp_ref.loss_and_grad(p - e, dims, x, t)[0]) / (2 * h)
    assert np.abs(fd - g).max() < 1e-7 * max(1.0, np.abs(g).max())
    assert abs(loss - np.mean((mlp_ref.forward(p, dims, x) - t) ** 2)) < 1e-14


def test_linear_layer_closed_form():
    rng = np.random.default_rng(12)
    dims = [4, 3]
    W = rng.normal(size=(3, 4))
    b = rng.normal(size=3)
    p = np.concatenate([W.reshape(-1), b])
    x = rng.normal(size=(6, 4))
    t = rng.normal(size=(6, 3))
    loss, g = mlp_ref.loss_and_grad(p, dims, x, t)
    y = x @ W.T + b
    dW = 2.0 / y.size * (y - t).T @ x
    db = 2.0 / y.size * (y - t).sum(0)
    assert np.allclose(g, np.concatenate([dW.reshape(-1), db]), rtol=1e-13, atol=1e-14)
    assert np.isclose(loss, np.mean((y - t) ** 2), rtol=1e-14)
