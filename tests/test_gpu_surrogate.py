"""Surrogate MLP kernels against the CPU oracle (oracle/mlp_ref.py).

The oracle rounds to bfloat16 where the kernels store bfloat16 and
accumulates in float64; the kernels accumulate in float32 and can round a
stored activation to the neighbouring bfloat16 value, so the tolerances are
relative 2-norms: forward outputs <= 1e-2, gradients <= 2e-2 (and <= 5e-2
against the unrounded float64 gradient).
"""
import sys

import numpy as np
import pytest

from oracle import mlp_ref
from sbewh_amd.surrogate import DIMS, MLPEnsemble

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def batch(rng, n):
    x = rng.uniform(-1, 1, size=(n, DIMS[0])).astype(np.float32)
    t = (np.sin(x @ rng.normal(size=(DIMS[0], DIMS[-1]))) + 0.1 * rng.normal(size=(n, DIMS[-1]))).astype(np.float32)
    return x, t


def test_forward_matches_oracle():
    say("[gpu] mlp forward 5 x (6-256-256-256-40), 300 rows")
    ens = MLPEnsemble(DIMS, 5, seed=41)
    rng = np.random.default_rng(1)
    x, _ = batch(rng, 300)
    y = ens.forward(x)
    for e in range(5):
        p = ens.get_params(e).astype(np.float64)
        ref = mlp_ref.forward(p, list(DIMS), x.astype(np.float64), bf16=True)
        assert rel2(y[e], ref) <= 1e-2, (e, rel2(y[e], ref))


def test_forward_rows_independent_of_batch_size():
    """a row's output does not depend on the other rows of the call, also when
    a call has fewer rows than an earlier one (the buffers keep their capacity)"""
    say("[gpu] mlp forward batch-size independence")
    ens = MLPEnsemble(DIMS, 5, seed=44)
    rng = np.random.default_rng(2)
    x, _ = batch(rng, 300)
    full = ens.forward(x)
    part = ens.forward(x[:5])
    one = ens.forward(x[7:8])
    assert np.array_equal(part, full[:, :5])
    assert np.array_equal(one, full[:, 7:8])


@pytest.mark.parametrize("rows", [64, 257])
def test_gradient_matches_oracle(rows):
    say(f"[gpu] mlp gradient, {rows} rows")
    ens = MLPEnsemble(DIMS, 5, seed=42)
    rng = np.random.default_rng(rows)
    x, t = batch(rng, rows)
    loss, g = ens.gradient(x, t)
    for e in range(5):
        p = ens.get_params(e).astype(np.float64)
        lr, gr = mlp_ref.loss_and_grad(p, list(DIMS), x.astype(np.float64), t.astype(np.float64), bf16=True)
        l64, g64 = mlp_ref.loss_and_grad(p, list(DIMS), x.astype(np.float64), t.astype(np.float64), bf16=False)
        assert abs(loss[e] - lr) <= 1e-3 * lr
        assert rel2(g[e], gr) <= 2e-2, (e, rel2(g[e], gr))
        assert rel2(g[e], g64) <= 5e-2, (e, rel2(g[e], g64))


def test_training_reduces_loss_and_members_differ():
    say("[gpu] mlp training, 2048 rows, 60 epochs")
    rng = np.random.default_rng(5)
    x, t = batch(rng, 2048)
    ens = MLPEnsemble(DIMS, 5, seed=43)
    hist, secs = ens.train(x, t, batch=256, epochs=60, lr=2e-3)
    say(f"  loss {hist[0].mean():.4f} -> {hist[-1].mean():.4f} in {secs:.2f} s")
    assert np.all(hist[-1] < 0.25 * hist[0])
    assert np.all(np.isfinite(hist))
    # members start and shuffle differently
    assert not np.array_equal(ens.get_params(0), ens.get_params(1))
    # the trained network fits better than the mean predictor
    y = ens.forward(x).mean(0)
    assert np.mean((y - t) ** 2) < 0.5 * np.mean((t - t.mean(0)) ** 2)
