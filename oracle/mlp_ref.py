"""CPU oracle of the surrogate MLP (test infrastructure only: imported by tests/
to check the HIP kernels, never by the product).

Restates, in numpy float64, the network of include/sbewh_surrogate.h: SiLU
hidden layers, linear output, mean-squared-error loss
L = mean_{b,o} (y_bo - t_bo)^2, and its gradient by backpropagation.  With
``bf16=True`` it rounds at exactly the points where the HIP kernels store
BF16 (inputs, weights, hidden activations, output gradient, hidden-layer
gradients), so the remaining difference to the GPU is FP32 versus FP64
accumulation.  Pinned in tests/test_surrogate_oracle.py by central finite
differences of the unrounded loss and by a one-layer closed form.
"""
from __future__ import annotations

import numpy as np


def bf16_round(a: np.ndarray) -> np.ndarray:
    """round to the nearest bfloat16 (ties to even), returned as float64"""
    f = np.asarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32).astype(np.float64)
    nan = np.isnan(f)
    if nan.any():
        out[nan] = np.nan
    return out


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def unpack(params: np.ndarray, dims: list):
    """[(W_l, b_l)] views of one member's flat parameter vector"""
    out, off = [], 0
    for l in range(len(dims) - 1):
        W = params[off:off + dims[l + 1] * dims[l]].reshape(dims[l + 1], dims[l])
        off += dims[l + 1] * dims[l]
        b = params[off:off + dims[l + 1]]
        off += dims[l + 1]
        out.append((W, b))
    assert off == params.size
    return out


def forward(params, dims, x, bf16=False):
    r = bf16_round if bf16 else (lambda a: np.asarray(a, np.float64))
    layers = unpack(np.asarray(params, np.float64), dims)
    h = r(x)
    for l, (W, b) in enumerate(layers):
        z = h @ r(W).T + b
        if l + 1 < len(layers):
            h = r(silu(z))
        else:
            return z


def loss_and_grad(params, dims, x, t, bf16=False):
    """(loss, flat gradient) of one member on one batch"""
    r = bf16_round if bf16 else (lambda a: np.asarray(a, np.float64))
    params = np.asarray(params, np.float64)
    layers = unpack(params, dims)
    hs, zs = [r(x)], []
    h = hs[0]
    for l, (W, b) in enumerate(layers):
        z = h @ r(W).T + b
        zs.append(z)
        if l + 1 < len(layers):
            h = r(silu(z))
            hs.append(h)
    y = zs[-1]
    cnt = y.size
    d = y - t
    loss = float((d * d).sum() / cnt)
    g = r(2.0 * d / cnt)
    grads = []
    for l in range(len(layers) - 1, -1, -1):
        W = r(layers[l][0])
        grads.append((g.T @ hs[l], g.sum(0)))
        if l > 0:
            g = r((g @ W) * dsilu(zs[l - 1]))
    flat = []
    for dW, db in reversed(grads):
        flat.append(dW.reshape(-1))
        flat.append(db)
    return loss, np.concatenate(flat)
