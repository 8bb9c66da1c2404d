"""CPU oracle of the learned-predictor network (test infrastructure only).

Restates, in numpy float64, the computation of csrc/mlp.hip: layers
Z = A_prev W' + b with W (out x in) row-major, tanh on hidden layers,
identity output; the forward reads bf16 copies of the weights (round to
nearest even of the float32 master) and float biases; the backward
differentiates that computation.  Parameter vector layout: per layer W then b.

This module has no counterpart in the reference (the reference continuation
has no learned predictor); it pins the GPU kernels by an independent
restatement plus finite differences in the tests.
"""
from __future__ import annotations

import numpy as np


def bf16_round(w):
    """float32 -> bf16 (round to nearest even) -> float64, as the kernels do."""
    u = np.asarray(w, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(params, widths):
    out, off = [], 0
    for nin, nout in zip(widths[:-1], widths[1:]):
        W = params[off:off + nout * nin].reshape(nout, nin)
        off += nout * nin
        b = params[off:off + nout]
        off += nout
        out.append((W, b))
    assert off == params.size
    return out


def forward(params, widths, X):
    """Returns (Y, activations list [A_0=X, A_1, ..., A_L])."""
    layers = unpack(np.asarray(params, dtype=np.float32), widths)
    A = [np.asarray(X, dtype=np.float64)]
    for l, (W, b) in enumerate(layers):
        Z = A[-1] @ bf16_round(W).T + b.astype(np.float64)
        A.append(np.tanh(Z) if l + 1 < len(layers) else Z)
    return A[-1], A


def backward(params, widths, X, dY):
    """Gradient of sum(dY * Y) w.r.t. the parameter vector."""
    layers = unpack(np.asarray(params, dtype=np.float32), widths)
    _, A = forward(params, widths, X)
    grads = []
    cur = np.asarray(dY, dtype=np.float64)
    for l in range(len(layers) - 1, -1, -1):
        W, _ = layers[l]
        if l + 1 < len(layers):
            cur = cur * (1.0 - A[l + 1] ** 2)
        dW = cur.T @ A[l]
        db = cur.sum(axis=0)
        grads.append((dW, db))
        cur = cur @ bf16_round(W)
    out = []
    for dW, db in reversed(grads):
        out.append(dW.ravel())
        out.append(db)
    return np.concatenate(out)


def predictor_loss(Y, G, g, h, lam_t, w_lam):
    """Per-sample loss delta'G delta - 2 delta'g + h + w (y_l - lam_t)^2 and
    d(mean loss)/dY."""
    Y = np.asarray(Y, dtype=np.float64)
    B, no = Y.shape
    nc = no - 1
    d = Y[:, :nc]
    Gd = np.einsum("sij,sj->si", G.reshape(B, nc, nc), d)
    e = Y[:, nc] - lam_t
    L = (d * Gd).sum(1) - 2 * (d * g.reshape(B, nc)).sum(1) + h + w_lam * e * e
    dY = np.zeros_like(Y)
    dY[:, :nc] = 2 * (Gd - g.reshape(B, nc)) / B
    dY[:, nc] = 2 * w_lam * e / B
    return L, dY
