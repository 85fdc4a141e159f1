"""CPU restatement of the learned part's network (numpy, FP64 arithmetic).

TEST INFRASTRUCTURE ONLY: imported by tests/ (and nothing in the product).

The network has no counterpart in the reference; this file restates the
algorithm of plegma_amd/csrc/plg_mlp.hip:
  x -> [Linear -> SiLU] x (L-1) -> Linear -> y,   loss = mean (y - t)^2
with the parameter vector laid out per layer as W [out][in] then b [out].
With emulate=True the bf16 rounding points of the device kernels are
reproduced (inputs, weights, hidden activations and the back-propagated
dZ operands are rounded to bf16, round-to-nearest-even), so that the device
gradient can be compared at FP32-accumulation accuracy; with emulate=False
everything is FP64 (used for finite-difference checks of the analytic
gradient).
"""
from __future__ import annotations

import numpy as np


def bf16(x):
    """round to the nearest bf16 (ties to even), returned as float64"""
    f = np.ascontiguousarray(np.asarray(x, dtype=np.float64).astype(np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def layers(p, dims):
    """[(W [out][in], b [out])] views of the flat parameter vector"""
    out, o = [], 0
    for i, j in zip(dims[:-1], dims[1:]):
        W = p[o:o + i * j].reshape(j, i)
        o += i * j
        b = p[o:o + j]
        o += j
        out.append((W, b))
    assert o == len(p), "parameter vector does not match dims"
    return out


def nparams(dims):
    return sum(i * j + j for i, j in zip(dims[:-1], dims[1:]))


def silu(z):
    return z / (1.0 + np.exp(-z))


def silu_grad(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(p, dims, x, emulate=True):
    """y and the cache (layer inputs X_l, pre-activations Z_l)"""
    q = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    p = np.asarray(p, dtype=np.float64)
    X = [q(x)]
    Z = []
    Ls = layers(p, dims)
    for l, (W, b) in enumerate(Ls):
        z = X[-1] @ q(W).T + b
        Z.append(z)
        if l < len(Ls) - 1:
            X.append(q(silu(z)))
    return Z[-1], (X, Z)


def loss_grad(p, dims, x, t, emulate=True):
    """loss = mean (y - t)^2 and its gradient (flat, FP64)"""
    q = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    p = np.asarray(p, dtype=np.float64)
    y, (X, Z) = forward(p, dims, x, emulate)
    e = y - t
    loss = float(np.mean(e * e))
    dz = 2.0 * e / e.size
    Ls = layers(p, dims)
    grads = [None] * len(Ls)
    for l in range(len(Ls) - 1, -1, -1):
        W, _ = Ls[l]
        dzq = q(dz)
        gW = dzq.T @ X[l]
        gb = dz.sum(axis=0)
        grads[l] = (gW, gb)
        if l > 0:
            dx = dzq @ q(W)
            dz = dx * silu_grad(Z[l - 1])
    g = np.concatenate([np.concatenate([gW.reshape(-1), gb]) for gW, gb in grads])
    return loss, g


def bias_corrected(f_pred, f_bc, y_bc):
    """per-source corrected predictions f(x_i) + mean_bc (y - f(x)) and the
    estimate of the population mean of y over the predicted sources"""
    delta = np.mean(y_bc - f_bc, axis=0)
    corr = f_pred + delta
    return corr, corr.mean(axis=0)
