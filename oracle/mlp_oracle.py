"""FP64 restatement of the learned preconditioner network (TEST INFRASTRUCTURE ONLY).

Only tests/ may import this module; it is the checker for the HIP MLP kernels of
mlxmolkit_archive_amd/csrc/precond.hip. The network has no counterpart in the reference: it
is the per-atom MLP of the learned initial inverse Hessian (104 -> 64 -> 64 -> 6, ReLU;
outputs = Cholesky factor of a 3x3 SPD block with softplus(.) + 1e-3 on the diagonal).

Pinned by tests/test_mlp_oracle.py: central finite differences of the FP64 loss against the
analytic gradient for every parameter group, and closed forms of the block (L L^T is SPD,
zero outputs give softplus(0) + 1e-3 squared on the diagonal).
"""
from __future__ import annotations

import numpy as np

N_IN, N_H, N_OUT = 104, 64, 6


def unpack(params):
    p = np.asarray(params, np.float64)
    sizes = [N_H * N_IN, N_H, N_H * N_H, N_H, N_OUT * N_H, N_OUT]
    out, o = [], 0
    for s in sizes:
        out.append(p[o:o + s])
        o += s
    w1, b1, w2, b2, w3, b3 = out
    return w1.reshape(N_H, N_IN), b1, w2.reshape(N_H, N_H), b2, w3.reshape(N_OUT, N_H), b3


def _sp(v):
    return np.where(v > 20.0, v, np.log1p(np.exp(np.minimum(v, 20.0)))) + 1e-3


def _sig(v):
    return 1.0 / (1.0 + np.exp(-v))


def forward(params, feat):
    """(blocks (n, 9), cache) in FP64."""
    w1, b1, w2, b2, w3, b3 = unpack(params)
    x = np.asarray(feat, np.float64)
    z1 = x @ w1.T + b1
    h1 = np.maximum(z1, 0.0)
    z2 = h1 @ w2.T + b2
    h2 = np.maximum(z2, 0.0)
    o = h2 @ w3.T + b3
    L = np.zeros((len(x), 3, 3))
    L[:, 0, 0] = _sp(o[:, 0])
    L[:, 1, 0] = o[:, 1]
    L[:, 1, 1] = _sp(o[:, 2])
    L[:, 2, 0] = o[:, 3]
    L[:, 2, 1] = o[:, 4]
    L[:, 2, 2] = _sp(o[:, 5])
    P = L @ L.transpose(0, 2, 1)
    return P.reshape(-1, 9), (x, z1, h1, z2, h2, o, L)


def loss_grad(params, feat, target, weight=None):
    """(loss sum, per-atom loss, gradient) of sum_a w_a ||P_a - T_a||_F^2 in FP64."""
    w1, b1, w2, b2, w3, b3 = unpack(params)
    P, (x, z1, h1, z2, h2, o, L) = forward(params, feat)
    T = np.asarray(target, np.float64).reshape(-1, 9)
    w = np.ones(len(x)) if weight is None else np.asarray(weight, np.float64)
    per = ((P - T) ** 2).sum(1)
    G = (2.0 * w[:, None] * (P - T)).reshape(-1, 3, 3)
    dL = (G + G.transpose(0, 2, 1)) @ L
    do = np.stack([dL[:, 0, 0] * _sig(o[:, 0]), dL[:, 1, 0], dL[:, 1, 1] * _sig(o[:, 2]), dL[:, 2, 0], dL[:, 2, 1],
                   dL[:, 2, 2] * _sig(o[:, 5])], 1)
    gw3 = do.T @ h2
    gb3 = do.sum(0)
    dz2 = (do @ w3) * (z2 > 0)
    gw2 = dz2.T @ h1
    gb2 = dz2.sum(0)
    dz1 = (dz2 @ w2) * (z1 > 0)
    gw1 = dz1.T @ x
    gb1 = dz1.sum(0)
    g = np.concatenate([gw1.ravel(), gb1, gw2.ravel(), gb2, gw3.ravel(), gb3])
    return float((w * per).sum()), per, g
