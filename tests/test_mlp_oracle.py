"""Pins the FP64 oracle of the learned-preconditioner network (oracle/mlp_oracle.py)."""
from __future__ import annotations

import numpy as np
import pytest

import mlp_oracle as mo


def _params(seed=0, scale=0.3):
    g = np.random.default_rng(seed)
    n = mo.N_H * mo.N_IN + mo.N_H + mo.N_H * mo.N_H + mo.N_H + mo.N_OUT * mo.N_H + mo.N_OUT
    return g.normal(0, scale, n) / np.sqrt(16)


def test_blocks_are_spd_and_zero_output_closed_form():
    p = _params(1)
    x = np.random.default_rng(2).normal(0, 1, (20, mo.N_IN))
    P, _ = mo.forward(p, x)
    for b in P.reshape(-1, 3, 3):
        assert np.allclose(b, b.T)
        assert np.linalg.eigvalsh(b).min() > 0
    z = np.zeros_like(p)  # all outputs 0: L = diag(softplus(0) + 1e-3)
    P0, _ = mo.forward(z, x[:1])
    d = (np.log(2.0) + 1e-3) ** 2
    assert np.allclose(P0.reshape(3, 3), d * np.eye(3))


@pytest.mark.parametrize("seed", [0, 3])
def test_gradient_matches_central_differences(seed):
    g = np.random.default_rng(seed)
    p = _params(seed)
    x = g.normal(0, 1, (6, mo.N_IN))
    t = g.normal(0, 0.5, (6, 9))
    w = g.uniform(0.5, 2.0, 6)
    _, _, grad = mo.loss_grad(p, x, t, w)
    # every parameter group: a handful of indices each
    sizes = [mo.N_H * mo.N_IN, mo.N_H, mo.N_H * mo.N_H, mo.N_H, mo.N_OUT * mo.N_H, mo.N_OUT]
    starts = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    h = 1e-6
    for s0, sz in zip(starts, sizes):
        for k in g.choice(sz, size=min(5, sz), replace=False):
            i = s0 + k
            pp, pm = p.copy(), p.copy()
            pp[i] += h
            pm[i] -= h
            fd = (mo.loss_grad(pp, x, t, w)[0] - mo.loss_grad(pm, x, t, w)[0]) / (2 * h)
            assert abs(fd - grad[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, grad[i])
