"""The oracle's FP32-emulation analysis switch: off by default and restored on exit; when on,
coordinates, term energies and gradients take float values while the sums stay FP64."""
from __future__ import annotations

import numpy as np

import oracle
from golden_io import load_small


def test_fp32_emulation_rounds_and_restores():
    params, counts, z = load_small()
    p = params[0]
    n3 = 3 * p.n_atoms
    x = z["x"][:n3].astype(np.float64) + 1e-9  # not float-representable
    e0, g0 = oracle.energy_grad(p, x)
    with oracle.fp32_emulation():
        e1, g1 = oracle.energy_grad(p, x)
        ob = oracle.OracleBatch([p], [1])
        xo, eo, st, cnt = oracle.minimize_nk(ob, x, max_iters=50, grad_tol=1e-4, n_threads=1)
    # gradient entries are floats; the energy moves by term rounding only
    assert np.array_equal(g1, g1.astype(np.float32).astype(np.float64))
    assert abs(e1 - e0) <= 1e-6 * max(1.0, oracle.terms_abs(p, x).sum())
    assert np.abs(g1 - g0).max() <= 1e-6 * max(1.0, np.abs(g0).max())
    # the minimizer's coordinates are floats throughout
    assert np.array_equal(xo, xo.astype(np.float32).astype(np.float64))
    # off again: the FP64 results come back bit for bit
    e2, g2 = oracle.energy_grad(p, x)
    assert e2 == e0 and np.array_equal(g2, g0)
