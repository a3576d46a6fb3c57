"""Golden fixtures on the CPU: the oracle reproduces its committed vectors, and the
built-in typer is measured against RDKit-relaxed geometries held by the reference.
"""
from __future__ import annotations

import json

import numpy as np
import pytest

import oracle
from golden_io import GOLDEN, load_small
from mlxmolkit_archive_amd.mmff_params import extract_mmff_params
from mlxmolkit_archive_amd.mol import mol_from_smiles


def test_oracle_reproduces_golden():
    params, counts, z = load_small()
    ob = oracle.OracleBatch(params, counts)
    x = z["x"].astype(np.float64)
    e, g, et = oracle.energy_grad_batch(ob, x, family=oracle.FAMILY_TG)
    np.testing.assert_allclose(e, z["e_tg"], rtol=1e-12, atol=1e-10)
    np.testing.assert_allclose(g, z["g_tg"], rtol=1e-10, atol=1e-9)
    e2, _, _ = oracle.energy_grad_batch(ob, x, family=oracle.FAMILY_MEGA)
    np.testing.assert_allclose(e2, z["e_mg"], rtol=1e-12, atol=1e-10)
    xb, eb, sb, cb = oracle.minimize_nk(ob, x, max_iters=2000, grad_tol=1e-4)
    np.testing.assert_allclose(eb, z["bfgs_e"], rtol=1e-9, atol=1e-8)
    assert np.array_equal(sb, z["bfgs_status"]) and np.array_equal(cb, z["bfgs_counters"])
    r = oracle.multi_mol(ob, x, max_iters=200)
    np.testing.assert_allclose(r[1], z["mm_e"], rtol=1e-9, atol=1e-8)
    assert r[5] == int(z["mm_used"])


def _fixture():
    return json.loads((GOLDEN / "rdkit_relaxed.json").read_text())


def test_smiles_parser_reproduces_rdkit_atom_order():
    """Heavy atoms in SMILES order, then hydrogens per heavy atom (RDKit AddHs order)."""
    for name, d in _fixture().items():
        m = mol_from_smiles(d["smiles"])
        assert m.atomic_numbers().tolist() == d["Z"], name


def test_typer_residual_forces_at_rdkit_minima():
    """Partial pin of the typing boundary: at RDKit-MMFF94-relaxed geometries the exact
    MMFF94 gradient is ~0. The built-in MMFF94-style parameters leave residual forces; this
    records them and guards against regressions: median per-molecule RMS force < 5 and
    every RMS < 50 kcal/mol/A (thiophene and HCl are the outliers: 5-ring heteroaromatic
    and H-Cl parameters are the roughest part of the built-in table)."""
    rows = []
    for name, d in _fixture().items():
        m = mol_from_smiles(d["smiles"])
        X = np.asarray(d["coords"], np.float64)
        m.set_conformers(X[None])
        try:
            p = extract_mmff_params(m)
        except ValueError:
            continue
        e, g = oracle.energy_grad(p, X)
        g3 = g.reshape(-1, 3)
        rows.append((name, np.sqrt((g3 ** 2).sum(1).mean()), np.abs(g3).max()))
    assert len(rows) >= 12
    rms = np.array([r[1] for r in rows])
    mx = np.array([r[2] for r in rows])
    print("\n".join(f"{n:20s} rms {a:7.3f} max {b:7.3f}" for n, a, b in rows))
    assert np.median(rms) < 5.0 and rms.max() < 50.0
