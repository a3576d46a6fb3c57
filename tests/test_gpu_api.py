"""GPU tests of the drop-in APIs, restating the reference's own contract tests
(tests/test_metal_multi_mol_convergence.py:39-93, tests/test_mmff_minimize_nk.py:25-58,
tests/test_macrocycle_bounds.py:82-88) on the package's own molecules and kernels."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from conftest import mol_params

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _mols(smiles, n_confs=2, seed=42):
    from mlxmolkit_archive_amd.embed import embed
    from mlxmolkit_archive_amd.mmff_params import extract_mmff_params
    from mlxmolkit_archive_amd.mol import mol_from_smiles
    out = []
    for i, s in enumerate(smiles):
        m = mol_from_smiles(s)
        m.set_conformers(np.zeros((1, m.GetNumAtoms(), 3)))
        p = extract_mmff_params(m)
        x = embed([p], [m.atomic_numbers()], [n_confs], seeds=np.arange(n_confs) + seed + 100 * i, noise=0.05)
        m.set_conformers(x.reshape(n_confs, -1, 3))
        out.append(m)
    return out


EASY = ["c1ccccc1", "Clc1ccccc1"]
HARD = ["CCCCCCCCCCCCCCCC"]


def _mm(mols, **kw):
    from mlxmolkit_archive_amd import mmff_optimize_metal_multi_mol
    return mmff_optimize_metal_multi_mol([(m, None) for m in mols], **kw)


def test_multi_mol_per_conf_iters_recorded_and_below_cap():
    res = _mm(_mols(EASY), max_iters=2000)
    for r in res.mol_results:
        assert r.per_conf_iters is not None and r.converged is not None
        assert r.converged.all()
        assert (r.per_conf_iters < 500).all() and (r.per_conf_iters > 0).all()


def test_multi_mol_converged_flag_agrees_with_grad_norms():
    tol = 1e-4
    for r in _mm(_mols(EASY), max_iters=2000, grad_tol=tol).mol_results:
        for conv, g in zip(r.converged, r.grad_norms):
            assert conv == (g <= tol)


def test_multi_mol_flexible_needs_more_iterations():
    res = _mm(_mols(EASY + HARD), max_iters=2000)
    rigid = np.concatenate([r.per_conf_iters for r in res.mol_results[:2]])
    assert res.mol_results[-1].per_conf_iters.min() > rigid.max()


def test_multi_mol_raising_cap_does_not_change_answer():
    a = _mm(_mols(EASY), max_iters=1000).mol_results
    b = _mm(_mols(EASY), max_iters=4000).mol_results
    for x, y in zip(a, b):
        np.testing.assert_allclose(x.energies, y.energies, atol=1e-4)
        np.testing.assert_array_equal(x.per_conf_iters, y.per_conf_iters)


def test_multi_mol_grad_tol_zero_disables_exit():
    for r in _mm(_mols(EASY), max_iters=100, grad_tol=0.0).mol_results:
        assert not r.converged.any() and (r.per_conf_iters == 100).all()


def test_multi_mol_writes_positions_back_in_place():
    mols = _mols(["CCO"], n_confs=3)
    before = [c.GetPositions() for c in mols[0].GetConformers()]
    res = _mm(mols, max_iters=200)
    r = res.mol_results[0]
    assert r.positions.shape == (3, 9, 3) and r.n_conformers == 3
    for k, c in enumerate(mols[0].GetConformers()):
        assert np.allclose(c.GetPositions(), r.positions[k], atol=1e-6)
        assert not np.allclose(c.GetPositions(), before[k])


def test_fused_multi_mol_result_layout():
    from mlxmolkit_archive_amd import mmff_optimize_metal_fused_multi_mol
    mols = _mols(EASY)
    res = mmff_optimize_metal_fused_multi_mol([(m, None) for m in mols], max_iters=300)
    for r, m in zip(res.mol_results, mols):
        assert r.n_iters == 300 and np.all(r.grad_norms == 0)
        assert r.per_conf_iters is None and r.converged is None
        assert r.energies.dtype == np.float32 and np.all(np.isfinite(r.energies))


def test_metal_batch_contract():
    """mmff_optimize_metal_batch (mmff_metal_optimizer.py:77-202): n_iters = max_iters, no
    per-conformer iterations or flags, grad_norms = ||g_sc||_2 (at least the largest per-atom
    norm), best-seen positions written back in place, energies at those positions."""
    from mlxmolkit_archive_amd import mmff_optimize_metal_batch
    mols = _mols(["CC(=O)OC"], n_confs=3)
    r = mmff_optimize_metal_batch(mols[0], max_iters=150)
    assert r.n_iters == 150 and r.n_conformers == 3
    assert r.per_conf_iters is None and r.converged is None
    assert r.positions.shape == (3, mols[0].GetNumAtoms(), 3) and r.energies.dtype == np.float32
    p = oracle_params(mols[0])
    for k, c in enumerate(mols[0].GetConformers()):
        assert np.allclose(c.GetPositions(), r.positions[k], atol=1e-6)
        e_ref = oracle.energy_grad(p, r.positions[k].reshape(-1).astype(np.float64), family=oracle.FAMILY_MEGA)[0]
        assert abs(r.energies[k] - e_ref) < 1e-3 + 1e-5 * abs(e_ref)
    assert np.all(r.grad_norms > 0) and np.all(np.isfinite(r.grad_norms))


def oracle_params(mol):
    from mlxmolkit_archive_amd.mmff_params import extract_mmff_params
    return extract_mmff_params(mol)


def test_minimize_nk_initial_energy_and_shapes():
    """tests/test_mmff_minimize_nk.py:25-58 restated: max_iters=0 energy equals the
    independent (oracle) energy; aspirin stays finite and < 50 kcal/mol; shapes kept."""
    from mlxmolkit_archive_amd import mmff_minimize_nk
    m, p, x = mol_params("c1ccccc1", noise=0.0)
    out_pos, e, conv = mmff_minimize_nk([p], [1], x, max_iters=0)
    assert e[0] == pytest.approx(oracle.energy_grad(p, x.astype(np.float64))[0], abs=1e-3)
    assert out_pos.shape == x.shape and e.dtype == np.float32 and conv.dtype == bool
    m, p, x = mol_params("CC(=O)OC1=CC=CC=C1C(=O)O", noise=0.0)
    out_pos, e, conv = mmff_minimize_nk([p], [1], x, max_iters=100)
    assert conv.shape == (1,) and np.isfinite(e[0]) and e[0] < 50.0 and out_pos.shape == x.shape


def test_minimize_nk_rejects_bad_positions():
    from mlxmolkit_archive_amd import mmff_minimize_nk
    m, p, x = mol_params("CCO")
    with pytest.raises(ValueError):
        mmff_minimize_nk([p], [1], x[:-3])


def test_energy_grad_metal_shapes_and_mega_family():
    from mlxmolkit_archive_amd import mmff_energy_grad_metal
    m, p, x = mol_params("CCOC(=O)C", k=3)
    e, g = mmff_energy_grad_metal(p, x.reshape(3, -1, 3))
    assert e.shape == (3,) and g.shape == (3, 3 * p.n_atoms)
    e_r = [oracle.energy_grad(p, x.reshape(3, -1)[k].astype(np.float64), family=oracle.FAMILY_MEGA)[0]
           for k in range(3)]
    assert np.allclose(e, e_r, rtol=1e-5, atol=1e-4)


def _pipeline_with_recorder(monkeypatch, smiles, k, **kw):
    """Run generate_conformers_nk(run_mmff=True) and record every mmff_minimize_nk call the
    MMFF stage makes (its inputs and outputs)."""
    import mlxmolkit_archive_amd.minimize as mn
    from mlxmolkit_archive_amd import generate_conformers_nk
    calls = []
    real = mn.mmff_minimize_nk

    def rec(plist, counts, pos, **a):
        x0 = pos.detach().cpu().numpy().copy() if hasattr(pos, "detach") else np.array(pos, copy=True)
        out = real(plist, counts, pos, **a)
        calls.append(dict(plist=plist, counts=list(counts), x0=x0, kw=a, out=out))
        return out

    monkeypatch.setattr(mn, "mmff_minimize_nk", rec)
    return generate_conformers_nk(smiles, n_confs_per_mol=k, run_mmff=True, **kw), calls


def test_pipeline_mmff_stage_matches_oracle(monkeypatch):
    """The MMFF stage of the pipeline (conformer_pipeline_v2.py:292-326) against the oracle:
    each chunk's minimizer call gets the reference's iteration budget 200 + 15 x atoms and
    the L-BFGS switch (off below 150 atoms); the returned conformers are that call's outputs,
    in order; and the FP64 oracle from the same starts reaches the same minima (0.1 kcal/mol)
    for at least 90% of them. A batch with a 152-atom molecule switches every chunk to L-BFGS."""
    smiles = ["CCO", "c1ccccc1C(=O)N", "CCCCCCCC", "CC(=O)OC1=CC=CC=C1C(=O)O", "CCN(CC)CC", "OCC(O)CO"]
    res, calls = _pipeline_with_recorder(monkeypatch, smiles, 5, max_confs_per_batch=12)
    assert res.total_conformers == 30 and len(calls) == res.n_batches >= 3
    n_atoms = [mr.n_atoms for mr in res.molecules]
    e_all, e_ref_all = [], []
    got = {i: 0 for i in range(len(smiles))}
    for call in calls:
        kw = call["kw"]
        assert kw["use_lbfgs"] is False
        ob = oracle.OracleBatch(call["plist"], call["counts"])
        assert kw["max_iters"] == 200 + 15 * max(p.n_atoms for p in call["plist"])
        out_pos, e32, conv, info = call["out"]
        _, e_ref, st_ref, _ = oracle.minimize_nk(ob, call["x0"].astype(np.float64), max_iters=kw["max_iters"],
                                                 grad_tol=1e-4)
        e_gpu, _, _ = oracle.energy_grad_batch(ob, out_pos.astype(np.float64))
        ok = conv & (st_ref == 0)
        e_all.append(e_gpu[ok])
        e_ref_all.append(e_ref[ok])
        # the pipeline hands back exactly this call's conformers, molecule by molecule
        off = 0
        c = 0
        for p, cnt in zip(call["plist"], call["counts"]):
            m = n_atoms.index(p.n_atoms) if n_atoms.count(p.n_atoms) == 1 else None
            for _ in range(cnt):
                if m is not None:
                    mr = res.molecules[m]
                    j = got[m]
                    assert np.array_equal(mr.positions_3d[j].reshape(-1), out_pos[off:off + 3 * p.n_atoms])
                    assert mr.energies[j] == float(info["energy_f64"][c]) and mr.converged[j] == bool(conv[c])
                    got[m] += 1
                off += 3 * p.n_atoms
                c += 1
    d = np.concatenate(e_all) - np.concatenate(e_ref_all)
    print(f"\npipeline MMFF stage: {len(d)} conformers converged on both sides, within 0.1 "
          f"{np.mean(np.abs(d) <= 0.1):.3f}, GPU lower {np.sum(d < -0.1)}, higher {np.sum(d > 0.1)}")
    assert len(d) >= 27 and np.mean(np.abs(d) <= 0.1) >= 0.9
    res2, calls2 = _pipeline_with_recorder(monkeypatch, ["C" * 50, "CCO"], 2)
    assert calls2 and all(c["kw"]["use_lbfgs"] is True for c in calls2)
    assert all(np.isfinite(e).all() for mr in res2.molecules for e in [np.array(mr.energies)])


def test_pipeline_run_mmff():
    """tests/test_macrocycle_bounds.py:82-88: the pipeline runs with run_mmff=True."""
    from mlxmolkit_archive_amd import generate_conformers_nk
    res = generate_conformers_nk(["CCO", "c1ccccc1C(=O)N", "CCCCCC"], n_confs_per_mol=4, run_mmff=True,
                                 max_confs_per_batch=5)
    assert res.total_conformers == 12 and res.n_batches == 3
    for mr in res.molecules:
        assert len(mr.positions_3d) == 4 and all(np.isfinite(mr.energies))
        assert all(mr.converged)
    res2 = generate_conformers_nk(["CCO"], n_confs_per_mol=2, run_mmff=False)
    assert all(np.isnan(res2.molecules[0].energies)) and all(res2.molecules[0].converged)
