"""GPU parity: the HIP path against the FP64 oracle and the committed golden vectors.

Tolerances (stated per check):
* per-term energies: |E_gpu - E_oracle| <= 1e-5 * sum over terms of |e_term| + 1e-6 kcal/mol
  (FP32 term arithmetic; the sum of absolute term energies is the scale its rounding acts on)
* gradients: max |g_gpu - g_oracle| <= 1e-5 * max |g_oracle| per conformer
* minimizers: final FP64 energy of the GPU geometry within 0.1 kcal/mol of the oracle
  minimization from the same start for >= 90% (BFGS) / >= 80% (L-BFGS) of conformers,
  median within 2e-3; the rest fall into a different local minimum after the FP32 and
  FP64 trajectories part (flexible molecules, hundreds of iterations)
* multi_mol: iterations used within 20 of the oracle's, converged flags agreeing for >= 90%
  of conformers, median |dE| < 0.01 kcal/mol where both converged
* reported FP64 energies equal the FP64 re-evaluation of the returned geometry to 1e-3
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from golden_io import load_small

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from mlxmolkit_archive_amd import _lib
    _lib.load()


def _eng():
    from mlxmolkit_archive_amd import engine
    return engine


def _slices(ob):
    for c in range(ob.n_confs):
        p_atoms = ob.tables[ob.conf_mol[c]].mol.n_atoms
        yield c, slice(int(ob.conf_off[c]), int(ob.conf_off[c]) + 3 * p_atoms)


def _term_scale(params, counts, x, family, diel=1):
    out = []
    k = 0
    off = 0
    for p, cnt in zip(params, counts):
        for _ in range(cnt):
            out.append(oracle.terms_abs(p, x[off:off + 3 * p.n_atoms], family=family, diel=diel))
            off += 3 * p.n_atoms
            k += 1
    return np.array(out)


def _check_eg(params, counts, x, family, diel=1):
    eng = _eng()
    db = eng.device_batch(params, counts)
    xd = torch.from_numpy(np.ascontiguousarray(x, np.float32)).cuda()
    e, g, et = eng.energy_grad(db, xd, family, diel, 0x7F, grad=True, terms=True)
    torch.cuda.synchronize()
    ob = oracle.OracleBatch(params, counts)
    x64 = np.asarray(x, np.float32).astype(np.float64)
    e_r, g_r, et_r = oracle.energy_grad_batch(ob, x64, family=family, diel=diel)
    scale = _term_scale(params, counts, x64, family, diel)
    et = et.cpu().numpy()
    assert np.all(np.abs(et - et_r) <= 1e-5 * scale + 1e-6), np.max(np.abs(et - et_r) / (scale + 1e-12), axis=0)
    assert np.all(np.abs(e.cpu().numpy() - e_r) <= 1e-5 * scale.sum(1) + 1e-6)
    g = g.cpu().numpy()
    for c, sl in _slices(ob):
        assert np.abs(g[sl] - g_r[sl]).max() <= 1e-5 * max(np.abs(g_r[sl]).max(), 1e-3), c


@pytest.mark.parametrize("family", [0, 1])
def test_energy_grad_per_term_small(small_wl, family):
    _check_eg(small_wl.params, small_wl.conf_counts, small_wl.x, family)


def test_energy_grad_per_term_medium_and_diel2(medium_wl):
    _check_eg(medium_wl.params, medium_wl.conf_counts, medium_wl.x, 0)
    _check_eg(medium_wl.params, medium_wl.conf_counts, medium_wl.x, 0, diel=2)
    _check_eg(medium_wl.params, medium_wl.conf_counts, medium_wl.x, 1)


def test_energy_grad_term_mask(small_wl):
    eng = _eng()
    wl = small_wl
    db = eng.device_batch(wl.params, wl.conf_counts)
    xd = torch.from_numpy(wl.x).cuda()
    ob = oracle.OracleBatch(wl.params, wl.conf_counts)
    for k in range(7):
        e, g, _ = eng.energy_grad(db, xd, 0, 1, 1 << k, grad=True)
        torch.cuda.synchronize()
        e_r, g_r, _ = oracle.energy_grad_batch(ob, wl.x.astype(np.float64), family=0, mask=1 << k)
        assert np.allclose(e.cpu().numpy(), e_r, rtol=2e-5, atol=2e-4), k


def _minimize_vs_oracle(params, counts, x, use_lbfgs=False, **kw):
    eng = _eng()
    db = eng.device_batch(params, counts)
    xd = torch.from_numpy(np.ascontiguousarray(x, np.float32)).cuda()
    res = eng.minimize_nk(db, xd, max_iters=2000, grad_tol=1e-4, use_lbfgs=use_lbfgs, lbfgs_m=8, **kw)
    torch.cuda.synchronize()
    ob = oracle.OracleBatch(params, counts)
    _, e_ref, st_ref, cnt_ref = oracle.minimize_nk(ob, np.asarray(x, np.float32).astype(np.float64),
                                                   max_iters=2000, grad_tol=1e-4, use_lbfgs=use_lbfgs, m_hist=8)
    xg = res.x.cpu().numpy()
    e_fin, _, _ = oracle.energy_grad_batch(ob, xg.astype(np.float64), family=0)
    return res, e_fin, e_ref, st_ref, cnt_ref


@pytest.mark.parametrize("use_lbfgs", [False, True])
def test_minimize_nk_matches_oracle(small_wl, use_lbfgs):
    wl = small_wl
    res, e_fin, e_ref, st_ref, cnt_ref = _minimize_vs_oracle(wl.params, wl.conf_counts, wl.x, use_lbfgs)
    st = res.status.cpu().numpy()
    assert (st == 0).all() and (st_ref == 0).all()
    dE = np.abs(e_fin - e_ref)
    assert np.mean(dE <= 0.1) >= (0.8 if use_lbfgs else 0.9) and np.median(dE) < 2e-3, dE
    assert np.abs(res.energy.cpu().numpy() - e_fin).max() < 1e-3
    it = res.counters.cpu().numpy()[:, 0]
    assert np.abs(np.median(it) - np.median(cnt_ref[:, 0])) <= 0.3 * np.median(cnt_ref[:, 0])


def test_minimize_nk_medium_all_buckets(medium_wl):
    wl = medium_wl
    res, e_fin, e_ref, st_ref, _ = _minimize_vs_oracle(wl.params, wl.conf_counts, wl.x, ls_grad=True)
    assert (res.status.cpu().numpy() == 0).all()
    dE = np.abs(e_fin - e_ref)
    assert np.mean(dE <= 0.1) >= 0.9 and np.median(dE) < 1e-2, dE


def test_placements_and_blocks_agree(small_wl):
    """Packed LDS, full HBM and packed HBM Hessians, block sizes and LDS table staging: same minima."""
    eng = _eng()
    wl = small_wl
    db = eng.device_batch(wl.params, wl.conf_counts)
    xd = torch.from_numpy(wl.x).cuda()
    outs = []
    for buckets, stage in ((((1 << 20, 256, 1),), True), (((1 << 20, 256, 2),), False),
                           (((1 << 20, 64, 1),), True), (((1 << 20, 512, 1),), False),
                           (((1 << 20, 256, 3),), False)):
        r = eng.NkRunner(db, max_iters=2000, grad_tol=1e-4, buckets=buckets, stage_tables=stage)
        res = r.run(xd)
        torch.cuda.synchronize()
        outs.append(res.energy.cpu().numpy().copy())
    for o in outs[1:]:
        assert np.median(np.abs(o - outs[0])) < 1e-3


@pytest.mark.parametrize("use_lbfgs", [False, True])
def test_global_hessian_bonded_staging_is_bit_identical(small_wl, use_lbfgs):
    """Beside a global Hessian (or L-BFGS history) the bonded term tables may be staged in LDS.
    The arithmetic is the same, so positions, energies, statuses and counters match the launch
    that reads every table from global memory bit for bit."""
    eng = _eng()
    wl = small_wl
    db = eng.device_batch(wl.params, wl.conf_counts)
    xd = torch.from_numpy(wl.x).cuda()
    outs = []
    for stage in (True, False):
        r = eng.NkRunner(db, max_iters=2000, grad_tol=1e-4, buckets=((1 << 20, 256, 2),), stage_tables=stage,
                         use_lbfgs=use_lbfgs, ls_grad=True)
        res = r.run(xd)
        torch.cuda.synchronize()
        outs.append([t.cpu().numpy().copy() for t in (res.x, res.energy, res.status, res.counters)])
    for a, b in zip(*outs):
        assert np.array_equal(a, b)
    assert (outs[0][2] <= 1).all()


def test_bit_reproducible(small_wl):
    """Fixed-point gradient sums and fixed-order reductions: identical results run to run."""
    eng = _eng()
    wl = small_wl
    db = eng.device_batch(wl.params, wl.conf_counts)
    xd = torch.from_numpy(wl.x).cuda()
    r = eng.NkRunner(db, max_iters=2000, grad_tol=1e-4)
    a = r.run(xd).x.cpu().numpy().copy()
    b = r.run(xd).x.cpu().numpy().copy()
    assert np.array_equal(a, b)


def test_max_iters_zero_and_tiny_molecules():
    """max_iters=0 returns the start energy with status 1 (tests/test_mmff_minimize_nk.py:25-40);
    a bare diatomic converges to r0."""
    from conftest import mol_params
    from mlxmolkit_archive_amd.mmff_params import MMFFParams
    eng = _eng()
    m, p, x = mol_params("c1ccccc1")
    db = eng.device_batch([p], [1])
    xd = torch.from_numpy(x).cuda()
    res = eng.minimize_nk(db, xd, max_iters=0)
    torch.cuda.synchronize()
    e0 = oracle.energy_grad(p, x.astype(np.float64))[0]
    assert res.energy.cpu().numpy()[0] == pytest.approx(e0, rel=1e-5)
    assert res.status.cpu().numpy()[0] == 1
    d = MMFFParams(n_atoms=2, bond_idx1=np.array([0], np.int32), bond_idx2=np.array([1], np.int32),
                   bond_kb=np.array([5.0], np.float32), bond_r0=np.array([1.4], np.float32))
    db = eng.device_batch([d], [1])
    res = eng.minimize_nk(db, torch.tensor([0, 0, 0, 1.9, 0.3, -0.2], dtype=torch.float32).cuda(),
                          max_iters=200, grad_tol=1e-8)
    xo = res.x.cpu().numpy()
    assert res.status.cpu().numpy()[0] == 0
    assert np.linalg.norm(xo[3:] - xo[:3]) == pytest.approx(1.4, abs=1e-4)


def test_golden_vectors():
    """The HIP path against the committed FP64 oracle vectors (tests/golden/oracle_small.npz)."""
    eng = _eng()
    params, counts, z = load_small()
    db = eng.device_batch(params, counts)
    xd = torch.from_numpy(z["x"]).cuda()
    e, g, et = eng.energy_grad(db, xd, 0, 1, 0x7F, grad=True, terms=True)
    e2, _, _ = eng.energy_grad(db, xd, 1, 1, 0x7F, grad=False)
    res = eng.minimize_nk(db, xd, max_iters=2000, grad_tol=1e-4)
    torch.cuda.synchronize()
    scale = _term_scale(params, counts, z["x"].astype(np.float64), 0).sum(1)
    assert np.all(np.abs(e.cpu().numpy() - z["e_tg"]) <= 1e-5 * scale + 1e-6)
    assert np.all(np.abs(e2.cpu().numpy() - z["e_mg"]) <= 1e-5 * scale + 1e-6)
    dE = np.abs(res.energy.cpu().numpy() - z["bfgs_e"])
    assert (res.status.cpu().numpy() == z["bfgs_status"]).all()
    assert np.mean(dE <= 0.1) >= 0.9 and np.median(dE) < 1e-3


def test_bb_fused_tracks_oracle_then_stays_bounded(small_wl):
    """BB has no convergence test, so FP32 and FP64 trajectories part ways eventually;
    the first 20 iterations must agree, and long runs stay below the start energy."""
    eng = _eng()
    wl = small_wl
    db = eng.device_batch(wl.params, wl.conf_counts)
    xd = torch.from_numpy(wl.x).cuda()
    ob = oracle.OracleBatch(wl.params, wl.conf_counts)
    x64 = wl.x.astype(np.float64)
    for k in (1, 5, 20):
        xo, e = eng.bb_fused(db, xd, max_iters=k)
        torch.cuda.synchronize()
        xr, er = oracle.bb_fused(ob, x64, max_iters=k)
        assert np.abs(xo.cpu().numpy() - xr).max() < 1e-3, k
        assert np.abs(e.cpu().numpy() - er).max() < 1e-2 + 1e-5 * np.abs(er).max(), k
    xo, e = eng.bb_fused(db, xd, max_iters=500)
    torch.cuda.synchronize()
    e0, _, _ = oracle.energy_grad_batch(ob, x64, family=1)
    assert np.all(np.isfinite(e.cpu().numpy())) and np.all(e.cpu().numpy() <= e0 + 1e-6)


def test_multimol_matches_oracle(small_wl):
    eng = _eng()
    wl = small_wl
    db = eng.device_batch(wl.params, wl.conf_counts)
    xd = torch.from_numpy(wl.x).cuda()
    ob = oracle.OracleBatch(wl.params, wl.conf_counts)
    for max_iters, tol in ((1000, 1e-4), (60, 0.0)):
        r = eng.multimol(db, xd, max_iters=max_iters, grad_tol=tol)
        torch.cuda.synchronize()
        xo, eo, gno, pcio, convo, usedo = oracle.multi_mol(ob, wl.x.astype(np.float64), max_iters=max_iters,
                                                           grad_tol=tol)
        conv = r["converged"].cpu().numpy().astype(bool)
        pci = r["per_conf_iters"].cpu().numpy()
        if tol == 0.0:
            assert not conv.any() and (pci == max_iters).all() and int(r["iters_used"].cpu()[0]) == max_iters
            continue
        used = int(r["iters_used"].cpu()[0])
        e = r["energy"].cpu().numpy()
        both = conv & convo
        print(f"\nmulti_mol: iters_used gpu {used} oracle {usedo}; converged gpu {conv.sum()} oracle {convo.sum()} "
              f"agree {np.mean(conv == convo):.3f}; median |dE| (both converged) {np.median(np.abs(e[both] - eo[both])):.2e}")
        assert abs(used - usedo) <= 20
        assert np.mean(conv == convo) >= 0.9
        assert both.sum() >= 0.8 * len(conv)
        assert np.median(np.abs(e[both] - eo[both])) < 0.01
        # converged flag agrees with the reported grad norm (reference contract)
        assert np.array_equal(conv, r["grad_norm"].cpu().numpy() <= tol) or tol == 0.0


def test_metal_batch_form_matches_oracle(small_wl):
    """The mmff_optimize_metal_batch form (no convergence exit, no polish pass, L2 norm of the
    last scaled gradient) against the oracle's batch form: after 20 iterations the geometries,
    energies and norms agree; after 200 iterations, with the trajectories still moving, the
    best-seen energies are as close to the oracle's as the oracle's own are to a restart of it
    from 1e-6 A away (a measure of how fast nearby trajectories separate)."""
    eng = _eng()
    wl = small_wl
    db = eng.device_batch(wl.params, wl.conf_counts)
    xd = torch.from_numpy(wl.x).cuda()
    ob = oracle.OracleBatch(wl.params, wl.conf_counts)
    for iters in (20, 200):
        r = eng.multimol(db, xd, max_iters=iters, grad_tol=0.0, batch_mode=True)
        torch.cuda.synchronize()
        xo, eo, gno, pcio, convo, usedo = oracle.multi_mol(ob, wl.x.astype(np.float64), max_iters=iters,
                                                           grad_tol=0.0, batch_mode=True)
        x, e, gn = r["x"].cpu().numpy(), r["energy"].cpu().numpy(), r["grad_norm"].cpu().numpy()
        assert int(r["iters_used"].cpu()[0]) == iters == usedo
        assert not r["converged"].cpu().numpy().any() and (r["per_conf_iters"].cpu().numpy() == iters).all()
        print(f"\nbatch form {iters} it: max |dx| {np.abs(x - xo).max():.2e}, max |dE| {np.abs(e - eo).max():.2e}, "
              f"max rel |d gnorm| {np.max(np.abs(gn - gno) / gno):.2e}")
        if iters == 20:
            assert np.abs(x - xo).max() < 1e-3
            assert np.abs(e - eo).max() < 1e-3 + 1e-5 * np.abs(eo).max()
            assert np.allclose(gn, gno, rtol=1e-2, atol=1e-5)
        else:
            xp = wl.x.astype(np.float64) + np.random.default_rng(5).normal(0.0, 1e-6, wl.x.shape)
            ep = oracle.multi_mol(ob, xp, max_iters=iters, grad_tol=0.0, batch_mode=True)[1]
            spread = float(np.median(np.abs(ep - eo)))
            print(f"median |dE| gpu vs oracle {np.median(np.abs(e - eo)):.3e}, oracle vs its restart {spread:.3e}")
            assert np.median(np.abs(e - eo)) < max(0.01, 3.0 * spread)


def test_empty_batch_is_a_no_op():
    eng = _eng()
    from conftest import mol_params
    m, p, x = mol_params("CCO")
    db = eng.device_batch([p], [0])
    r = eng.NkRunner(db)
    res = r.run(torch.empty(0, dtype=torch.float32, device="cuda"))
    torch.cuda.synchronize()
    assert db.n_confs == 0 and r.launches == [] and res.energy.numel() == 0


def test_graph_replay_matches_eager(small_wl):
    """The captured HIP graph of a runner reproduces the eager run bit for bit, replay after
    replay, with other eager work and host synchronisations in between."""
    eng = _eng()
    wl = small_wl
    db = eng.device_batch(wl.params, wl.conf_counts)
    xd = torch.from_numpy(wl.x).cuda()
    r = eng.NkRunner(db, max_iters=2000, grad_tol=1e-4, ls_grad=True)
    xo = torch.empty_like(xd)
    res = r.run(xd, xo)
    torch.cuda.synchronize()
    x_ref, e_ref = xo.cpu().numpy().copy(), res.energy.cpu().numpy().copy()
    st_ref, c_ref = res.status.cpu().numpy().copy(), res.counters.cpu().numpy().copy()
    g = r.capture(xd, xo)
    for _ in range(3):
        xo.zero_()
        r.e.zero_()
        r.cnt.zero_()
        eng.energy_grad(db, xd)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(xo.cpu().numpy(), x_ref)
        assert np.array_equal(r.e.cpu().numpy(), e_ref)
        assert np.array_equal(r.st.cpu().numpy(), st_ref)
        assert np.array_equal(r.cnt.cpu().numpy(), c_ref)


def test_autotuned_buckets_cover_batch_and_agree(medium_wl):
    """autotune_buckets returns a bucket per populated size class whose minima match the
    default plan's."""
    eng = _eng()
    wl = medium_wl
    db = eng.device_batch(wl.params, wl.conf_counts)
    xd = torch.from_numpy(wl.x).cuda()
    buckets, timings = eng.autotune_buckets(db, xd, sample=16, max_iters=2000, grad_tol=1e-4, ls_grad=True)
    assert buckets and all(len(t) > 0 for t in timings.values())
    a = eng.NkRunner(db, buckets=buckets, max_iters=2000, grad_tol=1e-4, ls_grad=True).run(xd)
    b = eng.NkRunner(db, max_iters=2000, grad_tol=1e-4, ls_grad=True).run(xd)
    torch.cuda.synchronize()
    ea, eb = a.energy.cpu().numpy(), b.energy.cpu().numpy()
    assert np.median(np.abs(ea - eb)) < 1e-3
    assert (a.status.cpu().numpy() == 0).mean() >= 0.9
