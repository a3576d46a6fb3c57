"""GPU tests of the learned preconditioner: the MLP kernels against the FP64 oracle, the
finite-difference Hessian blocks against the oracle's, and the trained network's effect on
BFGS iteration counts."""
from __future__ import annotations

import numpy as np
import pytest

import mlp_oracle as mo
import oracle

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")


def _model(seed=0, scale=1.0):
    from mlxmolkit_archive_amd.precond import PrecondMLP
    m = PrecondMLP(seed=seed)
    if scale != 1.0:
        m.params.mul_(scale)
    return m


def _feat(n, seed):
    g = np.random.default_rng(seed)
    f = np.zeros((n, mo.N_IN), np.float32)
    f[np.arange(n), g.integers(0, 99, n)] = 1.0
    f[:, 99:] = g.uniform(0, 2, (n, 5))
    return f


def test_mlp_forward_backward_match_oracle():
    m = _model(1)
    # larger output weights so that every parameter group carries gradient
    m.params[-(6 * 64 + 6):] = torch.from_numpy(np.random.default_rng(5).normal(0, 0.3, 6 * 64 + 6).astype(np.float32)).cuda()
    n = 300
    f = _feat(n, 2)
    g = np.random.default_rng(3)
    t = g.normal(0, 0.5, (n, 9)).astype(np.float32)
    w = g.uniform(0.5, 2.0, n).astype(np.float32)
    fd, td, wd = (torch.from_numpy(a).cuda() for a in (f, t, w))
    loss, grad, blocks = m.loss_grad(fd, td, wd)
    torch.cuda.synchronize()
    p = m.params.cpu().numpy().astype(np.float64)
    Pr, _ = mo.forward(p, f)
    tot, per, gr = mo.loss_grad(p, f, t, w)
    np.testing.assert_allclose(blocks.cpu().numpy(), Pr, rtol=1e-4, atol=1e-5)
    np.testing.assert_allclose(loss.cpu().numpy(), per, rtol=1e-4, atol=1e-5)
    gg = grad.cpu().numpy()
    assert np.max(np.abs(gg - gr)) <= 1e-4 * np.max(np.abs(gr))


def test_mlp_bf16_mfma_inference_matches_fp32():
    """BF16 operands, FP32 accumulation: blocks within bf16 rounding of the FP64 network;
    77 atoms exercise a partial 32-atom tile."""
    m = _model(4)
    m.params[-(6 * 64 + 6):] = torch.from_numpy(np.random.default_rng(6).normal(0, 0.2, 6 * 64 + 6).astype(np.float32)).cuda()
    for n in (77, 1000):
        f = _feat(n, 7 + n)
        b = m.blocks(torch.from_numpy(f).cuda())
        torch.cuda.synchronize()
        Pr, _ = mo.forward(m.params.cpu().numpy().astype(np.float64), f)
        err = np.abs(b.cpu().numpy() - Pr).max(axis=1) / np.abs(Pr).max(axis=1)
        assert np.median(err) < 1e-2 and err.max() < 5e-2, (n, np.median(err), err.max())


def test_hessian_blocks_match_oracle_finite_differences(small_wl):
    from mlxmolkit_archive_amd import engine
    from mlxmolkit_archive_amd.precond import hessian_diag_blocks
    wl = small_wl
    db = engine.device_batch(wl.params[:3], [1, 1, 1])
    n3 = [3 * p.n_atoms for p in wl.params[:3]]
    off = np.concatenate([[0], np.cumsum([3 * p.n_atoms * k for p, k in zip(wl.params, wl.conf_counts)])])
    x = np.concatenate([wl.x[off[i]:off[i] + n3[i]] for i in range(3)]).astype(np.float32)
    hb = hessian_diag_blocks(db, torch.from_numpy(x).cuda()).cpu().numpy()
    a0 = 0
    h = 1e-4
    for i, p in enumerate(wl.params[:3]):
        xi = x[sum(n3[:i]):sum(n3[:i]) + n3[i]].astype(np.float64)
        for a in (0, p.n_atoms // 2, p.n_atoms - 1):
            ref = np.zeros((3, 3))
            for k in range(3):
                xp, xm = xi.copy(), xi.copy()
                xp[3 * a + k] += h
                xm[3 * a + k] -= h
                gp = oracle.energy_grad(p, xp)[1]
                gm = oracle.energy_grad(p, xm)[1]
                ref[k] = (gp[3 * a:3 * a + 3] - gm[3 * a:3 * a + 3]) / (2 * h)
            ref = 0.5 * (ref + ref.T)
            got = hb[a0 + a]
            assert np.abs(got - ref).max() <= 2e-2 * np.abs(ref).max() + 0.5, (i, a, got, ref)
        a0 += p.n_atoms


def test_learned_precond_trains_and_keeps_bfgs_on_track():
    """Train on one seeded set of molecules, run held-out molecules with plain BFGS, with the
    learned initial inverse Hessian, and with its control: identity blocks scaled by 0.1
    through the same path (the learned start multiplies the network's blocks by 0.2, and a
    smaller first step alone saves iterations). The training loss must fall; the learned start
    must converge as often as plain BFGS, without more iterations than plain BFGS or than the
    control, with fewer energy evaluations than the control (the network's gain: its blocks'
    first steps need fewer line-search trials), and with mean final energies within 2 % of
    plain BFGS's."""
    from mlxmolkit_archive_amd import engine, molgen, precond, workloads
    tr = workloads.build("medium", n_mols=120, n_confs=2)
    te = workloads.build("medium", n_mols=60, n_confs=2, seed=molgen.SEED_BASE + 777)
    model, lg = precond.fit(tr, steps=500)
    assert lg.losses[-1] < 0.2 * lg.losses[0]
    db = engine.device_batch(te.params, te.conf_counts)
    x = torch.from_numpy(te.x).cuda()
    pc = precond.precond_for_batch(model, te.mols, te.params, te.conf_counts)
    n_at = int(te.atoms_per_conf().sum())
    ctrl = torch.from_numpy(np.tile(0.1 * np.eye(3, dtype=np.float32).reshape(1, 9), (n_at, 1))).cuda()
    out = {}
    for name, kw in (("plain", {}), ("learned", {"precond": pc}), ("identity x 0.1", {"precond": ctrl})):
        r = engine.NkRunner(db, max_iters=2000, grad_tol=1e-4, ls_grad=True, **kw).run(x)
        torch.cuda.synchronize()
        st = r.status.cpu().numpy()
        cnt = r.counters.cpu().numpy()
        ok = st == 0
        out[name] = dict(status0=float(ok.mean()), iters=float(cnt[ok, 0].mean()),
                         evals=float((cnt[ok, 1] + cnt[ok, 2]).mean()), mean_e=float(r.energy.cpu().numpy().mean()))
        print(name, {k: round(v, 3) for k, v in out[name].items()}, flush=True)
    pl, le, co = out["plain"], out["learned"], out["identity x 0.1"]
    print(f"learned vs plain: iterations x{pl['iters'] / le['iters']:.3f}, evaluations x{pl['evals'] / le['evals']:.3f}; "
          f"learned vs control: iterations x{co['iters'] / le['iters']:.3f}, evaluations x{co['evals'] / le['evals']:.3f}")
    assert le["status0"] >= pl["status0"] - 0.02
    assert le["iters"] <= 1.1 * pl["iters"]
    assert le["iters"] <= 1.05 * co["iters"]
    assert le["evals"] <= co["evals"]
    # which local minimum a floppy molecule reaches varies with the start of the search; the
    # mean final energy must stay within 2 % of plain BFGS's
    assert le["mean_e"] <= pl["mean_e"] + 0.02 * abs(pl["mean_e"])


def test_preconditioned_bfgs_matches_oracle(small_wl):
    """GPU BFGS from given per-atom blocks against the FP64 oracle from the same blocks:
    the same 0.1 kcal/mol bar as plain BFGS."""
    from mlxmolkit_archive_amd import engine
    wl = small_wl
    g = np.random.default_rng(11)
    n_at = int(wl.atoms_per_conf().sum())
    a = g.normal(0, 0.15, (n_at, 3, 3))
    blocks = (0.4 * np.eye(3)[None] + a @ a.transpose(0, 2, 1)).reshape(-1, 9).astype(np.float32)
    db = engine.device_batch(wl.params, wl.conf_counts)
    xd = torch.from_numpy(wl.x).cuda()
    pc = torch.from_numpy(blocks).cuda()
    res = engine.NkRunner(db, max_iters=2000, grad_tol=1e-4, precond=pc).run(xd)
    torch.cuda.synchronize()
    ob = oracle.OracleBatch(wl.params, wl.conf_counts)
    _, e_ref, st_ref, _ = oracle.minimize_nk(ob, wl.x.astype(np.float64), max_iters=2000, grad_tol=1e-4,
                                             h0=blocks.astype(np.float64))
    e_gpu, _, _ = oracle.energy_grad_batch(ob, res.x.cpu().numpy().astype(np.float64), family=oracle.FAMILY_TG)
    st = res.status.cpu().numpy()
    ok = (st == 0) & (st_ref == 0)
    dE = np.abs(e_gpu - e_ref)[ok]
    assert ok.mean() >= 0.9
    assert np.mean(dE <= 0.1) >= 0.85 and np.median(dE) < 5e-3
