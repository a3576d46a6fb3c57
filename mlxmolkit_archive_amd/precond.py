"""Learned per-atom initial inverse Hessian for the dense BFGS minimizer.

A per-atom MLP (``MLP_IN`` = 104 features -> 64 -> 64 -> 6, ReLU) predicts the Cholesky
factor of a 3x3 SPD block. ``k_minimize_nk`` starts BFGS from these blocks instead of the
identity. The blocks are scaled by 0.1 / (initial gradient scale), so that they act on the
unscaled gradient, and the minimizer restarts from the identity if a line search fails.
Everything runs on the GPU through the C ABI:

* inference: hand-written BF16 MFMA kernel with FP32 accumulation (``mmffhip_mlp_infer``);
* training: FP32 forward / backward kernels (``mmffhip_mlp_forward_backward``) and Adam
  (elementwise torch updates);
* targets: per-atom 3x3 diagonal blocks of the inverse of the exact MMFF Hessian at
  minimized geometries (the Hessian from central differences of the GPU gradient kernel);
  the inverse of each atom's own Hessian block is available as an alternative target.

``fit`` runs the whole recipe: minimize a training batch, build targets, train.

Features per atom (104): one-hot MMFF numeric type 1..99 (99), degree / 4, in-ring flag,
aromatic flag, sum of incident bond force constants / 10, and sum of force constants of
angles centred on the atom / 2.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from . import _lib, engine
from .mmff_typer import MMFFTyper

MLP_IN = 104
N_TYPES = 99
TARGET_SCALE = 0.1      # blocks approximate (0.1 * Hessian block)^-1, the identity's units
EIG_CLAMP = (1.0, 5000.0)  # Hessian block eigenvalues (kcal/mol/A^2) are clamped to this range


def atom_features(mol, params) -> np.ndarray:
    """(N, 104) float32 feature matrix of one molecule (package Mol + its MMFFParams)."""
    t = MMFFTyper(mol)
    n = mol.GetNumAtoms()
    f = np.zeros((n, MLP_IN), np.float32)
    types = np.asarray(t.types, dtype=np.int64)
    ok = (types >= 1) & (types <= N_TYPES)
    f[np.nonzero(ok)[0], types[ok] - 1] = 1.0
    f[:, 99] = np.array([len(list(mol.neighbors(i))) for i in range(n)], np.float32) / 4.0
    f[:, 100] = (t.ring_size > 0).astype(np.float32)
    f[:, 101] = np.asarray(t.arom, dtype=np.float32)[:n]
    kb = np.zeros(n, np.float64)
    np.add.at(kb, params.bond_idx1, params.bond_kb)
    np.add.at(kb, params.bond_idx2, params.bond_kb)
    ka = np.zeros(n, np.float64)
    np.add.at(ka, params.angle_idx2, params.angle_ka)
    f[:, 102] = (kb / 10.0).astype(np.float32)
    f[:, 103] = (ka / 2.0).astype(np.float32)
    return f


def conformer_features(mols: Sequence, params_list: Sequence, conf_counts: Sequence[int]) -> np.ndarray:
    """Features of every atom of every conformer, in the coordinate order of the batch."""
    per_mol = [atom_features(m, p) for m, p in zip(mols, params_list)]
    return np.concatenate([np.repeat(f[None], k, axis=0).reshape(-1, MLP_IN) for f, k in zip(per_mol, conf_counts)])


def _dims():
    L = _lib.load()
    a = [C.c_int32() for _ in range(4)]
    _lib.check(L.mmffhip_mlp_dims(*[C.byref(v) for v in a]))
    return tuple(v.value for v in a)


class PrecondMLP:
    """The per-atom network; ``params`` is one flat float32 device tensor
    (W1 64x104, b1, W2 64x64, b2, W3 6x64, b3, row-major)."""

    def __init__(self, seed: int = 0, device: str = "cuda"):
        torch = _lib.require_gpu()
        n_in, n_h, n_out, n_par = _dims()
        assert n_in == MLP_IN
        self.n_params = n_par
        g = np.random.default_rng(seed)
        w1 = g.normal(0, np.sqrt(2.0 / n_in), (n_h, n_in))
        w2 = g.normal(0, np.sqrt(2.0 / n_h), (n_h, n_h))
        w3 = g.normal(0, 0.01, (n_out, n_h))
        # start near the identity's scale: softplus(b) ~ 1 on the diagonal
        b3 = np.array([0.55, 0.0, 0.55, 0.0, 0.0, 0.55])
        flat = np.concatenate([w1.ravel(), np.zeros(n_h), w2.ravel(), np.zeros(n_h), w3.ravel(), b3])
        assert flat.size == n_par
        self.params = torch.from_numpy(flat.astype(np.float32)).to(device)

    def blocks(self, feat, stream=None):
        """(n_atoms, 9) float32 blocks from the BF16 MFMA inference kernel."""
        torch = _lib.require_gpu()
        n = int(feat.shape[0])
        out = torch.empty((n, 9), dtype=torch.float32, device=feat.device)
        L = _lib.load()
        _lib.check(L.mmffhip_mlp_infer(self.params.data_ptr(), feat.data_ptr(), n, out.data_ptr(),
                                       _lib.stream_ptr(stream)))
        return out

    def loss_grad(self, feat, target, weight=None, want_grad: bool = True, stream=None):
        """FP32 forward (+ backward): (per-atom loss, gradient of the weighted loss sum, blocks)."""
        torch = _lib.require_gpu()
        n = int(feat.shape[0])
        L = _lib.load()
        nbytes = C.c_size_t()
        _lib.check(L.mmffhip_mlp_workspace(n, C.byref(nbytes)))
        ws = torch.empty(int(nbytes.value), dtype=torch.uint8, device=feat.device)
        loss = torch.empty(n, dtype=torch.float32, device=feat.device)
        blocks = torch.empty((n, 9), dtype=torch.float32, device=feat.device)
        grad = torch.empty(self.n_params, dtype=torch.float32, device=feat.device) if want_grad else None
        _lib.check(L.mmffhip_mlp_forward_backward(self.params.data_ptr(), feat.data_ptr(), n, target.data_ptr(),
                                                  0 if weight is None else weight.data_ptr(), blocks.data_ptr(),
                                                  loss.data_ptr(), 0 if grad is None else grad.data_ptr(),
                                                  ws.data_ptr(), int(nbytes.value), _lib.stream_ptr(stream)))
        return loss, grad, blocks


def hessian_diag_blocks(db, x, h: float = 2e-3, chunk_atoms: int = 40000):
    """Exact-model 3x3 diagonal Hessian blocks (sum_c N_c, 3, 3) at coordinates x, by central
    differences of the GPU gradient: atom a's block column k is
    (g_a(x + h e_ak) - g_a(x - h e_ak)) / 2h. The 6 evaluations per atom are replicas of its
    conformer (same molecule tables) in one batched energy_grad call per chunk of atoms."""
    torch = _lib.require_gpu()
    hb = db.host
    n_atoms = hb.n_atoms[hb.conf_mol].astype(np.int64)
    total = int(n_atoms.sum())
    conf_of_atom = np.repeat(np.arange(hb.n_confs), n_atoms)
    atom_local = np.arange(total) - np.repeat(np.cumsum(n_atoms) - n_atoms, n_atoms)
    x = x.to(torch.float32)
    out = torch.empty((total, 3, 3), dtype=torch.float32, device=x.device)
    for a0 in range(0, total, chunk_atoms):
        a1 = min(total, a0 + chunk_atoms)
        ca, la = conf_of_atom[a0:a1], atom_local[a0:a1]
        reps = np.repeat(ca, 6)  # (atom, coordinate k, sign) replicas of each atom's conformer
        sub = _subset(hb, reps)
        n3 = 3 * hb.n_atoms[sub.conf_mol].astype(np.int64)
        sub_off = sub.conf_off.astype(np.int64)
        src = np.arange(int(n3.sum())) - np.repeat(sub_off, n3) + np.repeat(hb.conf_off[reps].astype(np.int64), n3)
        xr = x[torch.from_numpy(src).to(x.device)].clone()
        kk = np.tile(np.repeat(np.arange(3), 2), len(ca))
        sg = np.tile(np.array([1.0, -1.0] * 3), len(ca))
        pos = sub_off + 3 * np.repeat(la, 6) + kk
        xr[torch.from_numpy(pos).to(x.device)] += torch.from_numpy((h * sg).astype(np.float32)).to(x.device)
        from .packing import DeviceBatch
        _, g, _ = engine.energy_grad(DeviceBatch(sub, device=x.device), xr, grad=True)
        rows = (sub_off + 3 * np.repeat(la, 6))[:, None] + np.arange(3)[None, :]
        ga = g[torch.from_numpy(rows).to(x.device)].reshape(len(ca), 3, 2, 3)  # [atom, k, sign, row]
        blk = (ga[:, :, 0, :] - ga[:, :, 1, :]) / (2.0 * h)                   # [atom, column k, row]
        out[a0:a1] = 0.5 * (blk + blk.transpose(1, 2))
    return out


def inverse_hessian_diag_blocks(db, x, h: float = 2e-3, floor: float = 2.0, chunk_confs: int = 256):
    """Per-atom 3x3 diagonal blocks (sum_c N_c, 3, 3) of the inverse of the full MMFF Hessian
    at x: the Hessian of each conformer is built column by column from central differences
    of the GPU gradient (2n replicas per conformer), symmetrised, its eigenvalues floored at
    ``floor`` kcal/mol/A^2 (this also regularises the six rigid-body modes), inverted, and
    its diagonal blocks kept. Unlike the inverse of each atom's own block, these blocks carry
    the soft collective modes (torsions) that dominate a minimization's path."""
    torch = _lib.require_gpu()
    hb = db.host
    n_at = hb.n_atoms[hb.conf_mol].astype(np.int64)
    x = x.to(torch.float32)
    out = torch.empty((int(n_at.sum()), 3, 3), dtype=torch.float32, device=x.device)
    atom0 = np.concatenate([[0], np.cumsum(n_at)[:-1]])
    from .packing import DeviceBatch
    for c0 in range(0, hb.n_confs, chunk_confs):
        cs = np.arange(c0, min(hb.n_confs, c0 + chunk_confs))
        n3c = 3 * n_at[cs]
        reps = np.repeat(cs, 2 * n3c)                     # (conformer, coordinate, sign)
        sub = _subset(hb, reps)
        n3 = 3 * hb.n_atoms[sub.conf_mol].astype(np.int64)
        sub_off = sub.conf_off.astype(np.int64)
        src = np.arange(int(n3.sum())) - np.repeat(sub_off, n3) + np.repeat(hb.conf_off[reps].astype(np.int64), n3)
        xr = x[torch.from_numpy(src).to(x.device)].clone()
        coord = np.concatenate([np.repeat(np.arange(k), 2) for k in n3c])
        sg = np.tile([1.0, -1.0], len(reps) // 2)
        xr[torch.from_numpy(sub_off + coord).to(x.device)] += torch.from_numpy((h * sg).astype(np.float32)).to(x.device)
        _, g, _ = engine.energy_grad(DeviceBatch(sub, device=x.device), xr, grad=True)
        r0 = 0
        for c, k in zip(cs, n3c):
            o0 = int(sub_off[r0])  # the 2k replicas of a conformer are contiguous, k floats each
            gg = g[o0:o0 + 2 * k * k].to(torch.float64).reshape(k, 2, k)
            Hc = (gg[:, 0, :] - gg[:, 1, :]) / (2.0 * h)  # row = perturbed coordinate
            Hc = 0.5 * (Hc + Hc.T)
            ev, V = torch.linalg.eigh(Hc)
            Hi = (V * (1.0 / ev.clamp(min=floor))[None, :]) @ V.T
            nat = k // 3
            blk = Hi.reshape(nat, 3, nat, 3).diagonal(dim1=0, dim2=2).permute(2, 0, 1)
            out[atom0[c]:atom0[c] + nat] = blk.to(torch.float32)
            r0 += 2 * k
    return out


def targets_from_inverse_blocks(blocks):
    """Training targets from diagonal blocks of the inverse Hessian, in (0.1 * Hessian)^-1
    units (x 10), and per-atom loss weights 1 / mean(eig(T))^2."""
    torch = _lib.require_gpu()
    t = (0.5 * (blocks + blocks.transpose(1, 2))).to(torch.float64) / TARGET_SCALE
    ev = torch.linalg.eigvalsh(t)
    w = 1.0 / ev.clamp(min=1e-6).mean(dim=1) ** 2
    return t.reshape(-1, 9).to(torch.float32), (w / w.mean()).to(torch.float32)


def _subset(hb, confs):
    """HostBatch of the listed conformers (tables shared, coordinates re-packed densely)."""
    import dataclasses
    n3 = 3 * hb.n_atoms[hb.conf_mol[confs]].astype(np.int64)
    off = np.concatenate([[0], np.cumsum(n3)[:-1]]).astype(np.int32)
    return dataclasses.replace(hb, conf_mol=hb.conf_mol[confs].astype(np.int32), conf_off=off,
                               total_coords=int(n3.sum()),
                               conf_counts=np.bincount(hb.conf_mol[confs], minlength=hb.n_mols).astype(np.int32))


def targets_from_hessian(blocks):
    """Training targets (0.1 * clamp(B))^-1 and per-atom loss weights 1 / mean(eig(T))^2."""
    torch = _lib.require_gpu()
    b = 0.5 * (blocks + blocks.transpose(1, 2)).to(torch.float64)
    ev, V = torch.linalg.eigh(b)
    ev = ev.clamp(EIG_CLAMP[0], EIG_CLAMP[1])
    t = (V * (1.0 / (TARGET_SCALE * ev))[:, None, :]) @ V.transpose(1, 2)
    w = 1.0 / (1.0 / (TARGET_SCALE * ev)).mean(dim=1) ** 2
    return t.reshape(-1, 9).to(torch.float32), (w / w.mean()).to(torch.float32)


@dataclass
class TrainLog:
    losses: List[float]
    n_atoms: int


def train(model: PrecondMLP, feat, target, weight, steps: int = 400, lr: float = 3e-3, batch: int = 8192,
          seed: int = 0) -> TrainLog:
    """Adam on the weighted block loss, mini-batches of atoms; gradients from the GPU
    backward kernel."""
    torch = _lib.require_gpu()
    n = int(feat.shape[0])
    m = torch.zeros_like(model.params)
    v = torch.zeros_like(model.params)
    b1, b2, eps = 0.9, 0.999, 1e-8
    gen = torch.Generator(device="cpu").manual_seed(seed)
    losses = []
    for it in range(1, steps + 1):
        idx = torch.randint(0, n, (min(batch, n),), generator=gen).to(feat.device)
        f, t, w = feat[idx].contiguous(), target[idx].contiguous(), weight[idx].contiguous()
        loss, g, _ = model.loss_grad(f, t, w)
        g = g / len(idx)
        m.mul_(b1).add_(g, alpha=1 - b1)
        v.mul_(b2).addcmul_(g, g, value=1 - b2)
        mh = m / (1 - b1 ** it)
        vh = v / (1 - b2 ** it)
        model.params.sub_(lr * mh / (vh.sqrt() + eps))
        if it % 50 == 0 or it == 1:
            losses.append(float((loss * w).mean()))
    return TrainLog(losses=losses, n_atoms=n)


# Multiplier on the network's blocks when they start a minimization. The blocks approximate
# diagonal blocks of the inverse Hessian at a minimum; from a start geometry the full step they
# imply overshoots and the line search pays for it. Measured on the medium workload (1,000 x
# 10, network trained on a separate 300 x 2 set):
#   scale 1.0: 311.6 mean iterations, 1.76 evaluations per iteration, 176 ms per step
#   scale 0.2: 299.3 mean iterations, 1.37 evaluations per iteration, 158 ms per step
#   plain BFGS: 320.9 mean iterations, 1.40 evaluations per iteration, 173 ms per step
# (an identity start scaled by 0.1 gives 299.0, 1.46 and 166 ms).
START_SCALE = 0.2


def precond_for_batch(model: PrecondMLP, mols: Sequence, params_list: Sequence, conf_counts: Sequence[int],
                      device: str = "cuda", scale: float = START_SCALE):
    """(sum_c N_c, 9) device tensor in the layout NkRunner(precond=...) expects: the network's
    blocks times ``scale``."""
    torch = _lib.require_gpu()
    feat = torch.from_numpy(conformer_features(mols, params_list, conf_counts)).to(device)
    return (model.blocks(feat) * float(scale)).contiguous()


def fit(train_wl, steps: int = 800, lr: float = 3e-3, batch: int = 8192, seed: int = 0, target: str = "block-of-inverse",
        **nk_opts):
    """Train a PrecondMLP on a workload (mols, params, conf_counts, x): minimize it with plain
    BFGS on the GPU, take the converged conformers' Hessian targets, train. Returns
    (model, TrainLog)."""
    torch = _lib.require_gpu()
    db = engine.device_batch(train_wl.params, train_wl.conf_counts)
    x = torch.from_numpy(train_wl.x).cuda()
    opts = dict(max_iters=2000, grad_tol=1e-4, ls_grad=True)
    opts.update(nk_opts)
    res = engine.NkRunner(db, **opts).run(x)
    if target == "block-of-inverse":
        t, w = targets_from_inverse_blocks(inverse_hessian_diag_blocks(db, res.x))
    else:
        t, w = targets_from_hessian(hessian_diag_blocks(db, res.x))
    # atoms of conformers that did not converge do not train the network
    ok = np.repeat(res.status.cpu().numpy() == 0, train_wl.atoms_per_conf())
    w = w * torch.from_numpy(ok.astype(np.float32)).to(w.device)
    feat = torch.from_numpy(conformer_features(train_wl.mols, train_wl.params, train_wl.conf_counts)).to(x.device)
    model = PrecondMLP(seed=seed)
    log = train(model, feat, t, w, steps=steps, lr=lr, batch=batch, seed=seed)
    return model, log
