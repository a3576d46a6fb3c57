"""Drop-in for mlxmolkit/mmff_metal_optimizer.py on MI355X.

* ``mmff_optimize_metal_multi_mol`` (:217-535): lock-step L-BFGS with the reference's
  per-atom step cap, fixed gradient scale, best-x guard, sync every ``eval_freq``
  iterations, polish pass of ``max_iters // 4`` and the same result semantics
  (``per_conf_iters``, ``converged``, ``grad_norms`` of the last polish iterate,
  ``n_iters`` = iterations used by the main loop). Runs as a chain of device launches
  whose early exit is decided on the GPU (no host round trip per sync).
* ``mmff_optimize_metal_fused_multi_mol`` (:543-653): one launch of BB steepest descent;
  ``grad_norms = 0`` and ``n_iters = max_iters`` as in the reference (:646-647).
* ``mmff_optimize_metal_batch`` (:77-202): single molecule, the same device engine in its
  batch form: every one of ``max_iters`` iterations runs (no convergence exit), no polish
  pass, best-x guard at every sync point, ``grad_norms`` = ||g_sc||_2 of the last iterate
  (:186).

Conformers of the caller's molecules are written back in place (:514-519), vectorised per
conformer through ``SetPositions`` when available.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Optional, Tuple

import numpy as np

from . import _lib
from .engine import bb_fused, device_batch, multimol
from .mmff_params import extract_mmff_params


@dataclass
class MetalBatchResult:
    """Result of fused MMFF optimization (field layout of the reference :34-44)."""

    energies: np.ndarray
    positions: np.ndarray
    grad_norms: np.ndarray
    n_iters: int
    n_conformers: int
    per_conf_iters: Optional[np.ndarray] = None
    converged: Optional[np.ndarray] = None


@dataclass
class MultiMolResult:
    """Result of multi-molecule parallel MMFF optimisation (:210-214)."""

    mol_results: List[MetalBatchResult]


def _conf_ids(mol, conf_ids):
    return [c.GetId() for c in mol.GetConformers()] if conf_ids is None else list(conf_ids)


def _get_pos(mol, cid) -> np.ndarray:
    conf = mol.GetConformer(cid)
    if hasattr(conf, "GetPositions"):
        return np.asarray(conf.GetPositions(), dtype=np.float32)
    n = mol.GetNumAtoms()
    out = np.zeros((n, 3), np.float32)
    for a in range(n):
        p = conf.GetAtomPosition(a)
        out[a] = (p.x, p.y, p.z)
    return out


def _set_pos(mol, cid, pos: np.ndarray) -> None:
    conf = mol.GetConformer(cid)
    if hasattr(conf, "SetPositions"):
        conf.SetPositions(np.asarray(pos, np.float64))
        return
    for a in range(pos.shape[0]):
        conf.SetAtomPosition(a, pos[a].astype(float).tolist())


def _prepare(mol_specs):
    mols, all_ids, params, counts, xs = [], [], [], [], []
    for mol, conf_ids in mol_specs:
        ids = _conf_ids(mol, conf_ids)
        mols.append(mol)
        all_ids.append(ids)
        params.append(extract_mmff_params(mol, conf_id=ids[0]))  # MMFF94, as the reference (:275)
        counts.append(len(ids))
        for cid in ids:
            xs.append(_get_pos(mol, cid).reshape(-1))
    x = np.concatenate(xs).astype(np.float32) if xs else np.zeros(0, np.float32)
    return mols, all_ids, params, counts, x


def mmff_optimize_metal_multi_mol(mol_specs, max_iters: int = 200, eval_freq: int = 10,
                                  max_step_per_atom: float = 0.3, lbfgs_m: int = 10, grad_scale: float = 0.1,
                                  grad_tol: float = 1e-4) -> MultiMolResult:
    """Optimise many molecules' conformers in lock step on the GPU (reference :217-261)."""
    torch = _lib.require_gpu()
    mols, all_ids, params, counts, x_np = _prepare(mol_specs)
    db = device_batch(params, counts)
    x = torch.from_numpy(x_np).cuda()
    r = multimol(db, x, max_iters=max_iters, eval_freq=eval_freq, max_step_per_atom=max_step_per_atom,
                 lbfgs_m=lbfgs_m, grad_scale=grad_scale, grad_tol=grad_tol)
    torch.cuda.synchronize()
    xb = r["x"].cpu().numpy()
    e = r["energy"].cpu().numpy().astype(np.float32)
    gn = r["grad_norm"].cpu().numpy()
    pci = r["per_conf_iters"].cpu().numpy()
    conv = r["converged"].cpu().numpy().astype(bool)
    used = int(r["iters_used"].cpu().numpy()[0])
    return _unpack(mols, all_ids, params, counts, xb, e, gn, used, pci, conv)


def _unpack(mols, all_ids, params, counts, xb, e, gn, n_iters, pci, conv) -> MultiMolResult:
    results = []
    off = 0
    cc = 0
    for m, mol in enumerate(mols):
        C, N = counts[m], params[m].n_atoms
        pos = xb[off:off + C * 3 * N].reshape(C, N, 3)
        for k, cid in enumerate(all_ids[m]):
            _set_pos(mol, cid, pos[k])
        sl = slice(cc, cc + C)
        results.append(MetalBatchResult(
            energies=e[sl], positions=pos, grad_norms=gn[sl] if gn is not None else e[sl] * 0, n_iters=n_iters,
            n_conformers=C, per_conf_iters=None if pci is None else pci[sl].copy(),
            converged=None if conv is None else conv[sl].copy()))
        off += C * 3 * N
        cc += C
    return MultiMolResult(mol_results=results)


def mmff_optimize_metal_fused_multi_mol(mol_specs, max_iters: int = 500, max_step: float = 0.3,
                                        grad_lr: float = 0.1) -> MultiMolResult:
    """One launch of BB steepest descent for every conformer (reference :543-653)."""
    torch = _lib.require_gpu()
    mols, all_ids, params, counts, x_np = _prepare(mol_specs)
    db = device_batch(params, counts)
    x = torch.from_numpy(x_np).cuda()
    xo, e = bb_fused(db, x, max_iters=max_iters, max_step=max_step, grad_lr=grad_lr)
    torch.cuda.synchronize()
    e32 = e.cpu().numpy().astype(np.float32)
    return _unpack(mols, all_ids, params, counts, xo.cpu().numpy(), e32, e32 * 0, max_iters, None, None)


def mmff_optimize_metal_batch(mol, conf_ids=None, max_iters: int = 200, eval_freq: int = 10,
                              max_step_per_atom: float = 0.3, lbfgs_m: int = 10,
                              grad_scale: float = 0.1) -> MetalBatchResult:
    """Single molecule, many conformers (reference :77-202): L-BFGS for exactly ``max_iters``
    iterations with the best-x guard every ``eval_freq``, the best-seen positions written back
    to the conformers, and ``grad_norms`` = ||g_sc||_2 of the last iterate."""
    torch = _lib.require_gpu()
    mols, all_ids, params, counts, x_np = _prepare([(mol, conf_ids)])
    db = device_batch(params, counts)
    x = torch.from_numpy(x_np).cuda()
    r = multimol(db, x, max_iters=max_iters, eval_freq=eval_freq, max_step_per_atom=max_step_per_atom,
                 lbfgs_m=lbfgs_m, grad_scale=grad_scale, grad_tol=0.0, batch_mode=True)
    torch.cuda.synchronize()
    e = r["energy"].cpu().numpy().astype(np.float32)
    res = _unpack(mols, all_ids, params, counts, r["x"].cpu().numpy(), e, r["grad_norm"].cpu().numpy(), max_iters,
                  None, None)
    return res.mol_results[0]
