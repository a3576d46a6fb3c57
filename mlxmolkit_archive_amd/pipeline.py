"""``generate_conformers_nk`` with the ``run_mmff=True`` stage on MI355X.

Drop-in for mlxmolkit/conformer_pipeline_v2.py:357-481. Stage 4 (MMFF) follows the
reference (:292-326): per chunk, each molecule's first conformer seeds the MMFF parameter
extraction (with ``mmff_variant``), ``mmff_minimize_nk`` minimizes every conformer with
``max_iters = 200 + 15 * complexity`` when ``mmff_max_iters <= 0`` (:123-126, :169-170)
and L-BFGS when any molecule has >= 150 atoms and ``mmff_use_lbfgs`` is None (:448-452);
the MMFF energy replaces the embedding energy and ``converged`` requires MMFF status 0.
If parameter extraction fails for a chunk, MMFF is skipped and the chunk is reported
unconverged (:312-316, :341).

The DG -> 4D -> 3D -> ETK embedding stages are outside this package's scope; start
geometries come from the package's seeded restraint embedder, run on the GPU
(``embed_device``, no noise), and stay there for the MMFF stage: only each molecule's first
conformer is copied to the host, for the 100 A non-bonded filter of the parameter
extraction (:296-305). Without MMFF the reported energy is NaN (there is no DG/ETK energy to
report) with converged = True.
"""
from __future__ import annotations

import time
from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

from .embed import embed_device
from .mmff_params import extract_mmff_params
from .mol import mol_from_smiles

LBFGS_ATOM_THRESHOLD = 150


@dataclass
class ConformerResult:
    """Result for one molecule (reference :41-46)."""
    n_atoms: int
    positions_3d: List[np.ndarray]
    energies: List[float]
    converged: List[bool]


@dataclass
class PipelineResult:
    """Result for the full N x k pipeline (reference :49-56)."""
    molecules: List[ConformerResult]
    total_conformers: int
    total_time: float
    n_batches: int


def _build_chunk_schedule(n_confs_per_mol: Sequence[int], max_confs_per_batch: int) -> List[List[tuple]]:
    """Greedy chunking of (molecule, conformer range) into GPU-sized batches (:95-116)."""
    chunks, cur, count = [], [], 0
    for mol_idx, k in enumerate(n_confs_per_mol):
        remaining, offset = k, 0
        while remaining > 0:
            take = min(remaining, max_confs_per_batch - count)
            cur.append((mol_idx, offset, offset + take))
            count += take
            offset += take
            remaining -= take
            if count >= max_confs_per_batch:
                chunks.append(cur)
                cur, count = [], 0
    if cur:
        chunks.append(cur)
    return chunks


def _auto_iters(max_atoms: int, base: int, scale: float) -> int:
    return max(base, int(base + scale * max_atoms))


def _max_confs_per_batch(n_atoms: Sequence[int], max_memory_bytes: Optional[int]) -> int:
    """Conformers per chunk from free device memory (the reference reads macOS vm_stat, :63-88)."""
    if max_memory_bytes is None:
        try:
            import torch
            free, _ = torch.cuda.mem_get_info()
            max_memory_bytes = free // 2
        except Exception:
            max_memory_bytes = 4 * 1024 ** 3
    n = 3 * int(np.mean(n_atoms)) if len(n_atoms) else 90
    per_conf = n * 4 * 16 + 1024
    return max(1, int(max_memory_bytes // per_conf))


def generate_conformers_nk(smiles_list: Sequence[str], n_confs_per_mol=10, *, max_confs_per_batch: Optional[int] = None,
                           max_memory_bytes: Optional[int] = None, dg_max_iters: int = 0, etk_max_iters: int = 0,
                           fourth_dim_weight: float = 0.1, chiral_weight: float = 1.0, variant: str = "ETKDGv2",
                           run_mmff: bool = False, mmff_max_iters: int = 0, mmff_use_lbfgs: Optional[bool] = None,
                           mmff_variant: str = "MMFF94", seed: int = 42) -> PipelineResult:
    """Generate N x k conformers; with ``run_mmff=True`` minimize them with MMFF94 on the GPU.

    Arguments keep the reference signature; the embedding-stage options (dg/etk iterations,
    weights, ETKDG variant) are accepted and ignored because that stage is not part of
    this package.
    """
    from .minimize import mmff_minimize_nk

    t0 = time.time()
    N = len(smiles_list)
    k_list = [n_confs_per_mol] * N if isinstance(n_confs_per_mol, int) else list(n_confs_per_mol)
    mols = []
    for i, smi in enumerate(smiles_list):
        try:
            mols.append(mol_from_smiles(smi))
        except Exception as ex:
            raise ValueError(f"Invalid SMILES at index {i}: {smi}") from ex
    n_atoms = [m.GetNumAtoms() for m in mols]
    if max_confs_per_batch is None:
        max_confs_per_batch = _max_confs_per_batch(n_atoms, max_memory_bytes)
    chunks = _build_chunk_schedule(k_list, max_confs_per_batch)
    if mmff_use_lbfgs is None:
        mmff_use_lbfgs = max(n_atoms) >= LBFGS_ATOM_THRESHOLD if n_atoms else False
    results = [ConformerResult(n_atoms=n, positions_3d=[], energies=[], converged=[]) for n in n_atoms]
    # restraint tables for the embedder come from MMFF-style bond/angle targets of each molecule
    embed_params = []
    for m in mols:
        m.set_conformers(np.zeros((1, m.GetNumAtoms(), 3)))
        embed_params.append(extract_mmff_params(m, mmff_variant=mmff_variant))
    total = 0
    for ci, chunk in enumerate(chunks):
        order, kk = [], {}
        for mol_idx, a, b in chunk:
            if mol_idx not in kk:
                kk[mol_idx] = 0
                order.append(mol_idx)
            kk[mol_idx] += b - a
        counts = [kk[m] for m in order]
        seeds = np.arange(sum(counts), dtype=np.uint64) + np.uint64(seed + 10000 * ci)
        xdev = embed_device([embed_params[m] for m in order], [mols[m].atomic_numbers() for m in order], counts,
                            seeds=seeds, noise=0.0)
        energies = np.full(sum(counts), np.nan, np.float64)
        conv = np.ones(sum(counts), bool)
        pos3 = None
        if run_mmff:
            import torch
            conv[:] = False
            plist = []
            # each molecule's first conformer, gathered on the device and copied in one transfer
            firsts, off = [], 0
            for m_idx, c in zip(order, counts):
                firsts.append(np.arange(off, off + 3 * n_atoms[m_idx]))
                off += 3 * n_atoms[m_idx] * c
            first_x = xdev[torch.from_numpy(np.concatenate(firsts)).to(xdev.device)].cpu().numpy()
            try:
                off = 0
                for m_idx in order:
                    n = n_atoms[m_idx]
                    mols[m_idx].set_conformers(first_x[off:off + 3 * n].reshape(1, n, 3))
                    plist.append(extract_mmff_params(mols[m_idx], mmff_variant=mmff_variant))
                    off += 3 * n
            except Exception:
                plist = None
            if plist is not None:
                iters = mmff_max_iters if mmff_max_iters > 0 else _auto_iters(max(n_atoms[m] for m in order), 200, 15.0)
                pos3, e32, conv, info = mmff_minimize_nk(plist, counts, xdev, max_iters=iters,
                                                         use_lbfgs=bool(mmff_use_lbfgs), return_info=True)
                energies = info["energy_f64"]
        if pos3 is None:
            pos3 = xdev.cpu().numpy()
        off = 0
        c = 0
        for m_idx, cnt in zip(order, counts):
            n = n_atoms[m_idx]
            for _ in range(cnt):
                results[m_idx].positions_3d.append(pos3[off:off + 3 * n].reshape(n, 3).copy())
                results[m_idx].energies.append(float(energies[c]))
                results[m_idx].converged.append(bool(conv[c]))
                off += 3 * n
                c += 1
        total += sum(counts)
    return PipelineResult(molecules=results, total_conformers=total, total_time=time.time() - t0,
                          n_batches=len(chunks))
