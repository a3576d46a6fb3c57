"""Start geometries for the minimizer workloads (wraps csrc/embed.c).

Restraints per molecule come from its MMFF tables: every bond gives a 1-2 target r0,
every angle a 1-3 target from the law of cosines with theta0, and every other pair a
lower bound (sum-of-radii style wall). Conformer k of molecule m is seeded from
(SEED_BASE, family id, m, k), embedded, then perturbed by N(0, noise) per coordinate
(0.1 A for the benchmark workloads, following their definition of "embedding + noise").
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path
from typing import List, Optional, Sequence

import numpy as np

from ._build import lib_path

_LIB = None


class _EmbedMol(C.Structure):
    _fields_ = [("n_atoms", C.c_int32), ("n_restr", C.c_int32), ("ri", C.c_void_p), ("rj", C.c_void_p),
                ("rd", C.c_void_p), ("rkind", C.c_void_p)]


def _lib():
    global _LIB
    if _LIB is None:
        p = lib_path("libmmff_embed.so")
        if not p.exists():
            raise RuntimeError(f"{p} is missing: run __graft_entry__.build()")
        L = C.CDLL(str(p))
        L.mmff_embed_batch.restype = None
        L.mmff_embed_batch.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_float,
                                       C.c_void_p, C.c_int]
        _LIB = L
    return _LIB


def restraints(params, z: np.ndarray):
    """(ri, rj, d, kind) restraint arrays for one molecule's MMFFParams."""
    n = params.n_atoms
    tgt = {}
    for i, j, r0 in zip(params.bond_idx1, params.bond_idx2, params.bond_r0):
        tgt[(min(i, j), max(i, j))] = float(r0)
    r0map = dict(tgt)
    for i, j, k, t0 in zip(params.angle_idx1, params.angle_idx2, params.angle_idx3, params.angle_theta0):
        a = r0map.get((min(i, j), max(i, j)), 1.5)
        b = r0map.get((min(k, j), max(k, j)), 1.5)
        th = np.deg2rad(float(t0))
        key = (min(i, k), max(i, k))
        if key not in tgt:
            tgt[key] = float(np.sqrt(a * a + b * b - 2 * a * b * np.cos(th)))
    heavy = z > 1
    ri, rj, d, kind = [], [], [], []
    for (i, j), dd in tgt.items():
        ri.append(i), rj.append(j), d.append(dd), kind.append(0)
    iu, ju = np.triu_indices(n, 1)
    if tgt:
        keys = np.array([a * n + b for a, b in tgt], dtype=np.int64)
        mask = ~np.isin(iu.astype(np.int64) * n + ju, keys)
        iu, ju = iu[mask], ju[mask]
    hh = heavy[iu].astype(int) + heavy[ju].astype(int)
    lb = np.where(hh == 2, 3.0, np.where(hh == 1, 2.55, 2.2))
    ri = np.concatenate([np.array(ri, np.int32), iu.astype(np.int32)])
    rj = np.concatenate([np.array(rj, np.int32), ju.astype(np.int32)])
    d = np.concatenate([np.array(d, np.float32), lb.astype(np.float32)])
    kind = np.concatenate([np.array(kind, np.int8), np.ones(len(iu), np.int8)])
    return ri, rj, d, kind


def embed(params_list: Sequence, z_list: Sequence[np.ndarray], conf_counts: Sequence[int],
          seeds: Sequence[int], noise: float = 0.1, n_threads: int = 0) -> np.ndarray:
    """Flat float32 coordinates (sum over conformers of 3N) for a ragged batch."""
    import os
    keep = []
    mols = (_EmbedMol * len(params_list))()
    for k, (p, z) in enumerate(zip(params_list, z_list)):
        ri, rj, d, kind = restraints(p, np.asarray(z))
        keep += [ri, rj, d, kind]
        mols[k].n_atoms = p.n_atoms
        mols[k].n_restr = len(ri)
        mols[k].ri, mols[k].rj = ri.ctypes.data, rj.ctypes.data
        mols[k].rd, mols[k].rkind = d.ctypes.data, kind.ctypes.data
    conf_mol = np.repeat(np.arange(len(params_list), dtype=np.int32), np.asarray(conf_counts, int))
    n3 = np.array([3 * params_list[m].n_atoms for m in conf_mol], dtype=np.int64)
    off = np.zeros(len(conf_mol), dtype=np.int64)
    if len(off) > 1:
        off[1:] = np.cumsum(n3)[:-1]
    out = np.zeros(int(n3.sum()), dtype=np.float32)
    sd = np.ascontiguousarray(np.asarray(seeds, dtype=np.uint64))
    assert len(sd) == len(conf_mol)
    if n_threads <= 0:
        try:
            n_threads = len(os.sched_getaffinity(0))
        except AttributeError:  # pragma: no cover
            n_threads = os.cpu_count() or 1
    _lib().mmff_embed_batch(mols, conf_mol.ctypes.data, off.ctypes.data, sd.ctypes.data, len(conf_mol),
                            float(noise), out.ctypes.data, int(n_threads))
    return out


def restraint_table(params, z: np.ndarray):
    """One molecule's restraints in the device embedder's encoding: (ij uint32 = i | j << 16,
    d float32 = d0 for an equality spring or -d_min for a lower-bound wall)."""
    ri, rj, d, kind = restraints(params, np.asarray(z))
    ij = (ri.astype(np.uint32) & 0xFFFF) | ((rj.astype(np.uint32) & 0xFFFF) << 16)
    return ij.astype(np.uint32), np.where(kind == 0, d, -d).astype(np.float32)


def pack_restraints(params_list: Sequence, z_list: Sequence[np.ndarray], tables: Optional[Sequence] = None):
    """Ragged restraint tables of a batch for the device embedder: (mol_desc int32 [M, 4] =
    {n_atoms, offset, count, 0}, restr_ij uint32, restr_d float32), from ``restraint_table``
    per molecule (or from ``tables``, those per-molecule pairs made beforehand)."""
    if tables is None:
        tables = [restraint_table(p, z) for p, z in zip(params_list, z_list)]
    desc = np.zeros((len(params_list), 4), np.int32)
    off = 0
    for k, (p, (ij, dd)) in enumerate(zip(params_list, tables)):
        desc[k] = (p.n_atoms, off, len(ij), 0)
        off += len(ij)
    cat = lambda parts, dt: np.concatenate(parts).astype(dt) if len(parts) else np.zeros(0, dt)  # noqa: E731
    return desc, cat([t[0] for t in tables], np.uint32), cat([t[1] for t in tables], np.float32)


def embed_device(params_list: Sequence, z_list: Sequence[np.ndarray], conf_counts: Sequence[int],
                 seeds: Sequence[int], noise: float = 0.1, device: str = "cuda", stream=None,
                 tables: Optional[Sequence] = None):
    """The same seeded restraint embedding as ``embed``, run on the GPU (one workgroup per
    conformer, FP64, csrc/embed.hip). Returns the flat float32 coordinates as a device tensor,
    ready for the minimizer without a host round trip. Results agree with ``embed`` in
    distribution, not bit for bit (device FP64 math and fixed-point gradient sums).
    ``tables``: optional per-molecule ``restraint_table`` results made beforehand."""
    from . import _lib
    torch = _lib.require_gpu()
    L = _lib.load()
    counts = np.asarray(conf_counts, np.int64)
    n_atoms = np.array([p.n_atoms for p in params_list], np.int64)
    conf_mol = np.repeat(np.arange(len(params_list), dtype=np.int32), counts)
    n3 = 3 * n_atoms[conf_mol]
    off = np.zeros(len(conf_mol), np.int64)
    if len(off) > 1:
        off[1:] = np.cumsum(n3)[:-1]
    total = int(n3.sum())
    if total >= 2 ** 31:
        raise ValueError("batch exceeds 2^31 coordinates")
    sd = np.ascontiguousarray(np.asarray(seeds, dtype=np.uint64))
    assert len(sd) == len(conf_mol)
    desc, ij, dd = pack_restraints(params_list, z_list, tables)
    out = torch.empty(total, dtype=torch.float32, device=device)
    if len(conf_mol) == 0:
        return out
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(device)  # noqa: E731
    bufs = [t(desc.reshape(-1)), t(ij.view(np.int32) if len(ij) else np.zeros(1, np.int32)),
            t(dd if len(dd) else np.zeros(1, np.float32)), t(conf_mol), t(off.astype(np.int32)),
            t(sd.view(np.int64))]
    _lib.check(L.mmffhip_embed(*(int(b.data_ptr()) for b in bufs), len(conf_mol), int(n_atoms.max()), float(noise),
                               int(out.data_ptr()), _lib.stream_ptr(stream)))
    # the staging tensors must outlive the asynchronous launch
    out._embed_inputs = bufs
    return out
