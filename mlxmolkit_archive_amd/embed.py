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
from typing import List, Sequence

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
    mask = np.ones(len(iu), dtype=bool)
    if tgt:
        keys = set(tgt)
        mask = np.array([(a, b) not in keys for a, b in zip(iu.tolist(), ju.tolist())], dtype=bool)
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
