"""Seeded benchmark / test workloads for the minimizer.

Each workload is a deterministic function of its name and seed (base 20261202):

=========  ===================  ===========================================================
name       molecules x confs    molecules
=========  ===================  ===========================================================
small      50 x 50              10-25 heavy atoms, <= 1 ring (about 20-50 atoms with H)
medium     1,000 x 10           15-40 heavy atoms, <= 1 ring (about 30-90 atoms)
ragged     10,000 x 20          heavy atoms log-uniform 8-60, <= 3 fused/spiro rings
chain      2,000 x 8            flexible chains of 60-150 heavy atoms, 0-2 macrocycles
=========  ===================  ===========================================================

Start geometries: the package's embedder (seed per conformer) plus N(0, 0.1 A) noise per
coordinate. Molecule and conformer counts can be scaled down for tests.
"""
from __future__ import annotations

import sys
import time
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import molgen
from .embed import embed
from .mmff_params import MMFFParams, extract_mmff_params

SPECS = {
    "small": ("small", 50, 50),
    "medium": ("medium", 1000, 10),
    "ragged": ("ragged", 10000, 20),
    "chain": ("chain", 2000, 8),
}


@dataclass
class Workload:
    name: str
    mols: list
    params: List[MMFFParams]
    conf_counts: List[int]
    x: np.ndarray  # float32 flat start coordinates (sum over conformers of 3N)
    # wall seconds of the build's stages: molecule generation, typing (parameter tables),
    # embedding (start geometries)
    timings: dict = field(default_factory=dict)

    @property
    def n_confs(self) -> int:
        return int(sum(self.conf_counts))

    def atoms_per_conf(self) -> np.ndarray:
        return np.repeat([p.n_atoms for p in self.params], self.conf_counts)


def _typed(args):
    mols, variant = args
    out = []
    for m in mols:
        # the non-bonded 100 A filter of extract_mmff_params needs a conformer; the
        # generated molecules are far smaller than 100 A, so a zero geometry keeps every pair
        m.set_conformers(np.zeros((1, m.GetNumAtoms(), 3)))
        out.append(extract_mmff_params(m, mmff_variant=variant))
    return out


def _gen_typed(args):
    """Worker: generate the given molecules of a family, type them and make their embedder
    restraint tables (each is a function of its own index, so workers need no shared state)."""
    fam, idx, seed, variant, with_tables = args
    mols = [molgen.generate_one(fam, int(m), seed) for m in idx]
    params = _typed((mols, variant))
    tables = None
    if with_tables:
        from .embed import restraint_table
        tables = [restraint_table(p, m.atomic_numbers()) for p, m in zip(params, mols)]
    return mols, params, tables


def _pool_size(n_mols: int) -> int:
    """Worker processes for typing: only for large batches. The workers are started fresh
    (spawned, not forked), so a GPU runtime already initialised in this process, by torch or
    by a profiler, is never copied into them."""
    import os
    if n_mols < 2000:
        return 1
    try:
        cpus = len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        cpus = os.cpu_count() or 1
    return max(1, min(16, cpus))


def build(name: str, n_mols: Optional[int] = None, n_confs: Optional[int] = None,
          seed: int = molgen.SEED_BASE, noise: float = 0.1, variant: str = "MMFF94",
          n_threads: int = 0, embedder: str = "host") -> Workload:
    """``embedder``: "host" (csrc/embed.c, threaded) or "device" (the same method on the GPU,
    csrc/embed.hip; the start geometries then differ from the host's at the level of
    floating-point rounding amplified by the relaxation, not in distribution)."""
    fam, M, K = SPECS[name]
    M = M if n_mols is None else n_mols
    K = K if n_confs is None else n_confs
    t0 = time.perf_counter()
    procs = _pool_size(M)
    tables = None
    if procs > 1:
        # generation, typing and the device embedder's restraint tables in spawned workers
        import multiprocessing as mp
        with mp.get_context("spawn").Pool(procs) as pool:
            parts = pool.map(_gen_typed, [(fam, np.arange(i, M, procs), seed, variant, embedder == "device")
                                          for i in range(procs)])
        mols, params = [None] * M, [None] * M
        tables = [None] * M if embedder == "device" else None
        for i, (pm, pp, pt) in enumerate(parts):
            mols[i::procs] = pm
            params[i::procs] = pp
            if tables is not None:
                tables[i::procs] = pt
        t1 = t2 = time.perf_counter()
        tm = {"generate_type_tables_s": round(t2 - t0, 2), "processes": procs}
    else:
        mols = molgen.generate(fam, M, seed=seed)
        t1 = time.perf_counter()
        params = _typed((mols, variant))
        t2 = time.perf_counter()
        tm = {"generate_s": round(t1 - t0, 2), "type_s": round(t2 - t1, 2)}
    counts = [K] * M
    ids = np.arange(M * K, dtype=np.uint64)
    seeds = (np.uint64(seed) * np.uint64(1000003) + np.uint64(molgen._FAMILY_ID[fam]) * np.uint64(7919) * np.uint64(
        10 ** 9) + ids)
    z = [m.atomic_numbers() for m in mols]
    if embedder == "device":
        from .embed import embed_device
        x = embed_device(params, z, counts, seeds=seeds, noise=noise, tables=tables).cpu().numpy()
        tm["embed_device_s"] = round(time.perf_counter() - t2, 2)
        return Workload(name=name, mols=mols, params=params, conf_counts=counts, x=x, timings=tm)
    if embedder != "host":
        raise ValueError(f"embedder must be 'host' or 'device', not {embedder!r}")
    # large batches are embedded in slices of molecules, with a progress line per slice
    step = M if M * K <= 20000 else max(1, 20000 // K)
    parts = []
    for a in range(0, M, step):
        b = min(M, a + step)
        parts.append(embed(params[a:b], z[a:b], counts[a:b], seeds=seeds[a * K:b * K], noise=noise,
                           n_threads=n_threads))
        if step < M:
            print(f"[workloads] {name}: embedded molecules {b}/{M}", file=sys.stderr, flush=True)
    x = np.concatenate(parts) if len(parts) > 1 else parts[0]
    tm["embed_host_s"] = round(time.perf_counter() - t2, 2)
    return Workload(name=name, mols=mols, params=params, conf_counts=counts, x=x, timings=tm)
