"""Seeded synthetic molecule generators for the minimizer workloads.

Four families, one per workload size:

* ``small``   10-25 heavy atoms (about 20-50 atoms with H), <= 1 five/six ring
* ``medium``  15-40 heavy atoms (about 30-90 atoms), <= 1 ring
* ``ragged``  heavy count log-uniform in 8-60 (10-150 atoms), <= 3 fused or spiro rings
* ``chain``   flexible chains of 60-150 heavy atoms (150-400 atoms), 0-2 macrocycles of
              12-24 members, amide and ether linkers

Heavy atoms are drawn C 70 % / N 12 % / O 12 % / {F, Cl, S} 6 %, every valence is
satisfied and hydrogens are explicit. Unsaturation: carbonyls, C=C, nitriles and
aromatic six-rings (benzene / pyridine). Everything is a deterministic function of the
seed (numpy PCG64).
"""
from __future__ import annotations

import math
from typing import List, Optional, Tuple

import numpy as np

from .mol import Mol

MAXV = {6: 4, 7: 3, 8: 2, 16: 2, 9: 1, 17: 1}
SEED_BASE = 20261202


def _draw_element(rng) -> int:
    u = rng.random()
    if u < 0.70:
        return 6
    if u < 0.82:
        return 7
    if u < 0.94:
        return 8
    return int(rng.choice([9, 17, 16]))


class _Builder:
    def __init__(self, rng):
        self.rng = rng
        self.z: List[int] = []
        self.arom: List[bool] = []
        self.bonds: dict = {}  # (i,j) i<j -> order
        self.used: List[float] = []

    def add_atom(self, z: int, arom: bool = False) -> int:
        self.z.append(z)
        self.arom.append(arom)
        self.used.append(0.0)
        return len(self.z) - 1

    def free(self, i: int) -> float:
        return MAXV[self.z[i]] - self.used[i]

    def add_bond(self, i: int, j: int, order: float) -> None:
        key = (min(i, j), max(i, j))
        assert key not in self.bonds
        self.bonds[key] = order
        self.used[i] += order
        self.used[j] += order

    def has_bond(self, i, j):
        return (min(i, j), max(i, j)) in self.bonds

    # rings ---------------------------------------------------------------
    def ring(self, size: int, aromatic: bool, hetero: bool, fuse: Optional[Tuple[int, int]] = None,
             spiro: Optional[int] = None) -> List[int]:
        rng = self.rng
        atoms: List[int] = []
        if fuse is not None:
            atoms = [fuse[0], fuse[1]]
        elif spiro is not None:
            atoms = [spiro]
        n_new = size - len(atoms)
        het_pos = int(rng.integers(0, n_new)) if hetero else -1
        for k in range(n_new):
            z = 6
            if k == het_pos:
                z = 7 if aromatic else int(rng.choice([7, 8, 16]))
            atoms.append(self.add_atom(z, aromatic))
        if fuse is not None:
            # new ring path: fuse[1] -> new atoms -> fuse[0]
            path = [fuse[1]] + atoms[2:] + [fuse[0]]
            order = 1.5 if aromatic else 1.0
            for a, b in zip(path[:-1], path[1:]):
                self.add_bond(a, b, order)
            if aromatic:
                for a in fuse:
                    self.arom[a] = True
            return atoms
        order = 1.5 if aromatic else 1.0
        for k in range(size):
            a, b = atoms[k], atoms[(k + 1) % size]
            self.add_bond(a, b, order)
        return atoms

    def _aromatic_used(self, i: int) -> float:
        # an aromatic atom with two aromatic bonds uses 3 valence units (C) or all 3 (N)
        return self.used[i]

    def attach_tree(self, n_more: int) -> None:
        rng = self.rng
        for _ in range(n_more):
            z = _draw_element(rng)
            cand = [i for i in range(len(self.z)) if self._can_take(i, z)]
            if not cand:
                z = 6
                cand = [i for i in range(len(self.z)) if self._can_take(i, z)]
                if not cand:
                    return
            # prefer growing from recently added / carbon atoms for chain-like graphs
            w = np.array([1.0 + (2.0 if self.z[i] == 6 else 0.0) for i in cand])
            parent = int(rng.choice(cand, p=w / w.sum()))
            c = self.add_atom(z)
            self.add_bond(parent, c, 1.0)

    def _can_take(self, i: int, z: int) -> bool:
        if self.free(i) < 1.0 - 1e-9:
            return False
        if self.z[i] in (9, 17):
            return False
        if z in (9, 17) and self.z[i] != 6:
            return False
        if z in (8, 16, 7) and self.z[i] in (8, 16, 7) and not (self.z[i] == 16 and z == 16):
            # avoid N-N, N-O, O-O, O-S heteroatom chains (peroxides, hydrazines)
            return False
        return True

    def unsaturate(self, p_carbonyl=0.5, p_cc=0.15, p_nitrile=0.08) -> None:
        rng = self.rng
        bonds = list(self.bonds.items())
        rng.shuffle(bonds)
        for (i, j), o in bonds:
            if o != 1.0 or self.arom[i] or self.arom[j]:
                continue
            zi, zj = self.z[i], self.z[j]
            deg_i = sum(1 for k in self.bonds if i in k)
            deg_j = sum(1 for k in self.bonds if j in k)
            # carbonyl: terminal O on a carbon with spare valence
            if {zi, zj} == {6, 8}:
                c, ox = (i, j) if zi == 6 else (j, i)
                dox = deg_j if ox == j else deg_i
                if dox == 1 and self.free(c) >= 1 and not self._has_double(c) and rng.random() < p_carbonyl:
                    self._upgrade(i, j, 2.0)
            elif zi == 6 and zj == 6:
                if self.free(i) >= 1 and self.free(j) >= 1 and not self._has_double(i) and not self._has_double(j) \
                        and not self._in_ring_bond(i, j) and rng.random() < p_cc:
                    self._upgrade(i, j, 2.0)
            elif {zi, zj} == {6, 7}:
                c, nn = (i, j) if zi == 6 else (j, i)
                dn = deg_j if nn == j else deg_i
                dc = deg_i if c == i else deg_j
                if dn == 1 and dc == 2 and self.free(c) >= 2 and self.free(nn) >= 2 and not self._has_double(c) \
                        and rng.random() < p_nitrile:
                    self._upgrade(i, j, 3.0)

    def _has_double(self, i):
        return any(o >= 2.0 for (a, b), o in self.bonds.items() if i in (a, b))

    def _in_ring_bond(self, i, j):
        return False

    def _upgrade(self, i, j, order):
        key = (min(i, j), max(i, j))
        old = self.bonds[key]
        self.bonds[key] = order
        self.used[i] += order - old
        self.used[j] += order - old

    def finish(self, name: str = "") -> Mol:
        heavy = len(self.z)
        z = list(self.z)
        arom = list(self.arom)
        bonds = [(a, b, o) for (a, b), o in sorted(self.bonds.items())]
        for i in range(heavy):
            if self.z[i] == 7 and arom[i]:
                continue  # pyridine-type N carries no H
            used = self.used[i]
            if arom[i] and self.z[i] == 6:
                n_ar = sum(1 for (a, b), o in self.bonds.items() if i in (a, b) and o == 1.5)
                other = used - 1.5 * n_ar
                used = other + (3.0 if n_ar == 2 else 4.0 if n_ar == 3 else 1.5 * n_ar)
            nh = int(round(MAXV[self.z[i]] - used))
            for _ in range(max(nh, 0)):
                h = len(z)
                z.append(1)
                arom.append(False)
                bonds.append((i, h, 1.0))
        return Mol(z, bonds, arom, name=name)


def gen_molecule(rng, n_heavy: int, max_rings: int = 1, allow_fused: bool = False,
                 macrocycles: int = 0, chain_linkers: bool = False, name: str = "") -> Mol:
    b = _Builder(rng)
    n_rings = int(rng.integers(0, max_rings + 1)) if max_rings > 0 else 0
    ring_atoms: List[List[int]] = []
    budget = n_heavy
    if chain_linkers:
        _chain_backbone(b, rng, n_heavy, macrocycles)
        budget = n_heavy - len(b.z)
    for r in range(n_rings):
        size = int(rng.choice([5, 6]))
        if budget < size:
            break
        aromatic = size == 6 and rng.random() < 0.6
        hetero = rng.random() < 0.35
        if r == 0 or not ring_atoms:
            if b.z:
                # attach the ring to the existing graph through a single bond
                atoms = b.ring(size, aromatic, hetero)
                cand = [i for i in range(len(b.z) - size) if b._can_take(i, 6)]
                ra = [a for a in atoms if b.free(a) >= 1 and b.z[a] == 6]
                if cand and ra:
                    b.add_bond(int(rng.choice(cand)), int(rng.choice(ra)), 1.0)
            else:
                atoms = b.ring(size, aromatic, hetero)
            budget -= size
        else:
            prev = ring_atoms[int(rng.integers(0, len(ring_atoms)))]
            mode = rng.random() if allow_fused else 1.0
            if mode < 0.45:
                # fused: share a bond of the previous ring whose atoms have spare valence
                pairs = [(prev[k], prev[(k + 1) % len(prev)]) for k in range(len(prev))]
                pairs = [(p, q) for p, q in pairs if b.free(p) >= 1 and b.free(q) >= 1 and b.z[p] == 6
                         and b.z[q] == 6]
                if not pairs:
                    continue
                p, q = pairs[int(rng.integers(0, len(pairs)))]
                both_arom = b.arom[p] and b.arom[q]
                aro = aromatic and both_arom
                if both_arom and not aro:
                    size = 6 if size == 6 else 5
                atoms = b.ring(size, aro, hetero and not aro, fuse=(p, q))
                budget -= size - 2
            elif mode < 0.7:
                cand = [a for a in prev if b.z[a] == 6 and not b.arom[a] and b.free(a) >= 2]
                if not cand:
                    continue
                atoms = b.ring(size, False, hetero, spiro=int(rng.choice(cand)))
                budget -= size - 1
            else:
                atoms = b.ring(size, aromatic, hetero)
                cand = [i for i in range(len(b.z) - size) if b._can_take(i, 6)]
                ra = [a for a in atoms if b.free(a) >= 1 and b.z[a] == 6]
                if cand and ra:
                    b.add_bond(int(rng.choice(cand)), int(rng.choice(ra)), 1.0)
                budget -= size
        ring_atoms.append(atoms)
    if not b.z:
        b.add_atom(6)
        budget -= 1
    b.attach_tree(max(budget, 0))
    b.unsaturate()
    return b.finish(name)


def _chain_backbone(b: _Builder, rng, n_heavy: int, macrocycles: int) -> None:
    """Flexible backbone: alkyl segments joined by amide (C(=O)N) and ether (O) linkers."""
    target = int(n_heavy * 0.8)
    prev = b.add_atom(6)
    backbone = [prev]
    while len(b.z) < target:
        u = rng.random()
        if u < 0.12:  # amide linker -C(=O)-NH-
            c = b.add_atom(6)
            b.add_bond(prev, c, 1.0)
            o = b.add_atom(8)
            b.add_bond(c, o, 2.0)
            n = b.add_atom(7)
            b.add_bond(c, n, 1.0)
            backbone += [c, n]
            prev = n
        elif u < 0.22:  # ether
            o = b.add_atom(8)
            b.add_bond(prev, o, 1.0)
            c = b.add_atom(6)
            b.add_bond(o, c, 1.0)
            backbone += [o, c]
            prev = c
        else:
            c = b.add_atom(6)
            b.add_bond(prev, c, 1.0)
            backbone.append(c)
            prev = c
    n_macro = int(rng.integers(0, macrocycles + 1))
    used_pos = set()
    for _ in range(n_macro):
        L = int(rng.integers(12, 25))
        for _try in range(20):
            s = int(rng.integers(0, max(1, len(backbone) - L)))
            e = s + L - 1
            if e >= len(backbone):
                continue
            a, c = backbone[s], backbone[e]
            if a in used_pos or c in used_pos or b.z[a] != 6 or b.z[c] != 6:
                continue
            if b.free(a) >= 1 and b.free(c) >= 1 and not b.has_bond(a, c):
                b.add_bond(a, c, 1.0)
                used_pos.update((a, c))
                break


def generate_one(family: str, m: int, seed: int = SEED_BASE) -> Mol:
    """Molecule ``m`` of a workload family: a function of (seed, family, m) alone, so any
    subset of a family's molecules can be generated independently (in worker processes)."""
    rng = np.random.default_rng([seed, _FAMILY_ID[family], m])
    if family == "small":
        nh = int(rng.integers(10, 26))
        return gen_molecule(rng, nh, max_rings=1, name=f"small-{m}")
    if family == "medium":
        nh = int(rng.integers(15, 41))
        return gen_molecule(rng, nh, max_rings=1, name=f"medium-{m}")
    if family == "ragged":
        nh = int(round(math.exp(rng.uniform(math.log(8), math.log(60)))))
        return gen_molecule(rng, nh, max_rings=3, allow_fused=True, name=f"ragged-{m}")
    if family == "chain":
        nh = int(rng.integers(60, 151))
        return gen_molecule(rng, nh, max_rings=1, macrocycles=2, chain_linkers=True, name=f"chain-{m}")
    raise ValueError(family)


def generate(family: str, n_mols: int, seed: int = SEED_BASE) -> List[Mol]:
    """Deterministic list of molecules of one workload family."""
    if family not in _FAMILY_ID:
        raise ValueError(family)
    return [generate_one(family, m, seed) for m in range(n_mols)]


_FAMILY_ID = {"small": 1, "medium": 2, "ragged": 3, "chain": 4}
