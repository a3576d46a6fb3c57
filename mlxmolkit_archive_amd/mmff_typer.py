"""MMFF94-style atom typing and parameter assignment without RDKit.

The reference delegates typing, charges and every force-field parameter to RDKit's
``MMFFGetMoleculeProperties`` (mlxmolkit/mmff_params.py:122-309). RDKit and the MMFF
parameter files are absent here, so this module assigns MMFF94 symbolic types for the
H/C/N/O/F/S/Cl chemistry the workload generators produce and derives parameters with
MMFF's published functional forms:

* vdW: MMFF buffered 14-7 combination rules (Halgren 1992) from per-type
  (alpha, N, A, G, donor/acceptor) values, with the donor-acceptor scaling
  DARAD = 0.8, DAEPS = 0.5 that RDKit applies (mlxmolkit/mmff_params.py:298-301).
* charges: bond-charge increments, q_i = sum over bonds of +-bci (formal charges zero).
* bonds: tabulated r0/kb for common type pairs, otherwise a Schomaker-Stevenson radius
  rule with a Badger-style kb ~ r0^-6 fallback.
* angles / stretch-bend / out-of-plane / torsions: per-centre and per-bond-class rules.

These values are MMFF94-*style*: the workload only needs realistic magnitudes and term
counts, and parity of the minimizer is measured on whatever tables this module emits.
Agreement with RDKit's exact MMFF94 parameters is NOT claimed (parity unpinned at the
typing boundary); the RDKit-relaxed fixtures in tests/golden measure how far off it is.
When RDKit is importable, mmff_params.extract_mmff_params uses RDKit instead.
"""
from __future__ import annotations

import math
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from .mol import Mol

# --- per-type vdW parameters: alpha, N, A, G, DA ('D' donor, 'A' acceptor, '-' none) ---
VDW = {
    1: (1.050, 2.490, 3.890, 1.282, "-"),
    2: (1.350, 2.490, 3.890, 1.282, "-"),
    3: (1.100, 2.490, 3.890, 1.282, "-"),
    4: (1.300, 2.490, 3.890, 1.282, "-"),
    5: (0.250, 0.800, 4.200, 1.209, "-"),
    6: (0.700, 3.150, 3.890, 1.282, "A"),
    7: (0.650, 3.150, 3.890, 1.282, "A"),
    8: (1.150, 2.820, 3.890, 1.282, "A"),
    9: (0.900, 2.820, 3.890, 1.282, "A"),
    10: (1.000, 2.820, 3.890, 1.282, "-"),
    11: (0.350, 3.480, 3.890, 1.282, "A"),
    12: (2.300, 5.100, 3.320, 1.345, "A"),
    15: (3.000, 4.800, 3.320, 1.345, "A"),
    21: (0.150, 0.800, 4.200, 1.209, "D"),
    23: (0.150, 0.800, 4.200, 1.209, "D"),
    24: (0.150, 0.800, 4.200, 1.209, "D"),
    27: (0.150, 0.800, 4.200, 1.209, "D"),
    28: (0.150, 0.800, 4.200, 1.209, "D"),
    29: (0.150, 0.800, 4.200, 1.209, "D"),
    37: (1.350, 2.490, 3.890, 1.282, "-"),
    38: (0.850, 2.820, 3.890, 1.282, "A"),
    39: (1.100, 2.820, 3.890, 1.282, "-"),
    40: (1.000, 2.820, 3.890, 1.282, "-"),
    42: (1.000, 2.820, 3.890, 1.282, "A"),
    44: (3.000, 4.800, 3.320, 1.345, "-"),
    59: (0.650, 3.150, 3.890, 1.282, "-"),
    63: (1.350, 2.490, 3.890, 1.282, "-"),
    64: (1.350, 2.490, 3.890, 1.282, "-"),
    71: (0.250, 0.800, 4.200, 1.209, "-"),
}

# hybridization class per type: 3 = sp3, 2 = sp2/aromatic, 1 = sp
HYB = {1: 3, 2: 2, 3: 2, 4: 1, 6: 3, 7: 2, 8: 3, 9: 2, 10: 2, 11: 3, 12: 3, 15: 3, 37: 2, 38: 2,
       39: 2, 40: 2, 42: 1, 44: 2, 59: 2, 63: 2, 64: 2}
AROMATIC_TYPES = {37, 38, 39, 44, 59, 63, 64}

# bond-charge increments bci(I, J) for I < J: q_I += bci, q_J -= bci
BCI = {
    (1, 1): 0.0, (1, 2): 0.0, (1, 3): 0.0, (1, 4): 0.0, (1, 5): 0.0, (1, 6): 0.28, (1, 8): 0.27,
    (1, 9): 0.32, (1, 10): 0.27, (1, 11): 0.23, (1, 12): 0.18, (1, 15): 0.23, (1, 37): 0.0,
    (1, 38): 0.33, (1, 40): 0.27, (1, 63): 0.0, (1, 64): 0.0,
    (2, 2): 0.0, (2, 3): 0.0, (2, 5): -0.15, (2, 6): 0.29, (2, 8): 0.30, (2, 9): 0.35, (2, 11): 0.23,
    (2, 12): 0.14, (2, 15): 0.15, (2, 37): 0.0, (2, 40): 0.20,
    (3, 3): 0.0, (3, 5): -0.06, (3, 6): 0.15, (3, 7): 0.57, (3, 9): 0.38, (3, 10): 0.0,
    (3, 11): 0.23, (3, 12): 0.18, (3, 15): 0.15, (3, 37): 0.0, (3, 40): 0.10,
    (4, 5): -0.15, (4, 4): 0.0, (4, 37): 0.0, (4, 42): 0.43,
    (5, 37): 0.15, (5, 63): 0.15, (5, 64): 0.15,
    (6, 21): -0.40, (6, 24): -0.45, (6, 29): -0.45, (6, 37): -0.10, (6, 3): -0.15,
    (8, 23): -0.36, (9, 27): -0.40, (10, 28): -0.37, (15, 15): 0.0, (15, 37): 0.0, (15, 71): -0.18,
    (11, 37): -0.18, (12, 37): -0.10, (37, 37): 0.0, (37, 38): 0.16, (37, 40): 0.10,
    (38, 38): 0.0, (28, 40): 0.40, (63, 63): 0.0, (63, 64): 0.0, (64, 64): 0.0, (44, 63): -0.15,
    (37, 63): 0.0, (37, 64): 0.0, (39, 63): 0.22, (59, 63): -0.27, (23, 39): 0.0,
}
# partial bci fallback: bci(I,J) = PBCI[I] - PBCI[J]
PBCI = {1: 0.0, 2: -0.135, 3: -0.065, 4: -0.2, 5: 0.0, 6: -0.28, 7: -0.57, 8: -0.27, 9: -0.23,
        10: -0.27, 11: -0.23, 12: -0.18, 15: -0.20, 21: 0.12, 23: 0.09, 24: 0.17, 27: 0.17,
        28: 0.10, 29: 0.17, 37: -0.135, 38: -0.28, 39: -0.15, 40: -0.27, 42: -0.55, 44: -0.22,
        59: -0.20, 63: -0.135, 64: -0.135, 71: -0.02}

# tabulated (type-class pair, bond order) -> (kb, r0)
_BOND_TABLE = {
    ("C3", "C3", 1): (4.258, 1.508), ("C3", "C2", 1): (4.539, 1.482), ("C3", "CA", 1): (4.295, 1.499),
    ("C3", "C1", 1): (5.0, 1.465), ("C2", "C2", 2): (9.505, 1.333), ("C2", "C2", 1): (5.3, 1.460),
    ("CA", "CA", 1.5): (5.573, 1.374), ("CA", "C2", 1): (5.0, 1.475), ("C1", "C1", 3): (15.5, 1.200),
    ("C3", "H", 1): (4.766, 1.093), ("C2", "H", 1): (5.170, 1.083), ("CA", "H", 1): (5.306, 1.084),
    ("C1", "H", 1): (5.9, 1.065), ("C2", "O2", 2): (12.95, 1.222), ("C3", "O3", 1): (5.047, 1.418),
    ("C2", "O3", 1): (5.7, 1.355), ("CA", "O3", 1): (5.8, 1.373), ("O3", "H", 1): (7.811, 0.972),
    ("C3", "N3", 1): (5.084, 1.451), ("C2", "N3", 1): (6.0, 1.369), ("CA", "N3", 1): (6.1, 1.388),
    ("N3", "H", 1): (6.420, 1.015), ("C2", "N2", 2): (9.9, 1.279), ("C3", "N2", 1): (5.0, 1.462),
    ("CA", "NA", 1.5): (5.8, 1.342), ("C1", "N1", 3): (17.5, 1.154), ("N2", "H", 1): (6.3, 1.021),
    ("C3", "F", 1): (6.011, 1.360), ("C2", "F", 1): (6.4, 1.340), ("CA", "F", 1): (6.4, 1.350),
    ("C3", "Cl", 1): (2.974, 1.773), ("C2", "Cl", 1): (3.1, 1.730), ("CA", "Cl", 1): (3.1, 1.741),
    ("C3", "S3", 1): (2.893, 1.805), ("C2", "S3", 1): (3.0, 1.760), ("CA", "S3", 1): (3.3, 1.770),
    ("S3", "H", 1): (4.0, 1.340), ("S3", "S3", 1): (2.6, 2.040), ("CA", "SA", 1.5): (3.6, 1.720),
    ("C1", "C2", 1): (5.5, 1.430), ("C1", "CA", 1): (5.5, 1.433), ("N3", "N3", 1): (5.0, 1.43),
    ("C3", "O2", 2): (12.95, 1.222), ("O3", "O3", 1): (4.0, 1.45), ("CA", "OA", 1.5): (5.0, 1.360),
    ("NA", "H", 1): (6.6, 1.010), ("H", "Cl", 1): (4.6, 1.280),
}

_COV = {1: 0.33, 6: 0.77, 7: 0.73, 8: 0.72, 9: 0.74, 16: 1.03, 17: 1.00}
_CHI = {1: 2.20, 6: 2.50, 7: 3.07, 8: 3.50, 9: 4.10, 16: 2.44, 17: 2.83}


def _cls(t: int, z: int) -> str:
    if z == 1:
        return "H"
    if z == 6:
        return {1: "C3", 2: "C2", 3: "C2", 4: "C1", 37: "CA", 63: "CA", 64: "CA"}.get(t, "C3")
    if z == 7:
        return {8: "N3", 10: "N3", 40: "N3", 9: "N2", 38: "NA", 39: "NA", 42: "N1"}.get(t, "N3")
    if z == 8:
        return {7: "O2", 59: "OA"}.get(t, "O3")
    if z == 16:
        return "SA" if t == 44 else "S3"
    if z == 9:
        return "F"
    if z == 17:
        return "Cl"
    return "X"


def _row(z: int) -> int:
    return 0 if z == 1 else (1 if z <= 10 else 2)


def find_small_rings(mol: Mol, max_size: int = 8) -> List[List[int]]:
    """Simple cycles up to ``max_size`` atoms (each listed once, canonical order)."""
    n = mol.GetNumAtoms()
    rings = set()
    out = []
    for s in range(n):
        stack = [(s, [s])]
        while stack:
            u, path = stack.pop()
            for v in mol.neighbors(u):
                if v == s and len(path) >= 3:
                    key = tuple(sorted(path))
                    if key not in rings:
                        rings.add(key)
                        out.append(list(path))
                elif v > s and v not in path and len(path) < max_size:
                    stack.append((v, path + [v]))
    return out


class MMFFTyper:
    """Assigns MMFF94 numeric atom types and MMFF94-style parameters for one molecule."""

    def __init__(self, mol: Mol, variant: str = "MMFF94"):
        if variant not in ("MMFF94", "MMFF94s"):
            raise ValueError(f"unknown MMFF variant {variant!r}")
        self.mol = mol
        self.variant = variant
        self.z = mol.atomic_numbers()
        n = mol.GetNumAtoms()
        for a in mol.atoms:
            if a.z not in (1, 6, 7, 8, 9, 16, 17):
                raise ValueError(f"MMFF typing not available for element Z={a.z}")
        self.rings = find_small_rings(mol)
        self.ring_size = np.zeros(n, dtype=np.int32)
        for r in self.rings:
            for a in r:
                if self.ring_size[a] == 0 or len(r) < self.ring_size[a]:
                    self.ring_size[a] = len(r)
        self.arom = self._perceive_aromatic()
        # bonds inside a perceived aromatic ring count as order 1.5 whatever the input
        # (Kekule) order was
        self._arom_bond = set()
        for r in self.rings:
            if len(r) in (5, 6) and all(self.arom[a] for a in r):
                for k in range(len(r)):
                    a, b = r[k], r[(k + 1) % len(r)]
                    self._arom_bond.add((min(a, b), max(a, b)))
        self.types = self._assign_types()
        self.charges = self._charges()

    # -- perception --------------------------------------------------------
    def _perceive_aromatic(self) -> np.ndarray:
        mol = self.mol
        n = mol.GetNumAtoms()
        arom = np.array([a.aromatic for a in mol.atoms], dtype=bool)
        for r in self.rings:
            if len(r) not in (5, 6):
                continue
            orders = [mol.bond_order(r[k], r[(k + 1) % len(r)]) for k in range(len(r))]
            if len(r) == 6:
                ndbl = sum(1 for o in orders if o == 2.0)
                if all(o in (1.5,) for o in orders) or (ndbl == 3 and all(o in (1.0, 2.0) for o in orders)):
                    arom[r] = True
            else:
                ndbl = sum(1 for o in orders if o == 2.0)
                if all(o == 1.5 for o in orders) or ndbl == 2:
                    # need one heteroatom donating a lone pair
                    if any(self.z[a] in (7, 8, 16) for a in r):
                        arom[r] = True
        return arom

    def _multiple(self, i: int) -> Tuple[int, int, List[int]]:
        dbl, trip, dpart = 0, 0, []
        for j in self.mol.neighbors(i):
            o = self.mol.bond_order(i, j)
            if o == 2.0:
                dbl += 1
                dpart.append(j)
            elif o == 3.0:
                trip += 1
        return dbl, trip, dpart

    def _assign_types(self) -> List[int]:
        mol, z = self.mol, self.z
        n = mol.GetNumAtoms()
        t = [0] * n
        # heavy atoms first
        for i in range(n):
            zi = z[i]
            dbl, trip, dpart = self._multiple(i)
            nb = mol.neighbors(i)
            if zi == 6:
                if self.arom[i]:
                    rs = self.ring_size[i]
                    if rs == 5:
                        het = [j for j in nb if self.arom[j] and z[j] in (7, 8, 16)]
                        t[i] = 63 if het else 64
                    else:
                        t[i] = 37
                elif trip or dbl >= 2:
                    t[i] = 4
                elif dbl == 1:
                    t[i] = 3 if z[dpart[0]] in (8, 7, 16) else 2
                else:
                    t[i] = 1
            elif zi == 7:
                if self.arom[i]:
                    t[i] = 38 if len(nb) == 2 and self.ring_size[i] == 6 else (39 if len(nb) == 3 or any(
                        z[j] == 1 for j in nb) else 38)
                elif trip:
                    t[i] = 42
                elif dbl:
                    t[i] = 9
                else:
                    t[i] = 8
            elif zi == 8:
                if self.arom[i]:
                    t[i] = 59
                else:
                    t[i] = 7 if dbl else 6
            elif zi == 16:
                t[i] = 44 if self.arom[i] else 15
            elif zi == 9:
                t[i] = 11
            elif zi == 17:
                t[i] = 12
        # second pass: N environment (amide / enamine-aniline)
        for i in range(n):
            if t[i] == 8:
                for j in mol.neighbors(i):
                    if t[j] == 3 and any(z[k] == 8 and mol.bond_order(j, k) == 2.0 for k in mol.neighbors(j)):
                        t[i] = 10
                        break
                else:
                    for j in mol.neighbors(i):
                        if t[j] in (2, 37, 3, 63, 64):
                            t[i] = 40
                            break
        # hydrogens
        for i in range(n):
            if z[i] != 1:
                continue
            nb = mol.neighbors(i)
            if not nb:
                t[i] = 5
                continue
            p = nb[0]
            tp = t[p]
            if z[p] == 6:
                t[i] = 5
            elif z[p] == 8:
                acid = any(t[q] == 3 for q in mol.neighbors(p) if q != i)
                enol = any(t[q] in (2, 37) for q in mol.neighbors(p) if q != i)
                t[i] = 24 if acid else (29 if enol else 21)
            elif z[p] == 7:
                t[i] = {8: 23, 10: 28, 40: 28, 9: 27, 39: 23}.get(tp, 23)
            elif z[p] == 16:
                t[i] = 71
            else:
                t[i] = 5
        return t

    # -- charges -----------------------------------------------------------
    def bci(self, ti: int, tj: int) -> float:
        if ti == tj:
            return 0.0
        if ti < tj:
            key, sgn = (ti, tj), 1.0
        else:
            key, sgn = (tj, ti), -1.0
        if key in BCI:
            return sgn * BCI[key]
        return sgn * (PBCI.get(key[0], 0.0) - PBCI.get(key[1], 0.0))

    def _charges(self) -> np.ndarray:
        q = np.zeros(self.mol.GetNumAtoms())
        for b in self.mol.bonds:
            i, j = b.begin, b.end
            w = self.bci(self.types[i], self.types[j])
            q[i] += w
            q[j] -= w
        return q

    # -- helpers -----------------------------------------------------------
    def bond_order(self, i: int, j: int) -> float:
        if (min(i, j), max(i, j)) in self._arom_bond:
            return 1.5
        return self.mol.bond_order(i, j)

    def is_linear(self, i: int) -> bool:
        return HYB.get(self.types[i], 3) == 1

    def hyb(self, i: int) -> int:
        if self.z[i] == 1:
            return 0
        return HYB.get(self.types[i], 3)

    # -- parameters --------------------------------------------------------
    def bond(self, i: int, j: int) -> Tuple[float, float]:
        o = self.bond_order(i, j)
        ci, cj = _cls(self.types[i], self.z[i]), _cls(self.types[j], self.z[j])
        key_o = 1.5 if o == 1.5 else int(o)
        for key in ((ci, cj, key_o), (cj, ci, key_o)):
            if key in _BOND_TABLE:
                return _BOND_TABLE[key]
        # Schomaker-Stevenson radius rule + Badger-style stiffness
        ri, rj = _COV[self.z[i]], _COV[self.z[j]]
        shrink = {1: 0.0, 1.5: 0.07, 2: 0.10, 3: 0.17}[key_o]
        r0 = ri + rj - 2 * shrink - 0.085 * abs(_CHI[self.z[i]] - _CHI[self.z[j]]) ** 1.4
        kb = 4.3 * (1.51 / r0) ** 6 * {1: 1.0, 1.5: 1.3, 2: 2.2, 3: 3.6}[key_o]
        return kb, r0

    def angle(self, i: int, j: int, k: int) -> Tuple[float, float]:
        tj, zj = self.types[j], self.z[j]
        nH = int(self.z[i] == 1) + int(self.z[k] == 1)
        h = self.hyb(j)
        in_ring = self._same_small_ring(i, j, k)
        if h == 1:
            return 0.25, 180.0
        if zj == 6:
            if h == 3:
                t0 = {0: 109.6, 1: 110.5, 2: 108.8}[nH]
                ka = {0: 0.851, 1: 0.636, 2: 0.516}[nH]
            else:
                if tj == 3 and (self.types[i] == 7 or self.types[k] == 7):
                    t0, ka = 123.0, 0.72 if nH == 0 else 0.55
                elif tj == 3:
                    t0, ka = 114.0, 0.80
                else:
                    t0 = 120.0
                    ka = {0: 0.669, 1: 0.519, 2: 0.410}[nH]
        elif zj == 7:
            if h == 3:
                t0, ka = (106.4, 0.48) if nH == 2 else ((108.0, 0.60) if nH == 1 else (109.0, 0.85))
            elif tj == 9:
                t0, ka = 118.0, 0.75
            elif tj in (38,):
                t0, ka = 117.0, 0.80
            else:
                t0, ka = (118.5, 0.40) if nH == 2 else (120.0, 0.60 if nH else 0.80)
        elif zj == 8:
            t0, ka = (108.0, 0.72) if nH == 1 else ((104.5, 0.66) if nH == 2 else (106.5, 1.00))
        elif zj == 16:
            t0, ka = (96.5, 0.60) if nH else (99.0, 0.90)
            if tj == 44:
                t0, ka = 92.0, 1.00
        else:
            t0, ka = 109.5, 0.6
        if in_ring == 5:
            t0 = 108.0 if h == 2 else 104.5
        elif in_ring == 6 and h == 2:
            t0 = 120.0
        return ka, t0

    def _same_small_ring(self, i: int, j: int, k: int) -> int:
        best = 0
        for r in self.rings:
            if i in r and j in r and k in r and len(r) <= 6:
                if best == 0 or len(r) < best:
                    best = len(r)
        return best

    def strbend(self, i: int, j: int, k: int) -> Tuple[float, float]:
        ri, rj, rk = _row(self.z[i]), _row(self.z[j]), _row(self.z[k])
        table = {
            (0, 1, 0): (0.15, 0.15), (0, 1, 1): (0.10, 0.30), (1, 1, 0): (0.30, 0.10),
            (1, 1, 1): (0.30, 0.30), (0, 1, 2): (0.05, 0.35), (2, 1, 0): (0.35, 0.05),
            (1, 1, 2): (0.30, 0.45), (2, 1, 1): (0.45, 0.30), (2, 1, 2): (0.45, 0.45),
            (0, 2, 0): (0.00, 0.00), (0, 2, 1): (0.00, 0.15), (1, 2, 0): (0.15, 0.00),
            (1, 2, 1): (0.15, 0.15), (1, 2, 2): (0.15, 0.25), (2, 2, 1): (0.25, 0.15), (2, 2, 2): (0.25, 0.25),
            (0, 2, 2): (0.0, 0.25), (2, 2, 0): (0.25, 0.0),
        }
        return table.get((ri, rj, rk), (0.2, 0.2))

    def oop(self, j: int) -> Optional[float]:
        tj = self.types[j]
        if self.hyb(j) != 2:
            return None
        if tj == 3:
            return 0.130
        if tj in (37, 63, 64):
            return 0.040
        if tj == 2:
            return 0.030
        if tj == 10:
            return 0.080 if self.variant == "MMFF94s" else 0.015
        if tj == 40:
            return 0.090 if self.variant == "MMFF94s" else 0.030
        if tj in (39,):
            return 0.050
        return 0.030

    def torsion(self, i: int, j: int, k: int, l: int) -> Optional[Tuple[float, float, float]]:
        if self.is_linear(j) or self.is_linear(k):
            return None
        hj, hk = self.hyb(j), self.hyb(k)
        o = self.bond_order(j, k)
        tj, tk = self.types[j], self.types[k]
        heavy_ends = int(self.z[i] != 1) + int(self.z[l] != 1)
        soft = self.variant == "MMFF94s"
        if o == 2.0:
            return (0.0, 12.0, 0.0)
        if o == 1.5 or (tj in AROMATIC_TYPES and tk in AROMATIC_TYPES and self._same_ring(j, k)):
            return (0.0, 6.0, 0.0)
        if hj == 3 and hk == 3:
            zj, zk = self.z[j], self.z[k]
            v3 = 0.30
            if 8 in (zj, zk):
                v3 = 0.45
            elif 16 in (zj, zk):
                v3 = 0.40
            elif 7 in (zj, zk):
                v3 = 0.35
            if heavy_ends == 2:
                return (0.103, 0.681, 0.332 if v3 == 0.30 else v3)
            if heavy_ends == 1:
                return (0.0, 0.0, 0.267 if v3 == 0.30 else v3 * 0.9)
            return (0.0, 0.0, v3)
        if hj == 2 and hk == 2:
            amide = {tj, tk} in ({3, 10},) or (tj == 3 and tk == 10) or (tj == 10 and tk == 3)
            if amide:
                return (0.0, 4.5 if soft else 5.5, 0.0)
            if 40 in (tj, tk):
                return (0.0, 1.6 if soft else 2.0, 0.0)
            return (0.3, 1.8, 0.0)
        # sp2 - sp3 single bond
        if 40 in (tj, tk) or 10 in (tj, tk):
            return (0.0, 0.0, 0.05 if soft else 0.10)
        if heavy_ends == 2:
            return (0.20, 0.10, 0.15)
        return (0.0, 0.0, 0.05)

    def _same_ring(self, j: int, k: int) -> bool:
        return any(j in r and k in r for r in self.rings)

    def vdw(self, i: int, j: int) -> Tuple[float, float, float, float]:
        """(R*, eps, R*_DA_scaled, eps_DA_scaled) for the pair, MMFF combination rules."""
        ai, Ni, Ai, Gi, DAi = VDW[self.types[i]]
        aj, Nj, Aj, Gj, DAj = VDW[self.types[j]]
        Rii = Ai * ai ** 0.25
        Rjj = Aj * aj ** 0.25
        gam = (Rii - Rjj) / (Rii + Rjj)
        B = 0.0 if (DAi == "D" or DAj == "D") else 0.2
        Rij = 0.5 * (Rii + Rjj) * (1.0 + B * (1.0 - math.exp(-12.0 * gam * gam)))
        eps = 181.16 * Gi * Gj * ai * aj / (math.sqrt(ai / Ni) + math.sqrt(aj / Nj)) * Rij ** -6
        Rs, es = Rij, eps
        if (DAi == "D" and DAj == "A") or (DAi == "A" and DAj == "D"):
            Rs, es = Rij * 0.8, eps * 0.5
        return Rij, eps, Rs, es
