"""Minimal molecule model with the RDKit surface the MMFF drop-in API touches.

The reference API takes RDKit ``Chem.Mol`` objects (mlxmolkit/mmff_metal_optimizer.py:217-
535 reads ``GetConformers``, ``GetConformer(cid).GetAtomPosition`` and writes back with
``SetAtomPosition``). RDKit is optional: when it is importable those objects work
unchanged; otherwise this module supplies a ``Mol`` with the same methods, built from a
small SMILES parser (organic subset, aromatic lowercase atoms, branches, ring closures,
bracket atoms with explicit H) or from a generated graph.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Iterable, List, Optional, Sequence, Tuple

import numpy as np

ELEMENTS = {"H": 1, "C": 6, "N": 7, "O": 8, "F": 9, "S": 16, "Cl": 17, "Br": 35, "I": 53, "P": 15}
SYMBOL = {v: k for k, v in ELEMENTS.items()}
DEFAULT_VALENCE = {1: 1, 6: 4, 7: 3, 8: 2, 9: 1, 16: 2, 17: 1, 35: 1, 53: 1, 15: 3}


class Point3D:
    """Coordinate triple with the ``x/y/z`` attributes of RDKit's Point3D."""

    __slots__ = ("x", "y", "z")

    def __init__(self, x: float, y: float, z: float):
        self.x, self.y, self.z = float(x), float(y), float(z)

    def Distance(self, other: "Point3D") -> float:
        return float(np.sqrt((self.x - other.x) ** 2 + (self.y - other.y) ** 2 + (self.z - other.z) ** 2))

    def __iter__(self):
        return iter((self.x, self.y, self.z))


class Conformer:
    """One set of coordinates (RDKit ``Conformer`` subset)."""

    def __init__(self, n_atoms: int, positions: Optional[np.ndarray] = None, conf_id: int = 0):
        self._pos = np.zeros((n_atoms, 3), dtype=np.float64) if positions is None else np.array(
            positions, dtype=np.float64).reshape(n_atoms, 3)
        self._id = conf_id

    def GetId(self) -> int:
        return self._id

    def GetNumAtoms(self) -> int:
        return self._pos.shape[0]

    def GetAtomPosition(self, i: int) -> Point3D:
        p = self._pos[i]
        return Point3D(p[0], p[1], p[2])

    def SetAtomPosition(self, i: int, p) -> None:
        if isinstance(p, Point3D):
            p = (p.x, p.y, p.z)
        self._pos[i] = np.asarray(list(p), dtype=np.float64)

    def GetPositions(self) -> np.ndarray:
        return self._pos.copy()

    def SetPositions(self, pos: np.ndarray) -> None:
        self._pos[:] = np.asarray(pos, dtype=np.float64).reshape(self._pos.shape)


@dataclass
class Bond:
    begin: int
    end: int
    order: float  # 1, 2, 3 or 1.5 (aromatic)

    def GetBeginAtomIdx(self) -> int:
        return self.begin

    def GetEndAtomIdx(self) -> int:
        return self.end

    def GetBondTypeAsDouble(self) -> float:
        return float(self.order)

    def GetIsAromatic(self) -> bool:
        return self.order == 1.5


@dataclass
class Atom:
    idx: int
    z: int
    aromatic: bool = False
    charge: int = 0
    mol: "Mol" = field(default=None, repr=False)

    def GetIdx(self) -> int:
        return self.idx

    def GetAtomicNum(self) -> int:
        return self.z

    def GetSymbol(self) -> str:
        return SYMBOL[self.z]

    def GetIsAromatic(self) -> bool:
        return self.aromatic

    def GetFormalCharge(self) -> int:
        return self.charge

    def GetNeighbors(self) -> List["Atom"]:
        return [self.mol.atoms[j] for j in self.mol.neighbors(self.idx)]

    def GetDegree(self) -> int:
        return len(self.mol.neighbors(self.idx))


class Mol:
    """Molecular graph with explicit hydrogens and any number of conformers."""

    def __init__(self, z: Sequence[int], bonds: Iterable[Tuple[int, int, float]],
                 aromatic: Optional[Sequence[bool]] = None, name: str = ""):
        self.atoms: List[Atom] = []
        aromatic = aromatic if aromatic is not None else [False] * len(z)
        for i, zi in enumerate(z):
            self.atoms.append(Atom(i, int(zi), bool(aromatic[i]), 0, self))
        self.bonds: List[Bond] = [Bond(int(a), int(b), float(o)) for a, b, o in bonds]
        self._nbr: List[List[int]] = [[] for _ in self.atoms]
        self._bond_of = {}
        for k, b in enumerate(self.bonds):
            self._nbr[b.begin].append(b.end)
            self._nbr[b.end].append(b.begin)
            self._bond_of[(b.begin, b.end)] = k
            self._bond_of[(b.end, b.begin)] = k
        self._confs: List[Conformer] = []
        self.name = name

    # --- graph -----------------------------------------------------------
    def neighbors(self, i: int) -> List[int]:
        return self._nbr[i]

    def bond_order(self, i: int, j: int) -> float:
        k = self._bond_of.get((i, j))
        return 0.0 if k is None else self.bonds[k].order

    def GetNumAtoms(self) -> int:
        return len(self.atoms)

    def GetAtoms(self) -> List[Atom]:
        return list(self.atoms)

    def GetAtomWithIdx(self, i: int) -> Atom:
        return self.atoms[i]

    def GetBonds(self) -> List[Bond]:
        return list(self.bonds)

    def atomic_numbers(self) -> np.ndarray:
        return np.array([a.z for a in self.atoms], dtype=np.int32)

    def topological_distances(self) -> np.ndarray:
        """All-pairs shortest path lengths in bonds, like RDKit ``GetDistanceMatrix``
        (unreachable pairs get 10**6)."""
        from scipy.sparse import csr_matrix
        from scipy.sparse.csgraph import shortest_path
        n = len(self.atoms)
        if n == 0:
            return np.zeros((0, 0), dtype=np.int64)
        rows = [b.begin for b in self.bonds] + [b.end for b in self.bonds]
        cols = [b.end for b in self.bonds] + [b.begin for b in self.bonds]
        a = csr_matrix((np.ones(len(rows)), (rows, cols)), shape=(n, n))
        d = shortest_path(a, method="D", unweighted=True)
        d[~np.isfinite(d)] = 10 ** 6
        return d.astype(np.int64)

    # --- conformers ------------------------------------------------------
    def GetNumConformers(self) -> int:
        return len(self._confs)

    def GetConformers(self) -> List[Conformer]:
        return list(self._confs)

    def GetConformer(self, conf_id: int = -1) -> Conformer:
        if conf_id < 0:
            return self._confs[0]
        for c in self._confs:
            if c.GetId() == conf_id:
                return c
        raise ValueError(f"Bad Conformer Id {conf_id}")

    def AddConformer(self, conf: Conformer, assignId: bool = False) -> int:
        if assignId:
            conf._id = (max(c.GetId() for c in self._confs) + 1) if self._confs else 0
        self._confs.append(conf)
        return conf.GetId()

    def RemoveAllConformers(self) -> None:
        self._confs = []

    def set_conformers(self, coords: np.ndarray) -> None:
        """Replace all conformers with ``coords`` of shape (k, N, 3)."""
        coords = np.asarray(coords, dtype=np.float64)
        self._confs = [Conformer(self.GetNumAtoms(), coords[k], k) for k in range(coords.shape[0])]


# ---------------------------------------------------------------------------
# SMILES (organic subset) -> Mol with explicit hydrogens
# ---------------------------------------------------------------------------

_ORGANIC = ["Cl", "Br", "B", "C", "N", "O", "P", "S", "F", "I"]
_AROM = {"c": 6, "n": 7, "o": 8, "s": 16, "p": 15}


def mol_from_smiles(smiles: str, add_hs: bool = True, name: str = "") -> Mol:
    """Parse an organic-subset SMILES string; hydrogens become explicit atoms.

    Heavy atoms keep their SMILES order and hydrogens are appended after them in heavy-atom
    order, which is the atom order RDKit's ``AddHs`` produces.
    """
    z: List[int] = []
    arom: List[bool] = []
    exh: List[Optional[int]] = []
    bonds: List[Tuple[int, int, float]] = []
    stack: List[int] = []
    ring: dict = {}
    prev = -1
    pend_order: Optional[float] = None
    i = 0
    s = smiles
    while i < len(s):
        ch = s[i]
        if ch == "(":
            stack.append(prev)
            i += 1
            continue
        if ch == ")":
            prev = stack.pop()
            i += 1
            continue
        if ch in "-=#:":
            pend_order = {"-": 1.0, "=": 2.0, "#": 3.0, ":": 1.5}[ch]
            i += 1
            continue
        if ch == ".":  # a new disconnected fragment: the next atom bonds to nothing
            prev = -1
            pend_order = None
            i += 1
            continue
        if ch.isdigit() or ch == "%":
            if ch == "%":
                num = int(s[i + 1:i + 3])
                i += 3
            else:
                num = int(ch)
                i += 1
            if num in ring:
                j, o_open = ring.pop(num)
                o = pend_order or o_open
                if o is None:
                    o = 1.5 if (arom[j] and arom[prev]) else 1.0
                bonds.append((j, prev, o))
            else:
                ring[num] = (prev, pend_order)
            pend_order = None
            continue
        # atom
        hcount: Optional[int] = None
        if ch == "[":
            k = s.index("]", i)
            body = s[i + 1:k]
            i = k + 1
            m = 0
            sym = None
            if body[:2] in ("Cl", "Br"):
                sym, m = body[:2], 2
            elif body[0] in _AROM:
                sym, m = body[0], 1
            else:
                sym, m = body[0], 1
            rest = body[m:]
            hcount = 0
            if "H" in rest:
                p = rest.index("H")
                digits = ""
                q = p + 1
                while q < len(rest) and rest[q].isdigit():
                    digits += rest[q]
                    q += 1
                hcount = int(digits) if digits else 1
            is_arom = sym in _AROM
            zi = _AROM[sym] if is_arom else ELEMENTS[sym]
        else:
            if s[i:i + 2] in ("Cl", "Br"):
                sym = s[i:i + 2]
                i += 2
            else:
                sym = ch
                i += 1
            is_arom = sym in _AROM
            zi = _AROM[sym] if is_arom else ELEMENTS[sym]
        idx = len(z)
        z.append(zi)
        arom.append(is_arom)
        exh.append(hcount)
        if prev >= 0:
            o = pend_order
            if o is None:
                o = 1.5 if (arom[prev] and is_arom) else 1.0
            bonds.append((prev, idx, o))
        pend_order = None
        prev = idx
    nh = _implicit_h(z, arom, bonds, exh)
    heavy = len(z)
    if add_hs:
        for a in range(heavy):
            for _ in range(nh[a]):
                hidx = len(z)
                z.append(1)
                arom.append(False)
                bonds.append((a, hidx, 1.0))
    return Mol(z, bonds, arom, name=name or smiles)


def _implicit_h(z, arom, bonds, exh):
    n = len(z)
    vsum = np.zeros(n)
    narom = np.zeros(n, dtype=int)
    for a, b, o in bonds:
        for x in (a, b):
            if o == 1.5:
                narom[x] += 1
            else:
                vsum[x] += o
    out = []
    for i in range(n):
        if exh[i] is not None:
            out.append(exh[i])
            continue
        val = DEFAULT_VALENCE.get(z[i], 0)
        used = vsum[i]
        if arom[i]:
            # aromatic atoms: two ring bonds count as 3 valence units for C (one
            # double-bond equivalent); pyridine-type n uses 3, furan/thiophene o/s use 2
            if z[i] == 6:
                used += 3 if narom[i] >= 2 else narom[i]
                if narom[i] == 3:
                    used = vsum[i] + 4
            elif z[i] == 7:
                used += 3 if narom[i] >= 2 else narom[i]
            else:
                used += narom[i]
        else:
            used += narom[i]
        out.append(max(int(round(val - used)), 0))
    return out
