"""MMFF94/MMFF94s parameter tables per molecule (drop-in for mlxmolkit/mmff_params.py).

``MMFFParams`` keeps the reference field layout (mlxmolkit/mmff_params.py:30-81) so code
written against the reference reads the same arrays. ``extract_mmff_params`` follows the
reference's term enumeration rules exactly (mlxmolkit/mmff_params.py:129-320):

* bonds in bond order; angles per centre over neighbour pairs (ii < jj);
* stretch-bend per angle, skipped at linear (sp) centres (:182);
* out-of-plane: three permutations per degree-3 centre (:223-227);
* torsions per bond over neighbour pairs, excluding i1 == i3 and i4 in {i2, i1} (:242-253);
* non-bonded: all pairs at topological distance >= 3 within ``non_bonded_thresh`` of the
  chosen conformer (:284-292); ele pairs need |q| > 1e-10 on both atoms (:304), and
  ``ele_is_1_4`` flags topological distance exactly 3.

Parameters come from RDKit when the molecule is an RDKit ``Chem.Mol`` and RDKit is
importable (the reference's own source), otherwise from the built-in MMFF94-style typer
(:mod:`mmff_typer`, parity with RDKit's numbers unpinned). The O(N^2) pair loop is
vectorised with numpy instead of the reference's Python double loop.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .mol import Mol
from .mmff_typer import MMFFTyper

RELATION_1_2 = 0
RELATION_1_3 = 1
RELATION_1_4 = 2
RELATION_1_X = 3


def _i32():
    return np.empty(0, dtype=np.int32)


def _f32():
    return np.empty(0, dtype=np.float32)


@dataclass
class MMFFParams:
    """All MMFF94 parameters for a molecule, stored as flat numpy arrays."""

    n_atoms: int

    bond_idx1: np.ndarray = field(default_factory=_i32)
    bond_idx2: np.ndarray = field(default_factory=_i32)
    bond_kb: np.ndarray = field(default_factory=_f32)
    bond_r0: np.ndarray = field(default_factory=_f32)

    angle_idx1: np.ndarray = field(default_factory=_i32)
    angle_idx2: np.ndarray = field(default_factory=_i32)
    angle_idx3: np.ndarray = field(default_factory=_i32)
    angle_ka: np.ndarray = field(default_factory=_f32)
    angle_theta0: np.ndarray = field(default_factory=_f32)
    angle_is_linear: np.ndarray = field(default_factory=_i32)

    strbend_idx1: np.ndarray = field(default_factory=_i32)
    strbend_idx2: np.ndarray = field(default_factory=_i32)
    strbend_idx3: np.ndarray = field(default_factory=_i32)
    strbend_kba_ijk: np.ndarray = field(default_factory=_f32)
    strbend_kba_kji: np.ndarray = field(default_factory=_f32)
    strbend_r0_ij: np.ndarray = field(default_factory=_f32)
    strbend_r0_kj: np.ndarray = field(default_factory=_f32)
    strbend_theta0: np.ndarray = field(default_factory=_f32)

    oop_idx1: np.ndarray = field(default_factory=_i32)
    oop_idx2: np.ndarray = field(default_factory=_i32)
    oop_idx3: np.ndarray = field(default_factory=_i32)
    oop_idx4: np.ndarray = field(default_factory=_i32)
    oop_koop: np.ndarray = field(default_factory=_f32)

    torsion_idx1: np.ndarray = field(default_factory=_i32)
    torsion_idx2: np.ndarray = field(default_factory=_i32)
    torsion_idx3: np.ndarray = field(default_factory=_i32)
    torsion_idx4: np.ndarray = field(default_factory=_i32)
    torsion_V1: np.ndarray = field(default_factory=_f32)
    torsion_V2: np.ndarray = field(default_factory=_f32)
    torsion_V3: np.ndarray = field(default_factory=_f32)

    vdw_idx1: np.ndarray = field(default_factory=_i32)
    vdw_idx2: np.ndarray = field(default_factory=_i32)
    vdw_R_star: np.ndarray = field(default_factory=_f32)
    vdw_eps: np.ndarray = field(default_factory=_f32)

    ele_idx1: np.ndarray = field(default_factory=_i32)
    ele_idx2: np.ndarray = field(default_factory=_i32)
    ele_charge_term: np.ndarray = field(default_factory=_f32)
    ele_diel_model: int = 1
    ele_is_1_4: np.ndarray = field(default_factory=_i32)

    def term_counts(self) -> dict:
        return dict(bond=len(self.bond_idx1), angle=len(self.angle_idx1), strbend=len(self.strbend_idx1),
                    oop=len(self.oop_idx1), torsion=len(self.torsion_idx1), vdw=len(self.vdw_idx1),
                    ele=len(self.ele_idx1))


def _is_rdkit_mol(mol) -> bool:
    return type(mol).__module__.startswith("rdkit")


class _RDKitSource:
    """Parameter source backed by RDKit's MMFF (the reference's own provider)."""

    def __init__(self, mol, variant):
        from rdkit.Chem import rdForceFieldHelpers  # noqa: local optional import
        from rdkit import Chem
        self.Chem = Chem
        self.mol = mol
        self.mp = rdForceFieldHelpers.MMFFGetMoleculeProperties(mol, mmffVariant=variant)
        if self.mp is None:
            raise ValueError("Could not compute MMFF properties for molecule")

    def n_atoms(self):
        return self.mol.GetNumAtoms()

    def neighbors(self, i):
        return [a.GetIdx() for a in self.mol.GetAtomWithIdx(i).GetNeighbors()]

    def bonds(self):
        return [(b.GetBeginAtomIdx(), b.GetEndAtomIdx()) for b in self.mol.GetBonds()]

    def is_linear(self, i):
        return self.mol.GetAtomWithIdx(i).GetHybridization() == self.Chem.HybridizationType.SP

    def atom_type(self, i):
        return int(self.mp.GetMMFFAtomType(i))

    def bond(self, i, j):
        r = self.mp.GetMMFFBondStretchParams(self.mol, i, j)
        return None if not r else (r[1], r[2])

    def angle(self, i, j, k):
        r = self.mp.GetMMFFAngleBendParams(self.mol, i, j, k)
        return None if not r else (r[1], r[2])

    def strbend(self, i, j, k):
        r = self.mp.GetMMFFStretchBendParams(self.mol, i, j, k)
        return None if not r else (r[1], r[2])

    def oop(self, i, j, k, l):
        return self.mp.GetMMFFOopBendParams(self.mol, i, j, k, l)

    def torsion(self, i, j, k, l):
        r = self.mp.GetMMFFTorsionParams(self.mol, i, j, k, l)
        return None if not r else (r[1], r[2], r[3])

    def topo(self):
        return self.Chem.GetDistanceMatrix(self.mol)

    def charges(self):
        return np.array([self.mp.GetMMFFPartialCharge(i) for i in range(self.n_atoms())], dtype=np.float64)

    def vdw_pairs(self, ii, jj):
        R = np.empty(len(ii))
        E = np.empty(len(ii))
        ok = np.ones(len(ii), dtype=bool)
        for k, (i, j) in enumerate(zip(ii, jj)):
            v = self.mp.GetMMFFVdWParams(int(i), int(j))
            if v:
                R[k], E[k] = v[2], v[3]
            else:
                ok[k] = False
        return R, E, ok


class _OwnSource:
    """Parameter source backed by the built-in MMFF94-style typer."""

    def __init__(self, mol: Mol, variant: str):
        self.mol = mol
        self.t = MMFFTyper(mol, variant)

    def n_atoms(self):
        return self.mol.GetNumAtoms()

    def neighbors(self, i):
        return list(self.mol.neighbors(i))

    def bonds(self):
        return [(b.begin, b.end) for b in self.mol.bonds]

    def is_linear(self, i):
        return self.t.is_linear(i)

    def atom_type(self, i):
        return int(self.t.types[i])

    def bond(self, i, j):
        return self.t.bond(i, j)

    def angle(self, i, j, k):
        return self.t.angle(i, j, k)

    def strbend(self, i, j, k):
        return self.t.strbend(i, j, k)

    def oop(self, i, j, k, l):
        return self.t.oop(j)

    def torsion(self, i, j, k, l):
        return self.t.torsion(i, j, k, l)

    def topo(self):
        return self.mol.topological_distances()

    def charges(self):
        return self.t.charges

    def vdw_pairs(self, ii, jj):
        # parameters depend only on the two atom types: evaluate each type pair once
        types = np.asarray(self.t.types, dtype=np.int64)
        ii = np.asarray(ii, dtype=np.int64)
        jj = np.asarray(jj, dtype=np.int64)
        key = types[ii] * 1024 + types[jj]
        uk, first, inv = np.unique(key, return_index=True, return_inverse=True)
        table = np.array([self.t.vdw(int(ii[f]), int(jj[f]))[2:] for f in first], dtype=np.float64).reshape(-1, 2)
        return table[inv, 0], table[inv, 1], np.ones(len(ii), dtype=bool)


def _positions(mol, conf_id: int) -> np.ndarray:
    conf = mol.GetConformer(conf_id)
    if hasattr(conf, "GetPositions"):
        return np.asarray(conf.GetPositions(), dtype=np.float64)
    n = mol.GetNumAtoms()
    out = np.zeros((n, 3))
    for i in range(n):
        p = conf.GetAtomPosition(i)
        out[i] = (p.x, p.y, p.z)
    return out


# MMFF atom types whose angles RDKit treats as linear (its MMFFProp::linh flag), the rule the
# reference's second extractor follows (mlxmolkit/mmff_extract_shivam.py:16-17,127-129)
MMFF_LINEAR_TYPES = frozenset({4, 53, 61})


def extract_mmff_params(mol, conf_id: int = 0, non_bonded_thresh: float = 100.0,
                        mmff_variant: str = "MMFF94", *, linear_rule: str = "hybridization",
                        ignore_interfrag_interactions: bool = False) -> MMFFParams:
    """Extract MMFF94 or MMFF94s parameters (same signature as the reference :100-105).

    Keyword-only opt-ins (defaults reproduce mlxmolkit/mmff_params.py):

    * ``linear_rule``: which angle centres take the linear bend form and lose their
      stretch-bend terms. ``"hybridization"`` (default): sp centres, as mmff_params.py:168
      does; ``"mmff_types"``: centres of MMFF type 4, 53 or 61, RDKit's own rule
      (mmff_extract_shivam.py:127-129,204,230-238).
    * ``ignore_interfrag_interactions``: drop vdW and electrostatic pairs between
      disconnected fragments (mmff_extract_shivam.py:361-378; RDKit's MMFF default). The
      default keeps them, as mmff_params.py:284-309 does.

    Raises ``ValueError`` when the molecule cannot be typed, as the reference does (:123-124).
    """
    if linear_rule not in ("hybridization", "mmff_types"):
        raise ValueError(f"linear_rule must be 'hybridization' or 'mmff_types', not {linear_rule!r}")
    src = _RDKitSource(mol, mmff_variant) if _is_rdkit_mol(mol) else _OwnSource(mol, mmff_variant)
    if linear_rule == "mmff_types":
        src.is_linear = lambda i: src.atom_type(i) in MMFF_LINEAR_TYPES
    n = src.n_atoms()
    p = MMFFParams(n_atoms=n)

    # 1. bonds (:129-146)
    b1, b2, bk, br = [], [], [], []
    r0map = {}
    for i, j in src.bonds():
        r = src.bond(i, j)
        if r:
            kb, r0 = r
            b1.append(i), b2.append(j), bk.append(kb), br.append(r0)
            r0map[(i, j)] = r0
            r0map[(j, i)] = r0
    p.bond_idx1, p.bond_idx2 = np.array(b1, np.int32), np.array(b2, np.int32)
    p.bond_kb, p.bond_r0 = np.array(bk, np.float32), np.array(br, np.float32)

    # 2. angles (:148-173) and 3. stretch-bend (:175-210)
    a1, a2, a3, ak, at, al = [], [], [], [], [], []
    s1, s2, s3, sij, skj, sr1, sr2, st = [], [], [], [], [], [], [], []
    for c in range(n):
        nb = src.neighbors(c)
        if len(nb) < 2:
            continue
        lin = src.is_linear(c)
        for ii in range(len(nb)):
            for jj in range(ii + 1, len(nb)):
                i, j = nb[ii], nb[jj]
                ra = src.angle(i, c, j)
                if ra:
                    a1.append(i), a2.append(c), a3.append(j), ak.append(ra[0]), at.append(ra[1])
                    al.append(1 if lin else 0)
    for c in range(n):
        nb = src.neighbors(c)
        if len(nb) < 2 or src.is_linear(c):
            continue
        for ii in range(len(nb)):
            for jj in range(ii + 1, len(nb)):
                i, j = nb[ii], nb[jj]
                rs = src.strbend(i, c, j)
                ra = src.angle(i, c, j)
                if rs and ra:
                    s1.append(i), s2.append(c), s3.append(j)
                    sij.append(rs[0]), skj.append(rs[1])
                    sr1.append(r0map.get((i, c), 1.5)), sr2.append(r0map.get((j, c), 1.5))
                    st.append(ra[1])
    p.angle_idx1, p.angle_idx2, p.angle_idx3 = (np.array(a, np.int32) for a in (a1, a2, a3))
    p.angle_ka, p.angle_theta0 = np.array(ak, np.float32), np.array(at, np.float32)
    p.angle_is_linear = np.array(al, np.int32)
    p.strbend_idx1, p.strbend_idx2, p.strbend_idx3 = (np.array(a, np.int32) for a in (s1, s2, s3))
    p.strbend_kba_ijk, p.strbend_kba_kji = np.array(sij, np.float32), np.array(skj, np.float32)
    p.strbend_r0_ij, p.strbend_r0_kj = np.array(sr1, np.float32), np.array(sr2, np.float32)
    p.strbend_theta0 = np.array(st, np.float32)

    # 4. out-of-plane (:212-238)
    o1, o2, o3, o4, ok_ = [], [], [], [], []
    for c in range(n):
        nb = src.neighbors(c)
        if len(nb) != 3:
            continue
        k = src.oop(nb[0], c, nb[1], nb[2])
        if k is None:
            continue
        for q1, q3, q4 in ((nb[0], nb[1], nb[2]), (nb[0], nb[2], nb[1]), (nb[1], nb[2], nb[0])):
            o1.append(q1), o2.append(c), o3.append(q3), o4.append(q4), ok_.append(k)
    p.oop_idx1, p.oop_idx2, p.oop_idx3, p.oop_idx4 = (np.array(a, np.int32) for a in (o1, o2, o3, o4))
    p.oop_koop = np.array(ok_, np.float32)

    # 5. torsions (:240-270)
    t1, t2, t3, t4, v1, v2, v3 = [], [], [], [], [], [], []
    for i2, i3 in src.bonds():
        for i1 in src.neighbors(i2):
            if i1 == i3:
                continue
            for i4 in src.neighbors(i3):
                if i4 == i2 or i4 == i1:
                    continue
                r = src.torsion(i1, i2, i3, i4)
                if r:
                    t1.append(i1), t2.append(i2), t3.append(i3), t4.append(i4)
                    v1.append(r[0]), v2.append(r[1]), v3.append(r[2])
    p.torsion_idx1, p.torsion_idx2, p.torsion_idx3, p.torsion_idx4 = (
        np.array(a, np.int32) for a in (t1, t2, t3, t4))
    p.torsion_V1, p.torsion_V2, p.torsion_V3 = (np.array(a, np.float32) for a in (v1, v2, v3))

    # 6 & 7. non-bonded (:272-320), vectorised over the upper triangle
    dm = np.asarray(src.topo())
    rel = np.full((n, n), RELATION_1_X, dtype=np.int8)
    rel[dm == 1] = RELATION_1_2
    rel[dm == 2] = RELATION_1_3
    rel[dm == 3] = RELATION_1_4
    pos = _positions(mol, conf_id)
    iu, ju = np.triu_indices(n, 1)
    keep = rel[iu, ju] >= RELATION_1_4
    if ignore_interfrag_interactions:
        # unreachable pairs (different fragments) have no finite bond path
        keep &= dm[iu, ju] < n
    if keep.any():
        d = np.linalg.norm(pos[iu] - pos[ju], axis=1)
        keep &= d <= non_bonded_thresh
    ii, jj = iu[keep].astype(np.int32), ju[keep].astype(np.int32)
    R, E, okv = src.vdw_pairs(ii, jj)
    p.vdw_idx1, p.vdw_idx2 = ii[okv], jj[okv]
    p.vdw_R_star, p.vdw_eps = R[okv].astype(np.float32), E[okv].astype(np.float32)
    q = src.charges()
    okq = (np.abs(q[ii]) > 1e-10) & (np.abs(q[jj]) > 1e-10)
    p.ele_idx1, p.ele_idx2 = ii[okq], jj[okq]
    p.ele_charge_term = (q[ii] * q[jj])[okq].astype(np.float32)
    p.ele_is_1_4 = (rel[ii, jj] == RELATION_1_4)[okq].astype(np.int32)
    p.ele_diel_model = 0
    return p
