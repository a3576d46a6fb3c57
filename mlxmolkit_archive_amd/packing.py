"""Ragged device layout of MMFF term tables (replaces pack_multi_mol_for_metal and
_pack_mmff_for_nk, mlxmolkit/mmff_metal_kernel.py:594-687, mlxmolkit/mmff_minimize.py:40-173).

Differences from the reference layouts, by design:

* tables are stored once per molecule and conformers point at them through
  ``conf_mol`` (the nk packer replicates every table per conformer, :99-144);
* coordinates are ragged (sum of 3N per conformer) with no padding to ``max_dim``;
* the reference's vdW and electrostatic pair lists are merged into one pair list
  ``{R*, eps, q_i q_j * scale}`` (scale 0.75 for 1-4 pairs, :651-652), sorted by
  (j - i, i) so that the 64 lanes of a wave touch 64 distinct atoms per gradient update;
* SoA with 16-bit packed atom indices (a molecule has at most 65535 atoms).
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import List, Optional, Sequence

import numpy as np

DESC_STRIDE = 16


@dataclass
class HostBatch:
    mol_desc: np.ndarray
    bond_ij: np.ndarray
    bond_p: np.ndarray
    ang_idx: np.ndarray
    ang_p: np.ndarray
    sb_idx: np.ndarray
    sb_p: np.ndarray
    sb_t0: np.ndarray
    oop_idx: np.ndarray
    oop_k: np.ndarray
    tor_idx: np.ndarray
    tor_p: np.ndarray
    pair_ij: np.ndarray
    pair_p: np.ndarray
    conf_mol: np.ndarray
    conf_off: np.ndarray
    n_atoms: np.ndarray        # per molecule
    conf_counts: np.ndarray
    total_coords: int
    term_bytes: np.ndarray     # canonical reference term-table bytes per molecule

    @property
    def n_mols(self) -> int:
        return len(self.n_atoms)

    @property
    def n_confs(self) -> int:
        return len(self.conf_mol)

    @property
    def max_atoms(self) -> int:
        return int(self.n_atoms.max()) if len(self.n_atoms) else 1


def _pack16(a, b):
    return (np.asarray(a, np.uint32) & 0xFFFF) | ((np.asarray(b, np.uint32) & 0xFFFF) << 16)


def merge_pairs(p):
    """Union of the vdW and ele pair lists -> (i, j, R*, eps, qq*scale), sorted by (j-i, i)."""
    n = p.n_atoms
    vi, vj = np.asarray(p.vdw_idx1, np.int64), np.asarray(p.vdw_idx2, np.int64)
    ei, ej = np.asarray(p.ele_idx1, np.int64), np.asarray(p.ele_idx2, np.int64)
    vlo, vhi = np.minimum(vi, vj), np.maximum(vi, vj)
    elo, ehi = np.minimum(ei, ej), np.maximum(ei, ej)
    kv = vlo * n + vhi
    ke = elo * n + ehi
    keys = np.union1d(kv, ke)
    m = len(keys)
    R = np.ones(m, np.float32)
    eps = np.zeros(m, np.float32)
    qq = np.zeros(m, np.float32)
    if len(kv):
        pos = np.searchsorted(keys, kv)
        R[pos] = p.vdw_R_star
        eps[pos] = p.vdw_eps
    if len(ke):
        pos = np.searchsorted(keys, ke)
        scale = np.where(np.asarray(p.ele_is_1_4) != 0, np.float32(0.75), np.float32(1.0))
        qq[pos] = (np.asarray(p.ele_charge_term, np.float32) * scale).astype(np.float32)
    ii = keys // n
    jj = keys % n
    order = np.lexsort((ii, jj - ii))
    return ii[order], jj[order], R[order], eps[order], qq[order]


_INDEX_FIELDS = ("bond_idx1", "bond_idx2", "angle_idx1", "angle_idx2", "angle_idx3", "strbend_idx1", "strbend_idx2",
                 "strbend_idx3", "oop_idx1", "oop_idx2", "oop_idx3", "oop_idx4", "torsion_idx1", "torsion_idx2",
                 "torsion_idx3", "torsion_idx4", "vdw_idx1", "vdw_idx2", "ele_idx1", "ele_idx2")
_GROUPS = (("bond_idx1", "bond_kb", "bond_r0"), ("angle_idx1", "angle_ka", "angle_theta0", "angle_is_linear"),
           ("strbend_idx1", "strbend_kba_ijk", "strbend_kba_kji", "strbend_r0_ij", "strbend_r0_kj", "strbend_theta0"),
           ("oop_idx1", "oop_koop"), ("torsion_idx1", "torsion_V1", "torsion_V2", "torsion_V3"),
           ("vdw_idx1", "vdw_R_star", "vdw_eps"), ("ele_idx1", "ele_charge_term", "ele_is_1_4"))


def _validate(p) -> None:
    """Every atom index must address the molecule: the kernels index LDS with them unchecked."""
    n = int(p.n_atoms)
    for f in _INDEX_FIELDS:
        a = np.asarray(getattr(p, f))
        if a.size and (a.min() < 0 or a.max() >= n):
            raise ValueError(f"MMFFParams.{f} has an atom index outside [0, {n})")
    for grp in _GROUPS:
        ln = len(getattr(p, grp[0]))
        for f in grp[1:]:
            if len(getattr(p, f)) != ln:
                raise ValueError(f"MMFFParams.{f} length {len(getattr(p, f))} != {grp[0]} length {ln}")
    for a, b in (("bond_idx1", "bond_idx2"), ("vdw_idx1", "vdw_idx2"), ("ele_idx1", "ele_idx2")):
        if len(getattr(p, a)) != len(getattr(p, b)):
            raise ValueError(f"MMFFParams.{a} and {b} differ in length")


_TERM_BYTES = dict(bond=16, angle=20, strbend=32, oop=20, torsion=28, vdw=16, ele=16)


def canonical_term_bytes(p) -> int:
    """Reference mega-SoA bytes of one molecule's tables (mmff_metal_kernel.py:638-666)."""
    return (16 * len(p.bond_idx1) + 20 * len(p.angle_idx1) + 32 * len(p.strbend_idx1) + 20 * len(p.oop_idx1)
            + 28 * len(p.torsion_idx1) + 16 * len(p.vdw_idx1) + 16 * len(p.ele_idx1))


def pack(params_list: Sequence, conf_counts: Sequence[int]) -> HostBatch:
    M = len(params_list)
    desc = np.zeros((M, DESC_STRIDE), np.int32)
    parts = {k: [] for k in ("bond_ij", "bond_p", "ang_idx", "ang_p", "sb_idx", "sb_p", "sb_t0", "oop_idx", "oop_k",
                             "tor_idx", "tor_p", "pair_ij", "pair_p")}
    cur = dict(bond=0, ang=0, sb=0, oop=0, tor=0, pair=0)
    n_atoms = np.zeros(M, np.int32)
    tb = np.zeros(M, np.int64)
    for m, p in enumerate(params_list):
        if p.n_atoms > 65535:
            raise ValueError("molecules above 65535 atoms are not supported")
        n_atoms[m] = p.n_atoms
        _validate(p)
        tb[m] = canonical_term_bytes(p)
        nb, na, nsb = len(p.bond_idx1), len(p.angle_idx1), len(p.strbend_idx1)
        noop, nt = len(p.oop_idx1), len(p.torsion_idx1)
        pi, pj, R, eps, qq = merge_pairs(p)
        npair = len(pi)
        desc[m, :13] = (p.n_atoms, cur["bond"], nb, cur["ang"], na, cur["sb"], nsb, cur["oop"], noop, cur["tor"], nt,
                        cur["pair"], npair)
        cur["bond"] += nb
        cur["ang"] += na
        cur["sb"] += nsb
        cur["oop"] += noop
        cur["tor"] += nt
        cur["pair"] += npair
        parts["bond_ij"].append(_pack16(p.bond_idx1, p.bond_idx2))
        parts["bond_p"].append(np.stack([p.bond_kb, p.bond_r0], 1).astype(np.float32).reshape(-1, 2))
        lin = np.asarray(p.angle_is_linear, np.uint32) != 0
        parts["ang_idx"].append(np.stack([_pack16(p.angle_idx1, p.angle_idx2),
                                          _pack16(p.angle_idx3, lin.astype(np.uint32))], 1).reshape(-1, 2))
        parts["ang_p"].append(np.stack([p.angle_ka, p.angle_theta0], 1).astype(np.float32).reshape(-1, 2))
        parts["sb_idx"].append(np.stack([_pack16(p.strbend_idx1, p.strbend_idx2),
                                         np.asarray(p.strbend_idx3, np.uint32)], 1).reshape(-1, 2))
        parts["sb_p"].append(np.stack([p.strbend_kba_ijk, p.strbend_kba_kji, p.strbend_r0_ij, p.strbend_r0_kj],
                                      1).astype(np.float32).reshape(-1, 4))
        parts["sb_t0"].append(np.asarray(p.strbend_theta0, np.float32))
        parts["oop_idx"].append(np.stack([_pack16(p.oop_idx1, p.oop_idx2), _pack16(p.oop_idx3, p.oop_idx4)],
                                         1).reshape(-1, 2))
        parts["oop_k"].append(np.asarray(p.oop_koop, np.float32))
        parts["tor_idx"].append(np.stack([_pack16(p.torsion_idx1, p.torsion_idx2),
                                          _pack16(p.torsion_idx3, p.torsion_idx4)], 1).reshape(-1, 2))
        parts["tor_p"].append(np.stack([p.torsion_V1, p.torsion_V2, p.torsion_V3, np.zeros(nt, np.float32)],
                                       1).astype(np.float32).reshape(-1, 4))
        parts["pair_ij"].append(_pack16(pi, pj))
        # w = 1 / R*, rounded once here, so the pair loop needs no reciprocal of a table constant
        inv_r = (1.0 / np.maximum(np.asarray(R, np.float64), 1e-8)).astype(np.float32)
        parts["pair_p"].append(np.stack([R, eps, qq, inv_r], 1).astype(np.float32))

    def cat(k, shape_tail, dtype):
        arrs = [a for a in parts[k] if a.size]
        if not arrs:
            return np.zeros((1,) + shape_tail, dtype)
        return np.ascontiguousarray(np.concatenate(arrs).astype(dtype, copy=False))

    cc = np.asarray(conf_counts, np.int64)
    conf_mol = np.repeat(np.arange(M, dtype=np.int32), cc)
    n3 = 3 * n_atoms[conf_mol].astype(np.int64)
    conf_off = np.zeros(len(conf_mol), np.int64)
    if len(conf_off) > 1:
        conf_off[1:] = np.cumsum(n3)[:-1]
    total = int(n3.sum())
    if total >= 2 ** 31:
        raise ValueError("batch exceeds 2^31 coordinates; split it")
    return HostBatch(
        mol_desc=desc.reshape(-1), bond_ij=cat("bond_ij", (), np.uint32), bond_p=cat("bond_p", (2,), np.float32),
        ang_idx=cat("ang_idx", (2,), np.uint32), ang_p=cat("ang_p", (2,), np.float32),
        sb_idx=cat("sb_idx", (2,), np.uint32), sb_p=cat("sb_p", (4,), np.float32), sb_t0=cat("sb_t0", (), np.float32),
        oop_idx=cat("oop_idx", (2,), np.uint32), oop_k=cat("oop_k", (), np.float32),
        tor_idx=cat("tor_idx", (2,), np.uint32), tor_p=cat("tor_p", (4,), np.float32),
        pair_ij=cat("pair_ij", (), np.uint32), pair_p=cat("pair_p", (4,), np.float32),
        conf_mol=conf_mol, conf_off=conf_off.astype(np.int32), n_atoms=n_atoms, conf_counts=cc.astype(np.int32),
        total_coords=total, term_bytes=tb)


class DeviceBatch:
    """Device copy of a HostBatch plus the C-ABI descriptor pointing at it."""

    _FIELDS = ("mol_desc", "bond_ij", "bond_p", "ang_idx", "ang_p", "sb_idx", "sb_p", "sb_t0", "oop_idx", "oop_k",
               "tor_idx", "tor_p", "pair_ij", "pair_p", "conf_mol", "conf_off")

    def __init__(self, hb: HostBatch, device="cuda"):
        import torch
        from . import _lib
        self.host = hb
        self.t = {}
        for f in self._FIELDS:
            a = getattr(hb, f)
            if a.dtype == np.uint32:
                a = a.view(np.int32)
            self.t[f] = torch.from_numpy(np.ascontiguousarray(a)).to(device)
        s = _lib.Batch()
        for f in self._FIELDS:
            setattr(s, f, self.t[f].data_ptr())
        s.n_mols = hb.n_mols
        s.n_confs = hb.n_confs
        s.max_atoms = hb.max_atoms
        self.desc = s
        self.device = device

    @property
    def n_confs(self) -> int:
        return self.host.n_confs


def lpt_order(hb: HostBatch) -> np.ndarray:
    """Conformer order for the work queue: most expensive first (longest-processing-time)."""
    n = 3 * hb.n_atoms[hb.conf_mol].astype(np.float64)
    cost = n * n + hb.term_bytes[hb.conf_mol] / 16.0
    return np.argsort(-cost, kind="stable").astype(np.int32)
