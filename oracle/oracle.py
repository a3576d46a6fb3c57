"""ctypes front-end of the FP64 CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
module, and only as the checker. The product package never imports it.

Functions take reference-layout ``MMFFParams`` objects (one per molecule), a ragged
conformer list and flat float64 coordinates, and call the C restatement in
mmff_oracle.c (see its header for which test pins each function).
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path
from typing import List, Sequence

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIB = None

FAMILY_TG = 0
FAMILY_MEGA = 1
T_ALL = 127
TERM_NAMES = ("bond", "angle", "strbend", "oop", "torsion", "vdw", "ele")

_i32p = C.POINTER(C.c_int32)
_f32p = C.POINTER(C.c_float)


class OracleMol(C.Structure):
    _fields_ = [
        ("n_atoms", C.c_int32),
        ("nb", C.c_int32), ("b_i", _i32p), ("b_j", _i32p), ("b_kb", _f32p), ("b_r0", _f32p),
        ("na", C.c_int32), ("a_i", _i32p), ("a_j", _i32p), ("a_k", _i32p), ("a_ka", _f32p), ("a_t0", _f32p),
        ("a_lin", _i32p),
        ("nsb", C.c_int32), ("s_i", _i32p), ("s_j", _i32p), ("s_k", _i32p), ("s_kij", _f32p), ("s_kkj", _f32p),
        ("s_r0ij", _f32p), ("s_r0kj", _f32p), ("s_t0", _f32p),
        ("noop", C.c_int32), ("o_i", _i32p), ("o_j", _i32p), ("o_k", _i32p), ("o_l", _i32p), ("o_koop", _f32p),
        ("nt", C.c_int32), ("t_i", _i32p), ("t_j", _i32p), ("t_k", _i32p), ("t_l", _i32p), ("t_v1", _f32p),
        ("t_v2", _f32p), ("t_v3", _f32p),
        ("nv", C.c_int32), ("v_i", _i32p), ("v_j", _i32p), ("v_r", _f32p), ("v_eps", _f32p),
        ("ne", C.c_int32), ("e_i", _i32p), ("e_j", _i32p), ("e_qq", _f32p), ("e_14", _i32p),
    ]


def lib():
    """Load oracle/libmmff_oracle.so (built by ``make -C oracle`` / __graft_entry__.build)."""
    global _LIB
    if _LIB is None:
        path = _HERE / "libmmff_oracle.so"
        if not path.exists():
            raise RuntimeError(f"oracle library missing: {path} (run `make -C oracle`)")
        L = C.CDLL(str(path))
        vp = C.c_void_p
        L.oracle_energy_grad.restype = C.c_double
        L.oracle_energy_grad.argtypes = [vp, C.c_int, C.c_int, vp, vp, C.c_uint32, vp]
        L.oracle_energy_terms_abs.restype = None
        L.oracle_energy_terms_abs.argtypes = [vp, C.c_int, C.c_int, vp, C.c_uint32, vp]
        L.oracle_bfgs_nk.restype = C.c_int
        L.oracle_bfgs_nk.argtypes = [vp, vp, C.c_int, C.c_double, C.c_int, vp, vp]
        L.oracle_lbfgs_nk.restype = C.c_int
        L.oracle_lbfgs_nk.argtypes = [vp, vp, C.c_int, C.c_double, C.c_int, C.c_int, vp, vp]
        L.oracle_bb_fused.restype = None
        L.oracle_bb_fused.argtypes = [vp, vp, C.c_int, C.c_double, C.c_double, vp]
        L.oracle_bfgs_nk_batch.restype = None
        L.oracle_bfgs_nk_batch.argtypes = [vp, vp, vp, C.c_int, vp, C.c_int, C.c_double, C.c_int, C.c_int,
                                           C.c_int, vp, vp, vp, C.c_int]
        L.oracle_bfgs_nk_batch_h0.restype = None
        L.oracle_bfgs_nk_batch_h0.argtypes = [vp, vp, vp, C.c_int, vp, C.c_int, C.c_double, C.c_int, C.c_int,
                                              C.c_int, vp, vp, vp, vp, C.c_int]
        L.oracle_energy_grad_batch.restype = None
        L.oracle_energy_grad_batch.argtypes = [vp, vp, vp, C.c_int, C.c_int, C.c_int, vp, vp, vp, C.c_uint32,
                                               vp, C.c_int]
        L.oracle_bb_fused_batch.restype = None
        L.oracle_bb_fused_batch.argtypes = [vp, vp, vp, C.c_int, vp, C.c_int, C.c_double, C.c_double, vp,
                                            C.c_int]
        L.oracle_multi_mol.restype = C.c_int
        L.oracle_multi_mol.argtypes = [vp, vp, vp, C.c_int, vp, C.c_int, C.c_int, C.c_double, C.c_int,
                                       C.c_double, C.c_double, vp, vp, vp, vp, C.c_int]
        L.oracle_set_fp32_emulation.restype = None
        L.oracle_set_fp32_emulation.argtypes = [C.c_int]
        L.oracle_multi_mol_ex.restype = C.c_int
        L.oracle_multi_mol_ex.argtypes = [vp, vp, vp, C.c_int, vp, C.c_int, C.c_int, C.c_double, C.c_int,
                                          C.c_double, C.c_double, vp, vp, vp, vp, C.c_int, C.c_int]
        _LIB = L
    return _LIB


class _MolTables:
    """Holds the contiguous arrays an OracleMol points into (keeps them alive)."""

    def __init__(self, p):
        self.keep = []
        m = OracleMol()
        m.n_atoms = int(p.n_atoms)

        def I(a):
            a = np.ascontiguousarray(a, dtype=np.int32)
            if a.size == 0:
                a = np.zeros(1, np.int32)
            self.keep.append(a)
            return a.ctypes.data_as(_i32p)

        def F(a):
            a = np.ascontiguousarray(a, dtype=np.float32)
            if a.size == 0:
                a = np.zeros(1, np.float32)
            self.keep.append(a)
            return a.ctypes.data_as(_f32p)

        m.nb = len(p.bond_idx1)
        m.b_i, m.b_j, m.b_kb, m.b_r0 = I(p.bond_idx1), I(p.bond_idx2), F(p.bond_kb), F(p.bond_r0)
        m.na = len(p.angle_idx1)
        m.a_i, m.a_j, m.a_k = I(p.angle_idx1), I(p.angle_idx2), I(p.angle_idx3)
        m.a_ka, m.a_t0, m.a_lin = F(p.angle_ka), F(p.angle_theta0), I(p.angle_is_linear)
        m.nsb = len(p.strbend_idx1)
        m.s_i, m.s_j, m.s_k = I(p.strbend_idx1), I(p.strbend_idx2), I(p.strbend_idx3)
        m.s_kij, m.s_kkj = F(p.strbend_kba_ijk), F(p.strbend_kba_kji)
        m.s_r0ij, m.s_r0kj, m.s_t0 = F(p.strbend_r0_ij), F(p.strbend_r0_kj), F(p.strbend_theta0)
        m.noop = len(p.oop_idx1)
        m.o_i, m.o_j, m.o_k, m.o_l, m.o_koop = I(p.oop_idx1), I(p.oop_idx2), I(p.oop_idx3), I(p.oop_idx4), F(
            p.oop_koop)
        m.nt = len(p.torsion_idx1)
        m.t_i, m.t_j, m.t_k, m.t_l = I(p.torsion_idx1), I(p.torsion_idx2), I(p.torsion_idx3), I(p.torsion_idx4)
        m.t_v1, m.t_v2, m.t_v3 = F(p.torsion_V1), F(p.torsion_V2), F(p.torsion_V3)
        m.nv = len(p.vdw_idx1)
        m.v_i, m.v_j, m.v_r, m.v_eps = I(p.vdw_idx1), I(p.vdw_idx2), F(p.vdw_R_star), F(p.vdw_eps)
        m.ne = len(p.ele_idx1)
        m.e_i, m.e_j, m.e_qq, m.e_14 = I(p.ele_idx1), I(p.ele_idx2), F(p.ele_charge_term), I(p.ele_is_1_4)
        self.mol = m


class OracleBatch:
    """Ragged conformer batch: molecule tables + per-conformer molecule index and offset."""

    def __init__(self, params_list: Sequence, conf_counts: Sequence[int]):
        self.tables = [_MolTables(p) for p in params_list]
        arr = (OracleMol * len(self.tables))()
        for k, t in enumerate(self.tables):
            arr[k] = t.mol
        self.mols = arr
        self.conf_mol = np.repeat(np.arange(len(params_list), dtype=np.int32), np.asarray(conf_counts, int))
        n3 = np.array([3 * params_list[m].n_atoms for m in self.conf_mol], dtype=np.int64)
        self.conf_off = np.zeros(len(self.conf_mol), dtype=np.int64)
        if len(n3) > 1:
            self.conf_off[1:] = np.cumsum(n3)[:-1]
        self.total = int(n3.sum())
        self.n_confs = len(self.conf_mol)

    def _p(self, a):
        return a.ctypes.data_as(C.c_void_p)


def default_threads() -> int:
    try:
        return max(1, len(os.sched_getaffinity(0)))
    except AttributeError:  # pragma: no cover
        return os.cpu_count() or 1


def energy_grad(params, x, family=FAMILY_TG, diel=1, mask=T_ALL, with_terms=False):
    """Energy (and gradient, per-term energies) of one conformer; x is (N,3) or (3N,)."""
    t = _MolTables(params)
    xx = np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(-1))
    g = np.zeros_like(xx)
    et = np.zeros(7)
    e = lib().oracle_energy_grad(C.byref(t.mol), family, diel, xx.ctypes.data, g.ctypes.data, mask,
                                 et.ctypes.data)
    if with_terms:
        return e, g, et
    return e, g


def terms_abs(params, x, family=FAMILY_TG, diel=1, mask=T_ALL):
    """Per-term-type sums of |term energy| of one conformer (7 values)."""
    t = _MolTables(params)
    xx = np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(-1))
    out = np.zeros(7)
    lib().oracle_energy_terms_abs(C.byref(t.mol), family, diel, xx.ctypes.data, mask, out.ctypes.data)
    return out


def energy_grad_batch(batch: OracleBatch, x, family=FAMILY_TG, diel=1, mask=T_ALL, n_threads=None):
    xx = np.ascontiguousarray(x, dtype=np.float64)
    assert xx.size == batch.total
    e = np.zeros(batch.n_confs)
    g = np.zeros_like(xx)
    et = np.zeros(7 * batch.n_confs)
    lib().oracle_energy_grad_batch(batch.mols, batch._p(batch.conf_mol), batch._p(batch.conf_off), batch.n_confs,
                                   family, diel, xx.ctypes.data, e.ctypes.data, g.ctypes.data, mask,
                                   et.ctypes.data, n_threads or default_threads())
    return e, g, et.reshape(-1, 7)


class fp32_emulation:
    """Context manager for the oracle's analysis switch: inside it, the TG energies and the nk
    minimizers round coordinates, term energies and the gradient to float (the HIP kernels'
    storage formats) while keeping FP64 sums. Used only to study trajectory differences."""

    def __enter__(self):
        lib().oracle_set_fp32_emulation(1)
        return self

    def __exit__(self, *exc):
        lib().oracle_set_fp32_emulation(0)
        return False


def minimize_nk(batch: OracleBatch, x, max_iters=500, grad_tol=1e-4, use_lbfgs=False, m_hist=8, diel=1,
                n_threads=None, h0=None):
    """BFGS / L-BFGS of the run_mmff path on every conformer. Returns (x, E, status, counters).
    h0: optional per-atom 3x3 start blocks (total atoms x 9), the learned preconditioner."""
    xx = np.array(x, dtype=np.float64, copy=True).reshape(-1)
    assert xx.size == batch.total
    e = np.zeros(batch.n_confs)
    st = np.zeros(batch.n_confs, dtype=np.int32)
    cnt = np.zeros(3 * batch.n_confs, dtype=np.int32)
    hh = None
    if h0 is not None:
        hh = np.ascontiguousarray(np.asarray(h0, dtype=np.float64).reshape(-1))
        assert hh.size == 3 * batch.total
    lib().oracle_bfgs_nk_batch_h0(batch.mols, batch._p(batch.conf_mol), batch._p(batch.conf_off), batch.n_confs,
                                  xx.ctypes.data, int(max_iters), float(grad_tol), diel, int(bool(use_lbfgs)),
                                  int(m_hist), None if hh is None else hh.ctypes.data, e.ctypes.data,
                                  st.ctypes.data, cnt.ctypes.data, n_threads or default_threads())
    return xx, e, st, cnt.reshape(-1, 3)


def bb_fused(batch: OracleBatch, x, max_iters=500, max_step=0.3, grad_lr=0.1, n_threads=None):
    xx = np.array(x, dtype=np.float64, copy=True).reshape(-1)
    e = np.zeros(batch.n_confs)
    lib().oracle_bb_fused_batch(batch.mols, batch._p(batch.conf_mol), batch._p(batch.conf_off), batch.n_confs,
                                xx.ctypes.data, int(max_iters), float(max_step), float(grad_lr), e.ctypes.data,
                                n_threads or default_threads())
    return xx, e


def multi_mol(batch: OracleBatch, x, max_iters=200, eval_freq=10, max_step_per_atom=0.3, lbfgs_m=10,
              grad_scale=0.1, grad_tol=1e-4, n_threads=None, batch_mode=False):
    """Lock-step host L-BFGS. Returns (x_best, E, grad_norms, per_conf_iters, converged, iters_used).
    ``batch_mode``: the mmff_optimize_metal_batch form (no convergence exit, no polish pass,
    grad_norms = ||g_sc||_2 of the last iterate)."""
    xx = np.array(x, dtype=np.float64, copy=True).reshape(-1)
    e = np.zeros(batch.n_confs)
    gn = np.zeros(batch.n_confs)
    pci = np.zeros(batch.n_confs, dtype=np.int32)
    conv = np.zeros(batch.n_confs, dtype=np.int32)
    used = lib().oracle_multi_mol_ex(batch.mols, batch._p(batch.conf_mol), batch._p(batch.conf_off), batch.n_confs,
                                     xx.ctypes.data, int(max_iters), int(eval_freq), float(max_step_per_atom),
                                     int(lbfgs_m), float(grad_scale), float(grad_tol), e.ctypes.data,
                                     gn.ctypes.data, pci.ctypes.data, conv.ctypes.data,
                                     n_threads or default_threads(), 1 if batch_mode else 0)
    return xx, e, gn, pci, conv.astype(bool), int(used)
