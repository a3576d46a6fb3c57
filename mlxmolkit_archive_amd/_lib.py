"""ctypes binding of libmmff_hip.so (the C ABI declared in include/mmff_hip.h).

The product path has no CPU fallback: if the HIP library or a GPU is missing, every call
raises. Buffers are torch-ROCm tensors; their ``data_ptr()`` is what crosses the ABI.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

from ._build import lib_path

ABI_VERSION = 2

ERR = {0: "ok", -1: "invalid argument", -2: "HIP runtime error", -3: "size limit", -4: "no device"}

FAMILY_TG = 0
FAMILY_MEGA = 1
BFGS = 0
LBFGS = 1
HESS_AUTO, HESS_LDS, HESS_GLOBAL, HESS_GLOBAL_PACKED, HESS_GLOBAL_SYM = 0, 1, 2, 3, 4
MM_BATCH = 1  # mmffhip_multimol_run_ex flag: the mmff_optimize_metal_batch form

_vp = C.c_void_p


class Batch(C.Structure):
    _fields_ = [
        ("mol_desc", _vp), ("bond_ij", _vp), ("bond_p", _vp), ("ang_idx", _vp), ("ang_p", _vp),
        ("sb_idx", _vp), ("sb_p", _vp), ("sb_t0", _vp), ("oop_idx", _vp), ("oop_k", _vp),
        ("tor_idx", _vp), ("tor_p", _vp), ("pair_ij", _vp), ("pair_p", _vp), ("conf_mol", _vp),
        ("conf_off", _vp), ("n_mols", C.c_int32), ("n_confs", C.c_int32), ("max_atoms", C.c_int32),
    ]


class NkOpts(C.Structure):
    _fields_ = [
        ("algo", C.c_int32), ("max_iters", C.c_int32), ("grad_tol", C.c_float), ("diel", C.c_int32),
        ("lbfgs_m", C.c_int32), ("hess_place", C.c_int32), ("block", C.c_int32), ("ls_grad", C.c_int32),
        ("precond", _vp), ("table_lds_bytes", C.c_int32), ("xcd_offsets", _vp),
    ]


# symbol -> (restype, argtypes); the list is also what the export test checks
SIGNATURES = {
    "mmffhip_abi_version": (C.c_int, []),
    "mmffhip_last_error": (C.c_char_p, []),
    "mmffhip_device_info": (C.c_int, [_vp, _vp, _vp]),
    "mmffhip_energy_grad": (C.c_int, [_vp, C.c_int, C.c_int, _vp, _vp, _vp, _vp, C.c_uint32, _vp]),
    "mmffhip_minimize_nk_workspace": (C.c_int, [_vp, _vp, _vp]),
    "mmffhip_minimize_nk_plan": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "mmffhip_minimize_nk": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, C.c_size_t, _vp]),
    "mmffhip_minimize_nk_noalloc": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "mmffhip_bb_fused": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int32, C.c_float, C.c_float, _vp]),
    "mmffhip_multimol_workspace": (C.c_int, [_vp, C.c_int32, _vp]),
    "mmffhip_multimol_run": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, C.c_int32, C.c_int32, C.c_float,
                                       C.c_int32, C.c_float, C.c_float, _vp, C.c_size_t, _vp]),
    "mmffhip_multimol_run_ex": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, C.c_int32, C.c_int32, C.c_float,
                                          C.c_int32, C.c_float, C.c_float, C.c_int32, _vp, C.c_size_t, _vp]),
    "mmffhip_embed": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, C.c_int32, C.c_int32, C.c_float, _vp, _vp]),
    "mmffhip_mlp_dims": (C.c_int, [_vp, _vp, _vp, _vp]),
    "mmffhip_mlp_infer": (C.c_int, [_vp, _vp, C.c_int32, _vp, _vp]),
    "mmffhip_mlp_workspace": (C.c_int, [C.c_int32, _vp]),
    "mmffhip_mlp_forward_backward": (C.c_int, [_vp, _vp, C.c_int32, _vp, _vp, _vp, _vp, _vp, _vp, C.c_size_t, _vp]),
}

_LIB: Optional[C.CDLL] = None


def load() -> C.CDLL:
    """Load the HIP library (raises when it is missing: there is no CPU fallback)."""
    global _LIB
    if _LIB is None:
        p = lib_path("libmmff_hip.so")
        if not p.exists():
            raise RuntimeError(f"{p} is missing; build it with __graft_entry__.build()")
        L = C.CDLL(str(p))
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        v = L.mmffhip_abi_version()
        if v != ABI_VERSION:
            raise RuntimeError(f"libmmff_hip ABI {v} != expected {ABI_VERSION}")
        _LIB = L
    return _LIB


def check(rc: int) -> None:
    if rc != 0:
        msg = load().mmffhip_last_error().decode(errors="replace")
        raise RuntimeError(f"mmffhip error {rc} ({ERR.get(rc, '?')}): {msg}")


def require_gpu():
    """Return torch with a visible GPU, or raise (the minimizer only runs on the GPU)."""
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("the MI355X minimizer needs a ROCm GPU; none is visible")
    load()
    return torch


def stream_ptr(stream=None) -> int:
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return int(s.cuda_stream)
