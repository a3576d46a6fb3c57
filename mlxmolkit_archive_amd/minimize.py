"""``mmff_minimize_nk``: drop-in for mlxmolkit/mmff_minimize.py:231-346 on MI355X.

Same signature, defaults and return layout ``(out_pos, energies, converged)`` with
``converged = status == 0`` (status 0 conflates gradient convergence, TOLX and the ascent
bail-out exactly as the reference kernel does). Dense BFGS by default, L-BFGS with the
reference's history length 8 when ``use_lbfgs=True``. The whole minimization is one
persistent kernel launch (mmffhip_minimize_nk); keyword-only extras expose the MI355X
plan (Hessian placement, workgroup size, line-search gradient fusion, precision-exact
FP64 energies) without changing the reference behaviour by default.
"""
from __future__ import annotations

from typing import List

import numpy as np

from . import _lib
from .engine import NkRunner, device_batch

LBFGS_M = 8  # history length of the reference nk L-BFGS kernel (mlxmolkit/mmff_minimize.py:31)


def mmff_minimize_nk(mmff_params_list: List, conf_counts: List[int], positions_3d: np.ndarray, *,
                     max_iters: int = 500, grad_tol: float = 1e-4, use_lbfgs: bool = False,
                     lbfgs_m: int = LBFGS_M, diel: int = 1, hess_place: int = _lib.HESS_AUTO, block: int = 0,
                     ls_grad: bool = False, return_info: bool = False):
    """Run MMFF94 optimization entirely on the GPU.

    ``positions_3d`` may also be a float32 CUDA tensor (used in place on the device).
    Returns ``(out_pos float32 [sum 3N], energies float32 [C], converged bool [C])``; with
    ``return_info=True`` a fourth element holds FP64 energies, raw statuses and per-conformer
    counters {iterations, gradient evaluations, energy evaluations}.
    """
    torch = _lib.require_gpu()
    db = device_batch(mmff_params_list, conf_counts)
    if isinstance(positions_3d, torch.Tensor) and positions_3d.is_cuda:
        # start geometries already on the device (e.g. from embed_device): no host round trip
        x = positions_3d.reshape(-1).to(torch.float32).contiguous()
    else:
        x = torch.from_numpy(np.ascontiguousarray(np.asarray(positions_3d, dtype=np.float32).reshape(-1))).cuda()
    if x.numel() != db.host.total_coords:
        raise ValueError(f"positions_3d has {x.numel()} values, the batch needs {db.host.total_coords}")
    r = NkRunner(db, max_iters=max_iters, grad_tol=grad_tol, use_lbfgs=use_lbfgs, lbfgs_m=lbfgs_m, diel=diel,
                 hess_place=hess_place, block=block, ls_grad=ls_grad)
    res = r.run(x)
    torch.cuda.synchronize()
    out_pos = res.x.cpu().numpy()
    e64 = res.energy.cpu().numpy()
    st = res.status.cpu().numpy()
    conv = st == 0
    if return_info:
        return out_pos, e64.astype(np.float32), conv, dict(energy_f64=e64, status=st,
                                                           counters=res.counters.cpu().numpy())
    return out_pos, e64.astype(np.float32), conv
