"""Single-point MMFF energy/gradient API (drop-in for mlxmolkit/mmff_metal_kernel.py:512-775).

``pack_params_for_metal`` / ``pack_multi_mol_for_metal`` return device batches in this
build's ragged layout (see packing.py) instead of the reference's flat MLX buffers, and
``mmff_energy_grad_metal`` / ``mmff_energy_grad_metal_mega`` evaluate the mega formula
family (mmff_metal_kernel.py:274-505) on the GPU. ``mmff_energy_grad`` additionally
exposes the run_mmff (TG) family, per-term energies and a term mask.
"""
from __future__ import annotations

from typing import List, Sequence

import numpy as np

from . import _lib
from .engine import device_batch, energy_grad


def pack_params_for_metal(p, n_confs: int = 1):
    """Device batch for one molecule's conformers."""
    return device_batch([p], [n_confs])


def pack_multi_mol_for_metal(params_list: Sequence, conf_counts: Sequence[int]):
    """Device batch for many molecules (ragged, unpadded coordinates)."""
    return device_batch(params_list, conf_counts)


def mmff_energy_grad(params_list: Sequence, conf_counts: Sequence[int], positions, *, family: str = "mega",
                     diel: int = 1, term_mask: int = 0x7F, per_term: bool = False):
    """Energies (float64 [C]) and gradients (float32, same layout as positions) on the GPU.

    family "mega" follows mmff_metal_kernel.py:274-505; "tg" the run_mmff header
    (mmff_bfgs_header.metal:14-402). Returns numpy arrays.
    """
    torch = _lib.require_gpu()
    db = device_batch(params_list, conf_counts)
    x = torch.from_numpy(np.ascontiguousarray(np.asarray(positions, np.float32).reshape(-1))).cuda()
    fam = {"mega": _lib.FAMILY_MEGA, "tg": _lib.FAMILY_TG}[family]
    e, g, et = energy_grad(db, x, fam, diel, term_mask, grad=True, terms=per_term)
    torch.cuda.synchronize()
    out = (e.cpu().numpy(), g.cpu().numpy())
    if per_term:
        out = out + (et.cpu().numpy(),)
    return out


def mmff_energy_grad_metal(p, positions: np.ndarray):
    """(E [C], G [C, 3N]) for one molecule's conformers, positions (C, N, 3)."""
    pos = np.asarray(positions, np.float32)
    C = pos.shape[0]
    e, g = mmff_energy_grad([p], [C], pos.reshape(-1), family="mega")
    return e.astype(np.float32), g.reshape(C, -1)


def mmff_energy_grad_metal_mega(params_list: Sequence, conf_counts: Sequence[int], pos_flat: np.ndarray):
    """(E [C], G flat) for a ragged multi-molecule batch (mega formula family)."""
    e, g = mmff_energy_grad(params_list, conf_counts, pos_flat, family="mega")
    return e.astype(np.float32), g
