"""Loader for tests/golden/oracle_small.npz (term tables + oracle outputs)."""
from __future__ import annotations

from pathlib import Path

import numpy as np

from mlxmolkit_archive_amd.mmff_params import MMFFParams

GOLDEN = Path(__file__).resolve().parent / "golden"
FIELDS = ["bond_idx1", "bond_idx2", "bond_kb", "bond_r0", "angle_idx1", "angle_idx2", "angle_idx3", "angle_ka",
          "angle_theta0", "angle_is_linear", "strbend_idx1", "strbend_idx2", "strbend_idx3", "strbend_kba_ijk",
          "strbend_kba_kji", "strbend_r0_ij", "strbend_r0_kj", "strbend_theta0", "oop_idx1", "oop_idx2", "oop_idx3",
          "oop_idx4", "oop_koop", "torsion_idx1", "torsion_idx2", "torsion_idx3", "torsion_idx4", "torsion_V1",
          "torsion_V2", "torsion_V3", "vdw_idx1", "vdw_idx2", "vdw_R_star", "vdw_eps", "ele_idx1", "ele_idx2",
          "ele_charge_term", "ele_is_1_4"]


def load_small():
    z = np.load(GOLDEN / "oracle_small.npz", allow_pickle=False)
    params = []
    for m in range(int(z["n_mols"])):
        p = MMFFParams(n_atoms=int(z[f"m{m}_n_atoms"]))
        for f in FIELDS:
            setattr(p, f, z[f"m{m}_{f}"])
        params.append(p)
    return params, [int(c) for c in z["conf_counts"]], z
