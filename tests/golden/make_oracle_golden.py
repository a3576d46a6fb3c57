"""Writes tests/golden/oracle_small.npz: FP64 oracle vectors on a small seeded batch.

The batch (term tables + start coordinates) is stored in the file, so the fixture does
not depend on the typer. Outputs: TG/MEGA energies, TG per-term energies and gradients,
nk BFGS and L-BFGS results, BB fused (60 iterations) and multi_mol. Regenerate only when
the oracle's documented behaviour changes; tests/test_golden.py checks the oracle (CPU)
and the HIP path (GPU) against it.
"""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))

import oracle  # noqa: E402
from mlxmolkit_archive_amd import workloads  # noqa: E402

FIELDS = ["bond_idx1", "bond_idx2", "bond_kb", "bond_r0", "angle_idx1", "angle_idx2", "angle_idx3", "angle_ka",
          "angle_theta0", "angle_is_linear", "strbend_idx1", "strbend_idx2", "strbend_idx3", "strbend_kba_ijk",
          "strbend_kba_kji", "strbend_r0_ij", "strbend_r0_kj", "strbend_theta0", "oop_idx1", "oop_idx2", "oop_idx3",
          "oop_idx4", "oop_koop", "torsion_idx1", "torsion_idx2", "torsion_idx3", "torsion_idx4", "torsion_V1",
          "torsion_V2", "torsion_V3", "vdw_idx1", "vdw_idx2", "vdw_R_star", "vdw_eps", "ele_idx1", "ele_idx2",
          "ele_charge_term", "ele_is_1_4"]


def main():
    wl = workloads.build("small", n_mols=6, n_confs=2, seed=777)
    out = {"n_mols": np.int32(len(wl.params)), "conf_counts": np.asarray(wl.conf_counts, np.int32),
           "x": wl.x.astype(np.float32)}
    for m, p in enumerate(wl.params):
        out[f"m{m}_n_atoms"] = np.int32(p.n_atoms)
        for f in FIELDS:
            out[f"m{m}_{f}"] = np.asarray(getattr(p, f))
    ob = oracle.OracleBatch(wl.params, wl.conf_counts)
    x = wl.x.astype(np.float64)
    e_tg, g_tg, et_tg = oracle.energy_grad_batch(ob, x, family=oracle.FAMILY_TG)
    e_mg, g_mg, et_mg = oracle.energy_grad_batch(ob, x, family=oracle.FAMILY_MEGA)
    out.update(e_tg=e_tg, g_tg=g_tg, et_tg=et_tg, e_mg=e_mg, g_mg=g_mg, et_mg=et_mg)
    xb, eb, sb, cb = oracle.minimize_nk(ob, x, max_iters=2000, grad_tol=1e-4)
    xl, el, sl, cl = oracle.minimize_nk(ob, x, max_iters=2000, grad_tol=1e-4, use_lbfgs=True, m_hist=8)
    x60, e60 = oracle.bb_fused(ob, x, max_iters=60)
    xm, em, gm, pm, cm, um = oracle.multi_mol(ob, x, max_iters=200)
    out.update(bfgs_x=xb, bfgs_e=eb, bfgs_status=sb, bfgs_counters=cb, lbfgs_x=xl, lbfgs_e=el, lbfgs_status=sl,
               bb60_x=x60, bb60_e=e60, mm_x=xm, mm_e=em, mm_gnorm=gm, mm_pci=pm, mm_conv=cm, mm_used=np.int32(um))
    path = Path(__file__).resolve().parent / "oracle_small.npz"
    np.savez_compressed(path, **out)
    print(f"wrote {path}")


if __name__ == "__main__":
    main()
