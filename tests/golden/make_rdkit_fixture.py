"""Writes tests/golden/rdkit_relaxed.json from data files the reference's own tests hold.

Sources (read as data, never imported or executed):
  tests/_mopac_ref_generated.py  -- dict literal of SMILES + Z + coordinates after RDKit
                                    ETKDG(seed 1) + MMFFOptimizeMolecule (parsed with
                                    ast.literal_eval)
  tools/perf_series/geom.npz     -- Z + coordinates after MMFFOptimizeMolecule(maxIters=2000)
                                    (numpy.load, allow_pickle=False); SMILES from
                                    tools/perf_series/make_geom.py:7-12
Only molecules made of H/C/N/O/F/S/Cl are kept. Run once in the build container.
"""
import ast
import json
from pathlib import Path

import numpy as np

REF = Path("/root/reference")
OUT = Path(__file__).resolve().parent / "rdkit_relaxed.json"
ALLOWED = {1, 6, 7, 8, 9, 16, 17}


def main():
    txt = (REF / "tests/_mopac_ref_generated.py").read_text()
    tree = ast.parse(txt)
    data = None
    for node in tree.body:
        if isinstance(node, ast.Assign) and getattr(node.targets[0], "id", "") == "MOPAC_REF":
            # dict(...) calls -> convert to a literal by evaluating keyword values only
            out = {}
            for k, v in zip(node.value.keys, node.value.values):
                entry = {kw.arg: ast.literal_eval(kw.value) for kw in v.keywords}
                out[ast.literal_eval(k)] = entry
            data = out
    fixture = {}
    for name, d in data.items():
        if set(d["atoms"]) <= ALLOWED:
            fixture[name] = {"smiles": d["smiles"], "Z": list(d["atoms"]), "coords": [list(c) for c in d["coords"]],
                             "source": "tests/_mopac_ref_generated.py (RDKit ETKDGv3 seed 1 + MMFF94)"}
    g = np.load(REF / "tools/perf_series/geom.npz", allow_pickle=False)
    smiles = {"cholesterol": "CC(C)CCCC(C)C1CCC2C1(CCC3C2CC=C4C3(CCC(C4)O)C)C", "thioanisole": "CSc1ccccc1",
              "aspirin": "CC(=O)OC1=CC=CC=C1C(=O)O", "testosterone": "CC12CCC3C(CCC4=CC(=O)CCC34C)C1CCC2O"}
    for name, smi in smiles.items():
        Z = g[f"{name}_Z"].tolist()
        if set(Z) <= ALLOWED:
            fixture[name] = {"smiles": smi, "Z": Z, "coords": g[f"{name}_R"].round(6).tolist(),
                             "source": "tools/perf_series/geom.npz (RDKit MMFFOptimizeMolecule maxIters=2000)"}
    OUT.write_text(json.dumps(fixture, indent=1))
    print(f"wrote {OUT} with {len(fixture)} molecules")


if __name__ == "__main__":
    main()
