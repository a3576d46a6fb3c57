"""Shared pytest setup: the `gpu` marker, import paths and small seeded batches."""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
for p in (ROOT, ROOT / "oracle"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU and the built HIP library")


@pytest.fixture(scope="session")
def small_wl():
    from mlxmolkit_archive_amd import workloads
    return workloads.build("small", n_mols=8, n_confs=2)


@pytest.fixture(scope="session")
def medium_wl():
    from mlxmolkit_archive_amd import workloads
    return workloads.build("medium", n_mols=6, n_confs=2)


def mol_params(smiles: str, seed: int = 3, noise: float = 0.1, k: int = 1):
    """(params, flat start coordinates) of one SMILES molecule with k embedded conformers."""
    from mlxmolkit_archive_amd.embed import embed
    from mlxmolkit_archive_amd.mmff_params import extract_mmff_params
    from mlxmolkit_archive_amd.mol import mol_from_smiles
    m = mol_from_smiles(smiles)
    m.set_conformers(np.zeros((1, m.GetNumAtoms(), 3)))
    p = extract_mmff_params(m)
    x = embed([p], [m.atomic_numbers()], [k], seeds=np.arange(k) + seed, noise=noise)
    return m, p, x
