"""The C ABI library loads and exports every symbol include/mmff_hip.h declares (no
compute calls: this runs without a GPU)."""
from __future__ import annotations

import ctypes
import re
from pathlib import Path

from mlxmolkit_archive_amd import _lib
from mlxmolkit_archive_amd._build import lib_path

HEADER = Path(__file__).resolve().parents[1] / "include" / "mmff_hip.h"


def declared_symbols():
    txt = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(mmffhip_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(str(lib_path("libmmff_hip.so")))
    syms = declared_symbols()
    assert len(syms) >= 9
    for s in syms:
        assert hasattr(L, s), s
    assert set(syms) == set(_lib.SIGNATURES)  # the ctypes binding covers the whole ABI


def test_abi_version_and_struct_sizes():
    L = _lib.load()
    assert L.mmffhip_abi_version() == _lib.ABI_VERSION
    assert ctypes.sizeof(_lib.Batch) == 16 * 8 + 3 * 4 + 4  # 16 pointers + 3 int32 (+ padding)
    assert ctypes.sizeof(_lib.NkOpts) == 8 * 4 + 8 + 8 + 8  # ... precond, table_lds_bytes (+ pad), xcd_offsets


def test_oracle_is_not_linked_by_the_product():
    """The product library must not depend on the oracle in any form."""
    data = lib_path("libmmff_hip.so").read_bytes()
    assert b"oracle" not in data
