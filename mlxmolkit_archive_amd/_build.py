"""In-tree build of the package's native libraries.

* ``libmmff_hip.so``   hand-written HIP kernels for gfx950 + the extern "C" ABI
                       (csrc/mmff_hip.hip, csrc/precond.hip, csrc/embed.hip; declared in
                       include/mmff_hip.h)
* ``libmmff_embed.so`` host-side start-geometry embedder (csrc/embed.c)

Built in place next to this file so that the libraries travel with the source tree.
"""
from __future__ import annotations

import os
import shutil
import subprocess
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-ffp-contract=fast",
    # FP32 divide and sqrt at <= 2.5 ulp (v_rcp / v_sqrt) instead of the IEEE-exact
    # expansions; the term arithmetic is FP32 with a 1e-5 relative parity budget
    "-fno-hip-fp32-correctly-rounded-divide-sqrt",
]


def lib_path(name: str) -> Path:
    return PKG / name


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(map(str, cmd)) + "\n" + r.stdout + r.stderr)
    return r


def build_embed() -> Path:
    out = lib_path("libmmff_embed.so")
    tmp = out.with_suffix(".so.tmp")
    _run(["gcc", "-O2", "-fPIC", "-shared", "-pthread", "-std=c11", str(CSRC / "embed.c"), "-o", str(tmp), "-lm"])
    os.replace(tmp, out)
    return out


def hipcc() -> str:
    for c in ("/opt/rocm/bin/hipcc", shutil.which("hipcc") or ""):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def build_hip() -> Path:
    out = lib_path("libmmff_hip.so")
    tmp = out.with_suffix(".so.tmp")
    _run([hipcc(), *HIP_FLAGS, "-I", str(INCLUDE), "-I", str(CSRC), str(CSRC / "mmff_hip.hip"),
          str(CSRC / "precond.hip"), str(CSRC / "embed.hip"), "-o", str(tmp)])
    os.replace(tmp, out)
    return out


def build_all() -> None:
    build_embed()
    build_hip()
