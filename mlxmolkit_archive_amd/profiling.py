"""Profiling helpers: the plain-copy calibration run for rocprofv3 HBM counters.

``python -m mlxmolkit_archive_amd.profiling copy [MiB]`` copies a device buffer of known
size a few times; under ``rocprofv3 --pmc FETCH_SIZE`` (or ``WRITE_SIZE``) the copy
kernel's counter value per dispatch divided by the known byte count gives the scale
factor applied to the minimizer's counter values.
"""
from __future__ import annotations

import sys


def copy_calibration(mib: int = 1024, reps: int = 3) -> None:
    import torch
    n = mib * 1024 * 1024 // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(f"copied {reps} x {mib} MiB ({reps * n * 4} bytes read, same written)", flush=True)


if __name__ == "__main__":
    if len(sys.argv) >= 2 and sys.argv[1] == "copy":
        copy_calibration(int(sys.argv[2]) if len(sys.argv) > 2 else 1024)
    else:
        print("usage: python -m mlxmolkit_archive_amd.profiling copy [MiB]", file=sys.stderr)
        sys.exit(2)
