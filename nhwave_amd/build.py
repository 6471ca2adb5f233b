"""Builds libnhwave_amd.so (HIP, gfx950) in-tree with hipcc.

FP64 bit-parity with the reference needs the reference's operation sequence:
`-ffp-contract=off` keeps every multiply and add separately rounded (no FMA
contraction), and no fast-math flag is ever passed.  Packed-FP32 instruction
selection is disabled for all device code.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIB = PKG / "libnhwave_amd.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("NHW_OFFLOAD_ARCH", "gfx950")

SOURCES = ["nhw_capi.hip", "nhw_explicit.hip", "nhw_poisson.hip", "nhw_mg.hip", "nhw_batch.hip", "nhw_timing.hip",
           "nhw_output.hip", "nhw_forcing.hip", "nhw_viscous.hip", "nhw_noalloc.hip"]
FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-variable",
]


def build(verbose: bool = False, jobs: int = 0) -> Path:
    """Compiles every translation unit (in parallel, `jobs` at a time; 0: one
    per CPU up to 8) into _build/, then links the shared library.  Every call
    compiles every unit."""
    from concurrent.futures import ThreadPoolExecutor

    obj_dir = PKG / "_build"
    obj_dir.mkdir(exist_ok=True)
    cflags = [f for f in FLAGS if f != "-shared"]

    def compile_one(src: str):
        obj = obj_dir / (Path(src).stem + ".o")
        cmd = [HIPCC, *cflags, "-c", "-o", str(obj), str(CSRC / src)]
        if verbose:
            print(" ".join(cmd), flush=True)
        res = subprocess.run(cmd, capture_output=True, text=True)
        return src, obj, res

    jobs = jobs or min(8, os.cpu_count() or 1)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(compile_one, SOURCES))
    for src, _, res in results:
        if res.returncode != 0:
            sys.stderr.write(res.stdout + res.stderr)
            raise RuntimeError(f"hipcc failed ({res.returncode}) compiling {src}")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(LIB) + ".tmp",
           *[str(obj) for _, obj, _ in results]]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"hipcc failed ({res.returncode}) linking {LIB.name}")
    os.replace(str(LIB) + ".tmp", LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
