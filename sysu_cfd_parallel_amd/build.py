"""Build recipe of the HIP library: hipcc for gfx950, objects and the shared
library under sysu_cfd_parallel_amd/_build (in-tree, so the built .so travels
with the source tree).  Run as ``python -m sysu_cfd_parallel_amd.build``."""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OUT = PKG / "_build"
LIB = OUT / "libsimple355.so"

SOURCES = [
    "mesh/case_gen.cpp",
    "mesh/cases.cpp",
    "mesh/ingest.cpp",
    "mesh/batch.cpp",
    "mesh/gmsh.cpp",
    "mesh/reorder.cpp",
    "device/fv_kernels.hip",
    "device/keps_kernels.hip",
    "device/transient_kernels.hip",
    "device/krylov.hip",
    "device/geometry_kernels.hip",
    "device/scratch.hip",
    "device/context.hip",
    "host/navier.cpp",
    "host/output.cpp",
    "host/export.cpp",
    "capi.cpp",
]

ARCH = "gfx950"

# -ffp-contract=off keeps every expression's rounding equal to the oracle's;
# packed FP32 instructions are disabled for all device code.
COMMON = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
]


def hipcc() -> str:
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(exe).exists():
        raise RuntimeError("hipcc not found")
    return exe


def _compile(src: str) -> Path:
    s = CSRC / src
    o = OUT / (src.replace("/", "__") + ".o")
    cmd = [hipcc(), *COMMON, "-c", str(s), "-o", str(o)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return o


def build(verbose: bool = False) -> Path:
    OUT.mkdir(parents=True, exist_ok=True)
    workers = min(len(SOURCES), max(1, (os.cpu_count() or 2)))
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = OUT / "libsimple355.so.tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
