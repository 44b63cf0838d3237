"""Build the gfx950 HIP library libpsb_sph.so in-tree.

    python -m psb_project_amd.build          # or build_library() from Python

Every .hip file under csrc/ is compiled for gfx950 only, with FP contraction
off (the FP32 operation order is part of the numerical contract: fused
multiply-adds appear only where the source writes them) and without packed
FP32 instructions.  Objects go to psb_project_amd/_build/, the shared library
to psb_project_amd/libpsb_sph.so.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libpsb_sph.so"
INCLUDE = PKG.parent / "include"

ARCH = "gfx950"
HIPFLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-command-line-argument",
    f"-I{INCLUDE}",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm is required to build libpsb_sph.so)")


def _compile(src: Path, verbose: bool) -> Path:
    obj = BUILD / (src.stem + ".o")
    cmd = [hipcc(), *HIPFLAGS, "-c", str(src), "-o", str(obj)]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{r.stdout}\n{r.stderr}")
    warn = (r.stderr or "").strip()
    if verbose and warn:
        print(warn, flush=True)
    return obj


def build_library(verbose: bool = False, jobs: int | None = None) -> Path:
    """Compile every csrc/*.hip for gfx950 and link libpsb_sph.so (always rebuilds)."""
    BUILD.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip"))
    if not srcs:
        raise RuntimeError("no HIP sources found")
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


HOST = PKG / "host"
EXAMPLE = HOST / "example_dam_break"


def build_host_example(verbose: bool = False) -> Path:
    """Compile the C++ host-API example against libpsb_sph.so (rpath to it)."""
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", f"-I{INCLUDE}", f"-I{HOST}",
           str(HOST / "example_dam_break.cpp"), f"-L{PKG}", "-lpsb_sph",
           "-Wl,-rpath,$ORIGIN/..", "-o", str(EXAMPLE)]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"host example build failed:\n{r.stdout}\n{r.stderr}")
    return EXAMPLE


if __name__ == "__main__":
    path = build_library(verbose="-v" in sys.argv)
    build_host_example(verbose="-v" in sys.argv)
    print(path)
