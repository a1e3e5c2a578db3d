"""Build libfem_mi355x.so for gfx950 with hipcc (in-tree, no JIT cache).

    python sci-solver_fem_amd/build.py [--jobs N]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OUT = PKG / "lib"
LIB = OUT / "libfem_mi355x.so"
SOURCES = ["primitives.hip", "assembly.hip", "sparse.hip", "amg_setup.hip", "solve.hip",
           "mesh_io.cpp", "aggmis.cpp", "devpool.cpp", "capi.cpp"]
ARCH = os.environ.get("FEM_OFFLOAD_ARCH", "gfx950")


def hipcc() -> str:
    p = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    return p


def flags(src: str):
    f = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-function", "-Wno-unused-result",
         f"--offload-arch={ARCH}"]
    if src.endswith(".hip"):
        # keep packed FP32 instructions out of the device code
        f += ["-x", "hip", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
    return f


def compile_one(src: str) -> Path:
    obj = OUT / (src.replace(".", "_") + ".o")
    cmd = [hipcc()] + flags(src) + ["-I", str(CSRC), "-c", str(CSRC / src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(jobs: int = 8, verbose: bool = False) -> Path:
    OUT.mkdir(exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp)] + [str(o) for o in objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    for o in objs:
        o.unlink()
    build_examples()
    if verbose:
        print(f"built {LIB}")
    return LIB


def build_examples() -> None:
    """C++ drivers with the reference examples' command lines (link the C ABI)."""
    inc = PKG.parent / "include"
    for ex in ("example1", "example2"):
        cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I", str(inc), str(PKG / "examples" / f"{ex}.cpp"),
               "-L", str(OUT), "-lfem_mi355x", "-Wl,-rpath,$ORIGIN", "-o", str(OUT / ex)]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"example build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=min(8, os.cpu_count() or 1))
    a = ap.parse_args()
    try:
        build(a.jobs, verbose=True)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
