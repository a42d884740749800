"""Compiles libnbots_amd.so (HIP for gfx950) and libnbx_objects.so (host
C++, the reference's FEM input objects for callers without libnbots)
in-tree.

    python -m nbots_amd.build

libnbots_amd.so leaves the object accessors of include/nbots_amd.h
(NBX_IMPORTS_BEGIN .. NBX_IMPORTS_END) undefined: they resolve to libnbots
in a drop-in link, or to libnbx_objects.so loaded before it.

Flags:
  -ffp-contract=off   the reference's arithmetic has no FMA; assembly, SpMV
                      and BC kernels reproduce it bit for bit
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      no packed FP32 instructions
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libnbots_amd.so")
OBJECTS_SRC = os.path.join(HERE, "objects", "objects.cpp")
OBJECTS_LIB = os.path.join(HERE, "libnbx_objects.so")
CXX = os.environ.get("CXX", "g++")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

CFLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wall", "-Wno-unused-result", f"-I{os.path.join(ROOT, 'include')}"]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC)
                  if f.endswith((".hip", ".cpp")))


def _compile(src: str) -> str:
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [HIPCC, *CFLAGS, "-c", src, "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC, *CFLAGS, "-x", "hip", "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr[-6000:]}")
    return obj


def build_objects(verbose: bool = True) -> str:
    tmp = OBJECTS_LIB + ".tmp"
    cmd = [CXX, "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-Wextra",
           f"-I{os.path.join(ROOT, 'include')}", OBJECTS_SRC, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{CXX} failed on {OBJECTS_SRC}:\n{r.stderr[-6000:]}")
    os.replace(tmp, OBJECTS_LIB)
    if verbose:
        print(f"built {OBJECTS_LIB}")
    return OBJECTS_LIB


def build(verbose: bool = True) -> str:
    build_objects(verbose)
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-6000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    try:
        build()
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
