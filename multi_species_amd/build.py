"""Builds libfarrsight_hip.so (HIP kernels + C ABI + C++ host side) for gfx950.

Usage: ``python -m multi_species_amd.build``. Every call recompiles every
translation unit (one hipcc process per file, in parallel) and relinks.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libfarrsight_hip.so")
ARCH = os.environ.get("FS_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 instructions (v_pk_*_f32) stay off: some of their forms have
    # given wrong results on MI355X without any error
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unknown-pragmas",
    "-I", os.path.join(HERE, "..", "include"),
]


def sources():
    out = []
    for name in sorted(os.listdir(CSRC)):
        if name.endswith((".hip", ".cpp")):
            out.append(os.path.join(CSRC, name))
    return out


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC, *COMMON, "-c", src, "-o", obj]
    if src.endswith(".cpp"):
        cmd = [HIPCC, *COMMON, "-x", "hip", "-c", src, "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    # hipcc's host pass warns about the packed-fp32 feature it does not know
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = sources()
    jobs = min(len(srcs), max(1, min(8, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


SHIM_SRC = os.path.join(HERE, "shim", "fs_field_shim.cpp")
SHIM_LIB = os.path.join(HERE, "shim", "libfarrsight_field_shim.so")


def build_shim(reference: str, verbose: bool = False) -> str:
    """The reference-side ElectricField binding (shim/fs_field_shim.cpp),
    compiled against the reference's own inc/FieldStructure.hpp where it lies
    and linked to libfarrsight_hip.so. The base class's out-of-line members
    stay undefined here: the reference's compiled FieldStructure.cpp provides
    them in the program that loads the binding."""
    inc = os.path.join(reference, "inc")
    if not os.path.exists(os.path.join(inc, "FieldStructure.hpp")):
        raise FileNotFoundError(f"{inc}/FieldStructure.hpp not found")
    tmp = SHIM_LIB + ".tmp"
    cmd = ["g++", "-O2", "-std=c++14", "-fPIC", "-shared", "-w", "-I", inc,
           "-I", os.path.join(HERE, "..", "include"), "-o", tmp, SHIM_SRC,
           "-L", HERE, "-lfarrsight_hip", "-Wl,-rpath,$ORIGIN/.."]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"shim build failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, SHIM_LIB)
    if verbose:
        print(f"built {SHIM_LIB}")
    return SHIM_LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
