"""Compile the gfx950 HIP library (libpmgpu.so) in-tree with hipcc, and the
host front end (libpmfront.so: ProtoMol's input/output files) with g++.

Usage: python -m protomol_amd.build   (or protomol_amd.build.build()).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpmgpu.so")
SOURCES = ["pm_nonbonded.hip", "pm_pme.hip", "pm_md.hip", "pm_bonded.hip", "pm_krylov.hip", "pm_shake.hip",
           "pm_api.cpp"]
FRONT = os.path.join(HERE, "front")
FRONT_LIB = os.path.join(HERE, "libpmfront.so")
FRONT_SOURCES = ["pmf_read.cpp", "pmf_build.cpp", "pmf_write.cpp", "pmf_conf.cpp", "pmf_run.cpp", "pmf_api.cpp"]
PMRUN = os.path.join(HERE, "pmrun")
ROCM_LIB = os.environ.get("ROCM_LIB", "/opt/rocm/lib")
CXX = os.environ.get("CXX", "g++")
ARCH = os.environ.get("PM_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
         # packed FP32 instructions stay off (v_pk_* FP32 ops are not emitted)
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-I" + CSRC, "-I" + os.path.join(ROOT, "include")]


def _compile(src: str) -> str:
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS]
    if src.endswith(".cpp"):
        cmd += ["-x", "hip"]
    cmd += ["-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr[-4000:]}")
    return out


def build_front(verbose: bool = True) -> str:
    """Host-only library: readers, topology builder and writers (no HIP)."""
    tmp = FRONT_LIB + ".tmp"
    cmd = [CXX, "-O2", "-std=c++17", "-fPIC", "-Wall", "-shared", "-I" + FRONT, "-I" + os.path.join(ROOT, "include"),
           "-o", tmp] + [os.path.join(FRONT, f) for f in FRONT_SOURCES]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"front end build failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, FRONT_LIB)
    if verbose:
        print(f"built {FRONT_LIB}")
    return FRONT_LIB


def build(verbose: bool = True) -> str:
    build_front(verbose)
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    build_pmrun(verbose)
    return LIB


def build_pmrun(verbose: bool = True) -> str:
    """The .conf driver: links the front end and the device library."""
    tmp = PMRUN + ".tmp"
    cmd = [CXX, "-O2", "-std=c++17", "-Wall", "-I" + FRONT, "-I" + os.path.join(ROOT, "include"), "-o", tmp,
           os.path.join(FRONT, "pmrun.cpp"), "-L" + HERE, "-lpmfront", "-lpmgpu", "-L" + ROCM_LIB,
           "-Wl,-rpath,$ORIGIN", "-Wl,-rpath," + ROCM_LIB, "-Wl,--allow-shlib-undefined"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"pmrun link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, PMRUN)
    if verbose:
        print(f"built {PMRUN}")
    return PMRUN


if __name__ == "__main__":
    build()
    sys.exit(0)
