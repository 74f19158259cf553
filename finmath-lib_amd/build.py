"""Build the MI355X LMM engine: hipcc for gfx950, one shared library with a C ABI.

    python finmath-lib_amd/build.py            # builds finmath-lib_amd/lib/libfinmath_lmm.so

Every HIP translation unit is compiled with packed-FP32 instruction selection
disabled. The FP64 path kernel and the generator are compiled without fused
multiply-add contraction, so each product and sum rounds like the JVM reference.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "build")
LIB = os.path.join(HERE, "lib")
TARGET = os.path.join(LIB, "libfinmath_lmm.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = [
    "-std=c++17", "-O3", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(REPO, "include"), "-I", CSRC,
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]
NO_CONTRACT = ["-ffp-contract=off"]

# source -> extra flags
SOURCES = {
    "mt19937_gf2.cpp": NO_CONTRACT,
    "mt19937.hip": NO_CONTRACT,
    "lmm_step_fp64.hip": NO_CONTRACT,
    "lmm_step_fp32.hip": [],
    "lmm_step_fp64_logstate.hip": [],
    "lmm_step_general_fp64.hip": NO_CONTRACT,
    "lmm_step_pc_fp64.hip": NO_CONTRACT,
    "lmm_step_general_fp32.hip": [],
    "lmm_step_pc_fp32.hip": [],
    "lmm_step_lv_fp64.hip": NO_CONTRACT,
    "lmm_step_lv_fp32.hip": [],
    "lmm_stochvol.hip": NO_CONTRACT,
    "lmm_reduce.hip": NO_CONTRACT,
    "lmm_bermudan.hip": NO_CONTRACT,
    "rv.hip": NO_CONTRACT,
    "lmm_noalloc.hip": NO_CONTRACT,
    "capi.cpp": NO_CONTRACT,
}


def _compile(src: str, extra: list[str]) -> str:
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    lang = ["-x", "hip"] if src.endswith(".hip") else []
    cmd = [HIPCC, *COMMON, *extra, *lang, "-c", os.path.join(CSRC, src), "-o", out]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return out


def build(jobs: int | None = None, verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB, exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 4)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda kv: _compile(*kv), SOURCES.items()))
    tmp = TARGET + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, TARGET)
    if verbose:
        print(f"built {TARGET}")
    host = build_host()
    if verbose:
        print(f"built {host}")
    return TARGET


def build_host() -> str:
    """Host-side mirror of the reference API (plain C++ over the C ABI) + pybind11 module."""
    import sysconfig
    import pybind11
    out = os.path.join(HERE, "finmath_hip", "_finmath_host" + sysconfig.get_config_var("EXT_SUFFIX"))
    tmp = out + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-fvisibility=hidden",
           "-I", os.path.join(REPO, "include"), "-I", os.path.join(HERE, "host"),
           "-I", pybind11.get_include(), "-I", sysconfig.get_paths()["include"],
           os.path.join(HERE, "host", "finmath_host.cpp"), os.path.join(HERE, "host", "bindings.cpp"),
           "-L", LIB, "-lfinmath_lmm", "-Wl,-rpath,$ORIGIN/../lib", "-o", tmp]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"host build failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
