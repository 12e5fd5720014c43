"""Build libneko_amd.so in-tree: hipcc for the gfx950 kernels, g++ for the
host-only setup code, one shared library with a plain C ABI
(include/neko_amd.h).  No torch, no JIT cache: the .so sits next to this
file so it travels with the source tree."""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libneko_amd.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-I", CSRC, "-I", os.path.join(ROOT, "include"),
]
CXX_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-pthread", "-ffp-contract=off",
             "-I", CSRC, "-I", os.path.join(ROOT, "include"), "-I", "/opt/rocm/include",
             "-D__HIP_PLATFORM_AMD__"]

HIP_SOURCES = ["nm_kernels.hip", "nm_op.hip", "nm_cg.hip", "nm_abi.hip", "nm_hsmg.hip",
               "nm_gmres.hip"]
CXX_SOURCES = ["nm_setup.cpp", "nm_fdm_setup.cpp", "nm_mesh.cpp", "nm_curve.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def extra_hip_sources():
    """Every other *.hip in csrc is part of the library as well."""
    have = set(HIP_SOURCES)
    return sorted(f for f in os.listdir(CSRC) if f.endswith(".hip") and f not in have)


def build(verbose: bool = False) -> str:
    os.makedirs(OUT, exist_ok=True)
    jobs = []
    for src in HIP_SOURCES + extra_hip_sources():
        obj = os.path.join(OUT, src + ".o")
        jobs.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    for src in CXX_SOURCES:
        obj = os.path.join(OUT, src + ".o")
        jobs.append(["g++", *CXX_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    workers = min(len(jobs), max(1, min(8, os.cpu_count() or 1)))
    with cf.ThreadPoolExecutor(workers) as ex:
        list(ex.map(_run, jobs))
    objs = [j[-1] for j in jobs]
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs,
          "-lpthread"])
    shutil.move(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
