"""Build lj_gpu_amd/libljgpu.so (HIP for gfx950 + host C++), in-tree.

    python -m lj_gpu_amd.build          # or: python lj_gpu_amd/build.py

HIP sources are compiled by hipcc for gfx950 with packed FP32 instructions
disabled; host.cpp by g++ with -ffp-contract=off (the neighbour-list
predicate must round exactly as the reference's makepair); the objects are
linked by hipcc into one shared library with a C ABI (include/ljgpu.h).
The benchmark program lj_gpu_amd/force_ljgpu (harness/force_ljgpu.cpp, the
reference's cuda/force_cuda.cu main over that ABI) is linked against it.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libljgpu.so")
HARNESS = os.path.join(PKG, "force_ljgpu")

ARCH = os.environ.get("LJGPU_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

HIP_SOURCES = ["api.hip", "lists.hip", "md.hip", "tune.hip", "force_f64.hip", "force_mixed.hip",
               "force_kahan.hip", "force_aos3_f64.hip", "force_aos3_mixed.hip",
               "force_aos3_kahan.hip"]
HOST_SOURCES = ["host.cpp", "paircache.cpp"]

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # Packed FP32 (v_pk_*_f32) stays off in every HIP translation unit.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # FP64 atomic adds compile to the hardware instruction (no CAS loop).
    "-munsafe-fp-atomics",
    "-Wall", "-Wno-unused-command-line-argument",
]
HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-pthread", "-Wall"]
# The benchmark program over the C ABI (harness/force_ljgpu.cpp): a plain
# g++ link against the library, found next to the executable at run time.
HARNESS_FLAGS = ["-O2", "-std=c++17", "-Wall", f"-I{os.path.join(ROOT, 'include')}"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout)
    # hipcc's host pass warns about the GPU-only target feature; keep other output.
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if out.strip():
        print(out)


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    jobs = []
    objs = []
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        jobs.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in HOST_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        jobs.append([CXX, *HOST_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for cmd in jobs:
            if verbose:
                print(" ".join(cmd))
        list(ex.map(_run, jobs))
    tmp = LIB + ".tmp"
    link = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs, "-pthread"]
    if verbose:
        print(" ".join(link))
    _run(link)
    os.replace(tmp, LIB)
    exe = [CXX, *HARNESS_FLAGS, os.path.join(PKG, "harness", "force_ljgpu.cpp"), "-o",
           HARNESS + ".tmp", f"-L{PKG}", "-lljgpu", "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(exe))
    _run(exe)
    os.replace(HARNESS + ".tmp", HARNESS)
    return LIB


if __name__ == "__main__":
    path = build(verbose="-v" in sys.argv)
    print(path)
