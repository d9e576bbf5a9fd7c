"""Build libhykkt.so (gfx950) in-tree: hipcc for the HIP kernels, g++ for the
host C++ (symbolic analysis, orchestration, C ABI), hipcc to link.

    python -m hykkt_amd.build          # or hykkt_amd.build.build()
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
SRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libhykkt.so")
BIN = os.path.join(PKG, "bin")
DRIVERS = ["solver_driver", "solver_emulator"]
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = "gfx950"

HIP_SOURCES = {
    # formation / Ruiz / SpMV / CG kernels: explicit fma only, so results are
    # bit-identical to the CPU oracle's operation order
    "kernels_basic.hip": ["-ffp-contract=off"],
    "kernels_chol.hip": [],
    "kernels_dense.hip": [],
    "kernels_trsv.hip": [],
    "noalloc.hip": [],
}
CXX_SOURCES = ["symbolic.cpp", "amd.cpp", "cholplan.cpp", "mmio.cpp", "chol_engine.cpp",
               "profiler.cpp",
               "cg.cpp", "solver.cpp", "capi.cpp"]

HIP_FLAGS = ["-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
             # keep packed-FP32 instructions out of every HIP object
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-result"]
CXX_FLAGS = ["-O2", "-fPIC", "-std=c++17", f"-I{ROCM}/include", "-D__HIP_PLATFORM_AMD__",
             "-Wall", "-Wno-unused-result"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return r


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    jobs = []
    for src, extra in HIP_SOURCES.items():
        obj = os.path.join(OBJ, src + ".o")
        jobs.append([HIPCC, *HIP_FLAGS, *extra, "-c", os.path.join(SRC, src), "-o", obj])
    for src in CXX_SOURCES:
        obj = os.path.join(OBJ, src + ".o")
        jobs.append(["g++", *CXX_FLAGS, "-c", os.path.join(SRC, src), "-o", obj])
    workers = max(1, min(len(jobs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4))))
    with cf.ThreadPoolExecutor(workers) as ex:
        for r, cmd in zip(ex.map(_run, jobs), jobs):
            if verbose:
                print(" ".join(cmd))
    objs = [j[-1] for j in jobs]
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    # the reference's example drivers, rebuilt on the C++ HykktSolver header
    os.makedirs(BIN, exist_ok=True)
    for name in DRIVERS:
        out = os.path.join(BIN, name)
        _run(["g++", "-O2", "-std=c++17", "-Wall", f"-I{os.path.join(ROOT, 'include')}",
              os.path.join(PKG, "drivers", name + ".cpp"), "-o", out + ".tmp", f"-L{PKG}",
              "-lhykkt", "-Wl,-rpath,$ORIGIN/..", f"-Wl,-rpath,{ROCM}/lib"])
        os.replace(out + ".tmp", out)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
