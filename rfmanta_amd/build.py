"""Builds rfmanta_amd/libmh.so (the C-ABI library declared in include/mh.h) for gfx950.

HIP sources go through hipcc; the host C++ (SAH builder, C ABI) through g++ against the
HIP runtime headers. Every translation unit is compiled with -ffp-contract=off: the
kernels and the builder reproduce the reference's binary32 arithmetic operation for
operation, and a fused multiply-add would change results.

Usage: python -m rfmanta_amd.build [-v]
"""

from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
REPO = PKG.parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "libmh.so"
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
HIPCC = str(ROCM / "bin" / "hipcc")
ARCH = "gfx950"

COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall",
          f"-I{REPO / 'include'}", f"-I{CSRC}"]
HIP_FLAGS = [f"--offload-arch={ARCH}", "-fhip-fp32-correctly-rounded-divide-sqrt",
             "-fno-gpu-flush-denormals-to-zero", "-munsafe-fp-atomics",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
HOST_FLAGS = ["-D__HIP_PLATFORM_AMD__", f"-I{ROCM / 'include'}"]

HIP_SOURCES = ["trace.hip", "batch.hip", "batch_ks_pair.hip", "batch_ks_ref.hip",
               "batch_ks_soa.hip", "batch_wald_pair.hip", "batch_wald_ref.hip",
               "batch_wald_soa.hip", "refit.hip", "rotate.hip",
               "build.hip", "shade.hip", "genrays.hip", "util.hip", "noalloc.hip"]
HOST_SOURCES = ["sah_builder.cpp", "mh_api.cpp", "frame_api.cpp", "rays_api.cpp", "server.cpp",
                "ply.cpp", "obj.cpp"]


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if verbose and (r.stdout or r.stderr):
        print(r.stdout, r.stderr, flush=True)


def build(verbose: bool = False, save_temps: bool = False) -> Path:
    """Compiles every source and links libmh.so; returns its path."""
    OBJ.mkdir(exist_ok=True)
    jobs = []
    for src in HIP_SOURCES:
        obj = OBJ / (src + ".o")
        extra = ["-save-temps=obj"] if save_temps else []
        jobs.append([HIPCC, *COMMON, *HIP_FLAGS, *extra, "-c", str(CSRC / src), "-o", str(obj)])
    for src in HOST_SOURCES:
        obj = OBJ / (src + ".o")
        jobs.append(["g++", *COMMON, *HOST_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        list(ex.map(lambda c: _run(c, verbose), jobs))
    objs = [str(OBJ / (s + ".o")) for s in HIP_SOURCES + HOST_SOURCES]
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", str(tmp), *objs,
          "-lpthread"], verbose)
    os.replace(tmp, LIB)
    build_manta_harness(verbose)
    return LIB


MANTA = PKG / "manta"
HARNESS = MANTA / "libmanta_harness.so"


def build_manta_harness(verbose: bool = False) -> Path:
    """HipDynBVH (the Manta AccelerationStructure over the C ABI) and its driver."""
    inc = [f"-I{MANTA / 'compat'}", f"-I{MANTA}", f"-I{REPO / 'include'}"]
    tmp = HARNESS.with_suffix(".so.tmp")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", *inc,
          str(MANTA / "HipDynBVH.cc"), str(MANTA / "manta_harness.cc"), "-o", str(tmp),
          f"-L{PKG}", "-l:libmh.so", "-Wl,-rpath,$ORIGIN/..", "-pthread"], verbose)
    os.replace(tmp, HARNESS)
    return HARNESS


if __name__ == "__main__":
    build(verbose="-v" in sys.argv, save_temps="--save-temps" in sys.argv)
    print(LIB)
