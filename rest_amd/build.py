"""Build the MI355X J/K library (rest_amd/lib/librest_jk.so) with hipcc for gfx950.

Usage: python -m rest_amd.build   (or rest_amd.build.build())
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
OUT = PKG / "lib"
LIB = OUT / "librest_jk.so"
OBJ = OUT / "obj"

ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# Packed FP32 VALU ops are switched off for every device pass.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Wall", "-Wno-unused-function",
          "-I", str(PKG.parent / "include")]
SOURCES = ["rjk_api.cpp", "rjk_j.hip", "rjk_k.hip", "rjk_util.hip", "rsi_int.cpp", "rsi_gpu3c.hip", "rsi_build.hip",
           "rscf.hip", "rlg.hip", "rjk_batch.hip", "rjk_kbf.hip"]
# MFMA accumulators in VGPRs: the K tiles carry 64-128 accumulator registers
# across a software-pipelined loop; the default AGPR form makes the allocator
# copy them VGPR<->AGPR every iteration.
PER_FILE = {"rjk_k.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"], "rjk_batch.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"],
            "rjk_kbf.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"],
            "rsi_int.cpp": ["-fopenmp"]}
LIBS = ["-fopenmp", "-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib", "-lrocsolver", "-lrocblas"]


def _compile(src: Path) -> Path:
    obj = OBJ / (src.name + ".o")
    cmd = [HIPCC, *COMMON]
    if src.suffix == ".hip":
        cmd += ["-x", "hip", *DEVICE_FLAGS, *PER_FILE.get(src.name, [])]
    cmd += ["-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = False) -> Path:
    """Compile every source and link the shared library; returns its path."""
    if not Path(HIPCC).exists() and shutil.which("hipcc") is None:
        raise RuntimeError("hipcc not found")
    OBJ.mkdir(parents=True, exist_ok=True)
    srcs = [CSRC / s for s in SOURCES]
    with cf.ThreadPoolExecutor(max_workers=min(6, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", str(tmp), *map(str, objs), *LIBS]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
