"""Build recipe for the gfx950 library (and the host-side input generators).

Produces, in-tree:
  blaze_amd/lib/libblaze_gpu.so   HIP kernels + the extern "C" ABI (include/blaze_gpu.h)
  blaze_amd/lib/libbzg_gen.so     host-only synthetic-input generators (gen.c)
  blaze_amd/lib/libblaze_seam.so  cblas_dgemm / cblas_dgemv / dgesv_ / dgetrf_ / dgetrs_ over the ABI

hipcc cross-compiles for gfx950 without a GPU, so this runs anywhere ROCm is installed.
Every HIP translation unit is compiled with packed FP32 instructions disabled.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "lib"
OBJ = ROOT / "_build" / "obj"
INCLUDE = ROOT / "include"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wno-unused-result",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]

HIP_SOURCES = ["context.hip", "gemm.hip", "gemv.hip", "spmv.hip", "cg.hip", "lu.hip", "lu_mixed.hip", "getrf.hip", "ozaki.hip",
               "server.hip", "mmio.hip", "complex.hip"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")


def build(verbose: bool = True) -> Path:
    """Compile every HIP source for gfx950 and link libblaze_gpu.so; build libbzg_gen.so."""
    LIB.mkdir(parents=True, exist_ok=True)
    OBJ.mkdir(parents=True, exist_ok=True)
    srcs = [CSRC / s for s in HIP_SOURCES if (CSRC / s).exists()]

    def compile_one(src: Path) -> Path:
        obj = OBJ / (src.stem + ".o")
        _run([HIPCC, *HIP_FLAGS, "-c", str(src), "-o", str(obj)])
        return obj

    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(compile_one, srcs))
    out = LIB / "libblaze_gpu.so"
    tmp = LIB / "libblaze_gpu.so.tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp),
          *map(str, objs), "-lpthread"])
    os.replace(tmp, out)

    # the BLAS/LAPACK seam (host code over the C ABI), linked against the library above
    seam_out = LIB / "libblaze_seam.so"
    seam_tmp = LIB / "libblaze_seam.so.tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", f"-I{INCLUDE}", str(CSRC / "seam.cpp"), "-o",
          str(seam_tmp), f"-L{LIB}", "-lblaze_gpu", "-Wl,-rpath,$ORIGIN"])
    os.replace(seam_tmp, seam_out)

    gen_out = LIB / "libbzg_gen.so"
    gen_tmp = LIB / "libbzg_gen.so.tmp"
    _run(["gcc", "-O3", "-march=x86-64-v3", "-std=c11", "-fPIC", "-shared", "-fopenmp", f"-I{INCLUDE}",
          str(CSRC / "gen.c"), "-o", str(gen_tmp), "-lm"])
    os.replace(gen_tmp, gen_out)
    if verbose:
        print(f"built {out} and {gen_out}")
    return out


def clean() -> None:
    shutil.rmtree(ROOT / "_build", ignore_errors=True)
    for f in LIB.glob("*.so"):
        f.unlink()


if __name__ == "__main__":
    build()
