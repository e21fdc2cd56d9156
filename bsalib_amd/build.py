"""In-tree build of the GPU library and its Fortran front.

* bsalib_amd/_lib/libbsahip.so: hipcc compiles every HIP source for gfx950
  only; packed-FP32 instruction selection is disabled (``-packed-fp32-ops``).
* bsalib_amd/_lib/fortran/: amdflang compiles the Fortran front (modules
  BsaHip_C and BsaLib) into libbsalibf.a plus the .mod files a Fortran
  program needs (``use BsaLib``; link ``libbsalibf.a -lbsahip``).
* bsalib_amd/_lib/bsalib_fdriver: the Fortran test driver
  (tests/fortran/bsalib_driver.F90) linked against both.

No GPU is needed to build.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OUT_DIR = PKG / "_lib"
LIB = OUT_DIR / "libbsahip.so"
ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ROCM_LIB = "/opt/rocm/lib"      # rocSOLVER + rocBLAS: the POD eigendecomposition
FC = os.environ.get("FC", "/opt/rocm/bin/amdflang")
FDIR = PKG / "fortran"
FOUT = OUT_DIR / "fortran"
FLIB = FOUT / "libbsalibf.a"
FDRIVER_SRC = ROOT / "tests" / "fortran" / "bsalib_driver.F90"
FDRIVER = OUT_DIR / "bsalib_fdriver"

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-shared", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wno-unused-result", "-mllvm", "-disable-promote-alloca-to-lds",
]


def sources():
    return [CSRC / "bsahip_api.hip", CSRC / "bsacl_compat.cpp"]


def deps():
    return sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp")) + sorted(CSRC.glob("*.h")) + \
        sorted((ROOT / "include").glob("*.h"))


def build(verbose: bool = True) -> Path:
    OUT_DIR.mkdir(exist_ok=True)
    srcs = [s for s in sources() if s.exists()]
    cmd = [HIPCC, *HIP_FLAGS, "-I", str(ROOT / "include"), *map(str, srcs), "-o", str(LIB),
           "-L", ROCM_LIB, "-lrocsolver", "-lrocblas"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    return LIB


def build_fortran(verbose: bool = True) -> Path:
    """Fortran front (needs libbsahip.so for the driver link)."""
    FOUT.mkdir(parents=True, exist_ok=True)
    objs = []
    for unit in ("bsahip_c.F90", "bsalib.F90"):            # dependency order
        obj = FOUT / unit.replace(".F90", ".o")
        cmd = [FC, "-cpp", "-O2", "-fPIC", "-J", str(FOUT), "-c", str(FDIR / unit), "-o", str(obj)]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.check_call(cmd)
        objs.append(str(obj))
    if FLIB.exists():
        FLIB.unlink()
    subprocess.check_call(["ar", "rcs", str(FLIB), *objs])
    if FDRIVER_SRC.exists():
        obj = FOUT / "bsalib_driver.o"
        subprocess.check_call([FC, "-cpp", "-O2", "-I", str(FOUT), "-J", str(FOUT), "-c",
                               str(FDRIVER_SRC), "-o", str(obj)])
        cmd = [FC, "-o", str(FDRIVER), str(obj), str(FLIB), "-L", str(OUT_DIR), "-lbsahip",
               "-Wl,-rpath,$ORIGIN"]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.check_call(cmd)
    return FLIB


def fortran_needs_build() -> bool:
    if not (FLIB.exists() and FDRIVER.exists()):
        return True
    t = min(FLIB.stat().st_mtime, FDRIVER.stat().st_mtime)
    srcs = sorted(FDIR.glob("*.F90")) + [FDRIVER_SRC]
    return any(p.stat().st_mtime > t for p in srcs if p.exists())


def needs_build() -> bool:
    if not LIB.exists():
        return True
    t = LIB.stat().st_mtime
    return any(p.stat().st_mtime > t for p in deps())


if __name__ == "__main__":
    build()
    build_fortran()
    sys.exit(0)
