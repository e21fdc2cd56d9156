"""Build libbsahip.so (HIP, gfx950) in-tree: bsalib_amd/_lib/libbsahip.so.

hipcc compiles every HIP source for gfx950 only; packed-FP32 instruction
selection is disabled (``-packed-fp32-ops``).  No GPU is needed to build.
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
    cmd = [HIPCC, *HIP_FLAGS, "-I", str(ROOT / "include"), *map(str, srcs), "-o", str(LIB)]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.check_call(cmd)
    return LIB


def needs_build() -> bool:
    if not LIB.exists():
        return True
    t = LIB.stat().st_mtime
    return any(p.stat().st_mtime > t for p in deps())


if __name__ == "__main__":
    build()
    sys.exit(0)
