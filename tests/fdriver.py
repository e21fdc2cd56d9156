"""Runs the Fortran BsaLib front through its test driver (tests/fortran/bsalib_driver.F90,
built to bsalib_amd/_lib/bsalib_fdriver) and reads back what bsa_Run returned."""
from __future__ import annotations

import os
import subprocess
import tempfile

import numpy as np

from bsalib_amd import build
from oracle import oracle as O


def ensure_built():
    """The driver as __graft_entry__.build() left it: tests never compile (a
    GPU box runs the tree with the binaries built beside its sources)."""
    if not build.FDRIVER.exists():
        raise FileNotFoundError(f"{build.FDRIVER} is missing: run __graft_entry__.build() first")
    return build.FDRIVER


def run(prob, *opts, check=True, timeout=300, workdir=None):
    """Returns (outputs dict or None, completed process).  The driver runs in
    workdir (default: a temporary directory), where files it writes stay."""
    drv = ensure_built()
    with tempfile.TemporaryDirectory() as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        O._write_ref_input(prob, fin, 1)
        r = subprocess.run([str(drv), fin, fout, *opts], capture_output=True, text=True,
                           timeout=timeout, cwd=str(workdir or td))
        if r.returncode != 0:
            if check:
                raise RuntimeError(f"bsalib_fdriver failed ({r.returncode}):\n{r.stdout}\n{r.stderr}")
            return None, r
        raw = open(fout, "rb").read()
    nme = int(np.frombuffer(raw[:4], dtype=np.int32)[0])
    modes = np.frombuffer(raw[4:4 + 4 * nme], dtype=np.int32).copy()
    o = 4 + 4 * nme
    out = {"modes": modes}
    keys = ["m2mf", "m2mr", "m2o2mr", "m3mf", "m3mr", "m3mf_msh", "m3mr_msh"]
    if "post" in opts:
        keys += ["bkg", "res", "m2_diag", "m3_diag", "m2o2_diag", "sk_diag", "m2_full",
                 "m3_full", "m2o2_full", "sk_full", "peak_g_diag", "peak_pos_diag",
                 "peak_neg_diag", "peak_g_full", "peak_pos_full", "peak_neg_full"]
    if "slots" in opts:
        keys += ["slot_psd_f", "slot_psd_r", "slot_bisp_f", "slot_bisp_r"]
    for k in keys:
        n = int(np.frombuffer(raw[o:o + 4], dtype=np.int32)[0])
        o += 4
        out[k] = np.frombuffer(raw[o:o + 8 * n], dtype=np.float64).copy() if n else None
        o += 8 * n
    assert o == len(raw)
    return out, r
