"""CPU tests of the Fortran BsaLib front (module BsaLib, driven by
tests/fortran/bsalib_driver.F90): its validation aborts as the reference's
bsa_Run does, and without a GPU it stops instead of computing on the CPU."""
import os

import numpy as np
import pytest

from tests import fdriver
from tests.cases import CASES

no_device = pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")


@no_device
def test_front_without_gpu_aborts_loudly():
    out, r = fdriver.run(CASES["tiny_u"](), check=False)
    assert out is None and r.returncode != 0
    assert "GPU initialisation failed" in r.stdout


def test_front_rejects_structure_without_normalised_modes():
    """validateModalInfo_ (BsaLibImpl.F90:545-599) drops every mode whose max|phi|
    is not 1; with none left the run aborts before touching the device."""
    p = CASES["tiny_u"]()
    p.phi = p.phi * 0.5
    out, r = fdriver.run(p, check=False)
    assert out is None and r.returncode != 0
    assert "is not 1-normalised" in r.stdout
    assert "Effective number of modes to be kept is 0" in r.stdout


def test_fortran_library_and_modules_built():
    from bsalib_amd import build
    fdriver.ensure_built()
    assert build.FLIB.exists()
    assert (build.FOUT / "bsalib.mod").exists() and (build.FOUT / "bsahip_c.mod").exists()
