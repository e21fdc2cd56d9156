"""CPU tests of the C ABI: the library exists, loads, exports every symbol the
headers declare, and refuses to compute without a GPU (no CPU fallback)."""
import ctypes as C
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        txt = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        txt = re.sub(r"//[^\n]*", "", txt)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b((?:bsahip|bsacl)\w*)\s*\(", txt, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


@pytest.fixture(scope="module")
def libpath():
    from bsalib_amd import build
    if build.needs_build():
        build.build(verbose=False)
    return build.LIB


def test_headers_declare_the_expected_entry_points():
    names = declared_symbols()
    for must in ("bsahip_classic_run", "bsahip_bfm_points", "bsahip_classic_spectra", "bsaclInit",
                 "bsaclRun", "bsaclFinalise", "bsaclAcquireComputationFreqs", "bsaclSetKernelID"):
        assert must in names


def test_library_exports_every_declared_symbol(libpath):
    out = subprocess.run(["nm", "-D", "--defined-only", str(libpath)], check=True,
                         capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if line.strip()}
    missing = [n for n in declared_symbols() if n not in exported]
    assert not missing, missing


def test_library_loads_and_reports_version(libpath):
    L = C.CDLL(str(libpath))
    assert L.bsahip_version() >= 100


def test_no_cpu_fallback_without_gpu(libpath):
    L = C.CDLL(str(libpath))
    L.bsahip_last_error.restype = C.c_char_p
    rc = L.bsahip_init(0)
    if rc == 0:
        L.bsahip_finalise()
        pytest.skip("a GPU is present")
    assert rc == -2                                    # BSAHIP_EDEVICE
    assert b"device" in L.bsahip_last_error()
    from bsalib_amd import abi, synth
    with pytest.raises(abi.BsaHipError):
        abi.classic_run(synth.tower(nn=3, nm=2, nfreqs=5, df=0.2))
