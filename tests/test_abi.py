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


def _c_offsets(struct: str, fields):
    """offsetof of each field and sizeof, compiled from include/bsahip.h with gcc."""
    import tempfile
    body = "\n".join(f'   printf("%zu\\n", offsetof({struct}, {f}));' for f in fields)
    src = (f'#include <stdio.h>\n#include <stddef.h>\n#include "bsahip.h"\n'
           f'int main(void) {{\n{body}\n   printf("%zu\\n", sizeof({struct}));\n   return 0;\n}}\n')
    with tempfile.TemporaryDirectory() as td:
        c, exe = Path(td) / "probe.c", Path(td) / "probe"
        c.write_text(src)
        subprocess.run(["gcc", "-I", str(ROOT / "include"), str(c), "-o", str(exe)], check=True)
        return [int(x) for x in subprocess.run([str(exe)], check=True, capture_output=True,
                                               text=True).stdout.split()]


@pytest.mark.parametrize("cname,pyname", [("bsahip_problem", "Problem_t"),
                                          ("bsahip_mesher_cfg", "MesherCfg")])
def test_ctypes_mirrors_match_the_c_layout(cname, pyname):
    """The Python mirrors (and, field for field, the Fortran bind(c) types in
    bsalib_amd/fortran/bsahip_c.F90) must have the header's offsets."""
    from bsalib_amd import abi
    T = getattr(abi, pyname)
    names = [f[0] for f in T._fields_]
    want = _c_offsets(cname, names)
    got = [getattr(T, n).offset for n in names] + [C.sizeof(T)]
    assert got == want, list(zip(names + ["sizeof"], got, want))


def test_fortran_types_list_the_c_fields_in_order():
    """bind(c) derived types are laid out by field order: the Fortran types must
    name the same fields in the same order as the C structs."""
    from bsalib_amd import abi
    src = (ROOT / "bsalib_amd" / "fortran" / "bsahip_c.F90").read_text()
    for ftype, T in (("bsahip_problem_t", abi.Problem_t), ("bsahip_mesher_cfg_t", abi.MesherCfg)):
        block = re.search(rf"type, bind\(c\), public :: {ftype}\n(.*?)end type", src, re.S).group(1)
        fields = []
        for line in block.splitlines():
            if "::" not in line:
                continue
            decl = line.split("::", 1)[1]
            fields += [re.match(r"\s*(\w+)", part).group(1) for part in decl.split(",")]
        assert [f.lower() for f in fields] == [f[0].lower() for f in T._fields_]


def test_wrappers_check_sizes_before_calling_the_library():
    """frf_apply / bfm_points / bfm_points_cls refuse inconsistent arrays on the
    host (ValueError), so the C ABI never reads past a caller's buffer."""
    import numpy as np
    from bsalib_amd import abi
    from tests.cases import CASES
    p = CASES["tiny_uvw"]()
    b = abi.Bound(p)
    nm, f = b.nm, p.freqs()
    with pytest.raises(ValueError):
        abi.frf_apply(p, abi.FRF_PSD, np.zeros((f.size, nm)), f, bound=b)              # NM^2 columns
    with pytest.raises(ValueError):
        abi.frf_apply(p, abi.FRF_PSD, np.zeros((f.size, nm * nm)), f, flags=abi.ONLY_DIAG, bound=b)
    with pytest.raises(ValueError):
        abi.frf_apply(p, abi.FRF_GRID, np.zeros((f.size, f.size, nm)), f, f, bound=b)
    with pytest.raises(ValueError):
        abi.frf_apply(p, abi.FRF_POINTS, np.zeros((nm ** 3, 3)), f[:3], f[:2], bound=b)
    with pytest.raises(ValueError):
        abi.bfm_points(p, f[:3], f[:2], bound=b)
    with pytest.raises(ValueError):
        abi.bfm_points_cls(p, f[:3], f[:2], bound=b, only_diag=True)
