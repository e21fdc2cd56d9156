"""GPU parity of the BsaCL C ABI (bsaclRun) against the oracle's FP64
restatement of the reference BsaCL kernels (BsaCL/src/bsacl.cl:179-930)."""
import subprocess
import sys
import textwrap
from pathlib import Path

import numpy as np
import pytest

from bsalib_amd import host, synth
from bsalib_amd.bsacl import BsaCL
from oracle import oracle as O
from tests.conftest import rel_err

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parents[1]


def _problem(psd_id=5, comps=(1,), nn=6, nm=3, nfreqs=9):
    return synth.tower(nn=nn, nm=nm, comps=comps, nfreqs=nfreqs, df=0.1, seed=5,
                       psd_id=psd_id, zone_per_node=False)


def _run(prob, fi, fj, kid=2, h_in=None):
    modes = host.kept_modes(prob)
    psi = host.build_psi(prob, modes)
    cl = BsaCL()
    cl.acquire_problem(prob, psi, modes)
    assert cl.init(1) == 0
    assert cl.init_device_memory() == 0
    cl.acquire_computation_freqs(fi, fj)
    assert cl.set_kernel_id(kid) == 0
    dim = len(modes) ** 3
    h = np.zeros(dim) if h_in is None else h_in.copy()
    rc, out = cl.run(dim, h_res=h)
    cl.finalise()
    assert rc == 0
    return out


@pytest.mark.parametrize("psd_id,comps", [(5, (1,)), (1, (1,)), (5, (1, 3))])
def test_bsacl_kernel2_uniform_grid(gpu_lib, psd_id, comps):
    p = _problem(psd_id, comps)
    f = p.freqs()
    got = _run(p, f, f)
    want = O.bsacl_run(p, f, f)
    assert rel_err(got, want) < 1e-12


def test_bsacl_kernel2_arbitrary_frequencies(gpu_lib):
    p = _problem(5, (1, 2))
    fi = np.array([-0.35, -0.1, 0.05, 0.2, 0.41])
    fj = np.array([-0.3, 0.0, 0.15, 0.25])
    assert rel_err(_run(p, fi, fj), O.bsacl_run(p, fi, fj)) < 1e-12


def test_bsacl_kernel1_force_moment(gpu_lib):
    p = _problem(5, (1,))
    f = p.freqs()
    h0 = np.linspace(1.0, 2.0, 27)
    got = _run(p, f, f, kid=1, h_in=h0)
    fx = O.classic_factorised(p, rule="riemann", truncate=False)
    d = 2 * np.pi * p.df
    assert rel_err(got, h0 * d * d + fx["m3mf"]) < 1e-12


def test_bsacl_run_error_codes_on_gpu(gpu_lib):
    code = textwrap.dedent("""
    import sys, numpy as np
    sys.path.insert(0, %r)
    from bsalib_amd.bsacl import BsaCL
    from bsalib_amd import synth, host
    p = synth.tower(nn=4, nm=2, nfreqs=5, df=0.2, zone_per_node=False, psd_id=5)
    modes = host.kept_modes(p); psi = host.build_psi(p, modes)
    cl = BsaCL(); cl.acquire_problem(p, psi, modes)
    print('INIT', cl.init(1), 'DEVMEM', cl.init_device_memory())
    print('NOFREQ', cl.run(8)[0])
    f = p.freqs(); cl.acquire_computation_freqs(f, f)
    import ctypes
    print('NULLRES', cl.L.bsaclRun(ctypes.c_uint(0), ctypes.c_size_t(8), None))
    print('AFTER', cl.run(8)[0], cl.init(1))
    """ % str(ROOT))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = r.stdout
    assert "INIT 0 DEVMEM 0" in out
    assert "NOFREQ -20" in out
    assert "NULLRES -123" in out
    assert "AFTER -1 -999" in out


REF_CL_DRIVER = ROOT / "oracle" / "_ref" / "bsacl_module_driver"


@pytest.mark.skipif(not REF_CL_DRIVER.exists(),
                    reason="reference BsaCL Fortran binding not built (needs the reference sources)")
@pytest.mark.parametrize("psd_id,comps", [(5, (1,)), (5, (1, 3))])
def test_reference_fortran_bsacl_module_binds_unchanged(gpu_lib, tmp_path, psd_id, comps):
    """The reference's own module BsaCL (BsaCL/BsaCL_Fortran, compiled unmodified)
    linked against libbsahip.so, called in bsa_Run's order."""
    p = _problem(psd_id, comps)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    O._write_ref_input(p, str(fin), 1)
    r = subprocess.run([str(REF_CL_DRIVER), str(fin), str(fout)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = fout.read_bytes()
    rc = np.frombuffer(raw[:16], dtype=np.int32)
    assert rc.tolist() == [0, 0, 0, 0], (rc, r.stdout)
    got = np.frombuffer(raw[16:], dtype=np.float64)
    f = p.freqs()
    assert rel_err(got, O.bsacl_run(p, f, f)) < 1e-12
