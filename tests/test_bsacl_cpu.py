"""CPU tests of the BsaCL C ABI state machine and error codes (no GPU needed):
each scenario runs in a fresh process because a halt is sticky for the
process, as in the reference runtime (BsaCL/src/bsacl.c:129, 377-382)."""
import subprocess
import sys
import textwrap
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]

PRELUDE = """
import sys, numpy as np
sys.path.insert(0, %r)
from bsalib_amd.bsacl import BsaCL
from bsalib_amd import synth, host
cl = BsaCL()
p = synth.tower(nn=4, nm=2, nfreqs=5, df=0.2)
""" % str(ROOT)


def run(code):
    r = subprocess.run([sys.executable, "-c", PRELUDE + textwrap.dedent(code)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


@pytest.fixture(scope="module", autouse=True)
def built():
    from bsalib_amd import build
    if build.needs_build():
        build.build(verbose=False)


def test_freqs_before_init_halts_with_89():
    out = run("""
    cl.acquire_computation_freqs(np.zeros(3), np.zeros(3))
    print('INIT', cl.init(1))
    print('RUN', cl.run(8)[0])
    """)
    assert "Aborting  (-89)" in out
    assert "INIT -999" in out and "RUN -1" in out


def test_size_mismatch_halts():
    out = run("""
    cl.acquire_loaded_nodes_list(np.arange(1, 5))
    cl.acquire_wind_coeffs(np.zeros((1, 6, 3), order='F'))
    print('INIT', cl.init(1))
    """)
    assert "Aborting  (-1)" in out and "INIT -999" in out


@pytest.mark.parametrize("missing,code", [("tc", -83), ("nodes", -84), ("psi", -85), ("corr", -86)])
def test_init_device_memory_reports_the_first_missing_array(missing, code):
    out = run(f"""
    missing = {missing!r}
    modes = host.kept_modes(p); psi = host.build_psi(p, modes)
    if missing != 'tc': cl.acquire_turb_components_list(p.tc)
    if missing not in ('tc', 'nodes'): cl.acquire_loaded_nodes_list(p.n_load)
    if missing not in ('tc', 'nodes', 'psi'): cl.acquire_phi_times_c_mat(psi)
    if missing == 'corr':
        pass
    print('RC', cl.init_device_memory())
    print('INIT', cl.init(1))
    """)
    assert f"RC {code}" in out and "INIT -999" in out


def test_kernel_id_and_device_type():
    out = run("""
    print('K3', cl.set_kernel_id(3), 'K1', cl.set_kernel_id(1), 'K2', cl.set_kernel_id(2))
    cl.set_device_type(1)
    cl.set_device_type(9)
    print('INIT', cl.init(1))
    """)
    assert "K3 1 K1 0 K2 0" in out
    assert "Aborting  (-5)" in out and "INIT -999" in out


def test_init_without_gpu_reports_platform_error():
    out = run("""
    from bsalib_amd import abi
    import ctypes
    if abi.lib().bsahip_init(0) == 0:
        print('GPU')
    else:
        print('INIT', cl.init(1))
    """)
    if "GPU" in out:
        pytest.skip("a GPU is present")
    assert "INIT -2" in out
