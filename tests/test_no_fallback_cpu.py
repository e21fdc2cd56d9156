"""The product path has no CPU fallback: without the HIP library the bindings
refuse to load, and without a device every entry point returns an error
instead of computing (checked on a host with no GPU)."""
import os
import subprocess
import sys
from pathlib import Path

import pytest

from bsalib_amd import abi
from tests.cases import CASES

ROOT = Path(__file__).resolve().parent.parent


def test_missing_library_refuses_to_load():
    code = ("import sys; from pathlib import Path; sys.path.insert(0, %r)\n"
            "from bsalib_amd import abi\n"
            "abi.LIB_PATH = Path('/nonexistent/libbsahip.so')\n"
            "try:\n"
            "    abi.lib()\n"
            "except FileNotFoundError as e:\n"
            "    print('refused:', e)\n"
            "    sys.exit(0)\n"
            "sys.exit(3)\n") % str(ROOT)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "there is no CPU path" in r.stdout


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU driver is present on this host")
def test_entry_points_error_without_a_device():
    p = CASES["tiny_u"]()
    calls = [lambda: abi.init(0),
             lambda: abi.classic_run(p),
             lambda: abi.classic_run(p, prec="bf16x3"),
             lambda: abi.classic_spectra(p)]
    for call in calls:
        with pytest.raises(abi.BsaHipError) as e:
            call()
        assert e.value.code == abi.EDEVICE
        assert "no HIP device" in str(e.value)
