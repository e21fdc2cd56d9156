"""Mixed-precision (split-bf16 x3, FP32 accumulate, FP64 FRF/quadrature)
against the FP64 path.  Tolerances: 1e-5 normwise on m3, and the target of
the mixed path, skewness error <= 1e-4 (max|sk_mixed - sk_fp64| / max|sk_fp64|)."""
import numpy as np
import pytest

from bsalib_amd import host, synth
from tests.cases import CASES
from tests.conftest import load_golden, rel_err

pytestmark = pytest.mark.gpu


def _sk(r, nm):
    return host.skewness(r["m2mr"], r["m3mr"], nm)


@pytest.mark.parametrize("name", ["tiny_u", "tiny_uvw", "config1"])
def test_mixed_close_to_reference(gpu_lib, name):
    g = load_golden(name)
    r = gpu_lib.classic_run(CASES[name](), prec="bf16x3")
    assert rel_err(r["m3mf"], g["m3mf"]) < 1e-5
    assert rel_err(r["m3mr"], g["m3mr"]) < 1e-5
    for k in ("m2mf", "m2mr"):             # the modal PSD path stays FP64
        assert rel_err(r[k], g[k]) < 1e-12


@pytest.mark.parametrize("cfg", ["config2", "config3"])
def test_mixed_skewness_error_full_size(gpu_lib, cfg):
    p = getattr(synth, cfg)()
    b = gpu_lib.Bound(p)
    r64 = gpu_lib.classic_run(p, prec="fp64", bound=b)
    r16 = gpu_lib.classic_run(p, prec="bf16x3", bound=b)
    nm = b.nm
    assert rel_err(r16["m3mr"], r64["m3mr"]) < 1e-5
    assert rel_err(r16["m3mf"], r64["m3mf"]) < 1e-5
    s64, s16 = _sk(r64, nm), _sk(r16, nm)
    assert rel_err(s16, s64) < 1e-4
