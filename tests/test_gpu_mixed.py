"""Mixed precision (split-bf16 x3 and FP32 node contraction, FP32 accumulate,
FP64 FRF/quadrature) against the FP64 path.  Tolerances: 1e-5 normwise on m3,
and the target of the mixed path, skewness error <= 1e-4 (max|sk_mixed -
sk_fp64| / max|sk_fp64|); bispectra of selected entries with the mixed G
values (bsahip_classic_bispectra_sel) within 1e-5 in L2 of FP64."""
import numpy as np
import pytest

from bsalib_amd import host, synth
from tests.cases import CASES
from tests.conftest import load_golden, rel_err

pytestmark = pytest.mark.gpu


def _sk(r, nm):
    return host.skewness(r["m2mr"], r["m3mr"], nm)


@pytest.mark.parametrize("prec", ["bf16x3", "fp32"])
@pytest.mark.parametrize("name", ["tiny_u", "tiny_uvw", "config1"])
def test_mixed_close_to_reference(gpu_lib, name, prec):
    g = load_golden(name)
    r = gpu_lib.classic_run(CASES[name](), prec=prec)
    assert rel_err(r["m3mf"], g["m3mf"]) < 1e-5
    assert rel_err(r["m3mr"], g["m3mr"]) < 1e-5
    for k in ("m2mf", "m2mr"):             # the modal PSD path stays FP64
        assert rel_err(r[k], g[k]) < 1e-12


@pytest.mark.parametrize("prec", ["bf16x3", "fp32"])
@pytest.mark.parametrize("cfg", ["config2", "config3"])
def test_mixed_skewness_error_full_size(gpu_lib, cfg, prec):
    p = getattr(synth, cfg)()
    b = gpu_lib.Bound(p)
    r64 = gpu_lib.classic_run(p, prec="fp64", bound=b)
    r16 = gpu_lib.classic_run(p, prec=prec, bound=b)
    nm = b.nm
    assert rel_err(r16["m3mr"], r64["m3mr"]) < 1e-5
    assert rel_err(r16["m3mf"], r64["m3mf"]) < 1e-5
    s64, s16 = _sk(r64, nm), _sk(r16, nm)
    assert rel_err(s16, s64) < 1e-4


@pytest.mark.parametrize("prec", ["bf16x3", "fp32"])
def test_mixed_bispectra_close_to_fp64(gpu_lib, prec):
    """Selected entries of the tower of configuration 2 (513^2): the mixed G
    values differ from FP64 (they are not the FP64 path) but stay within 1e-5
    in L2 over the plane, force and response."""
    p = synth.config2()
    b = gpu_lib.Bound(p)
    sel = [1, 1 + 10 * 4 + 100 * 9, 1000, 5 + 10 * 5 + 100 * 5]
    r64 = gpu_lib.classic_bispectra_sel(p, sel, bound=b)
    rm = gpu_lib.classic_bispectra_sel(p, sel, bound=b, prec=prec)
    for k in ("bisp_f", "bisp_r"):
        e = np.linalg.norm(rm[k] - r64[k]) / np.linalg.norm(r64[k])
        assert 0.0 < e < 1e-5, (k, e)


def test_fp32_batch_equals_fp32_plans(gpu_lib):
    from bsalib_amd import abi
    ps = synth.batch(5, seed=4, max_nn=100)
    for q in ps:
        q.nfreqs, q.df = 65, 4.0 / 64
    B = abi.Batch(ps, prec="fp32")
    g0 = B.run(mode=0)
    B.close()
    for i, q in enumerate(ps):
        one = gpu_lib.classic_run(q, prec="fp32")
        assert rel_err(g0["m3mr"][i], one["m3mr"]) < 1e-12, i
        r64 = gpu_lib.classic_run(q, prec="fp64")
        assert rel_err(g0["m3mr"][i], r64["m3mr"]) < 1e-5, i
