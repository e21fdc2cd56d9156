"""GPU tests of the fixed-order mode (BSAHIP_DETERMINISTIC): every m3/m2 entry
sums its stored partials in one order with Kahan compensation, so repeated
runs give bitwise-equal moments, which equal the atomic sums to rounding
(1e-13) and the reference's golden moments (1e-12)."""
import numpy as np
import pytest

from bsalib_amd import synth
from tests.cases import CASES
from tests.conftest import load_golden, rel_err

pytestmark = pytest.mark.gpu

KEYS = ("m2mf", "m2mr", "m2o2mr", "m3mf", "m3mr")


@pytest.mark.parametrize("name", ["tiny_uvw", "config1"])
def test_fixed_order_matches_reference(gpu_lib, name):
    g = load_golden(name)
    r = gpu_lib.classic_run(CASES[name](), quad="vect", deterministic=True)
    for k in ("m2mf", "m2mr", "m3mf", "m3mr"):
        assert rel_err(r[k], g[k]) < 1e-12, k


@pytest.mark.parametrize("cfg,prec", [("config1", "fp64"), ("config2", "fp64"),
                                      ("config2", "bf16x3"), ("config2", "fp32"),
                                      ("config3", "bf16x3")])
def test_fixed_order_runs_are_bitwise_equal(gpu_lib, cfg, prec):
    p = getattr(synth, cfg)()
    b = gpu_lib.Bound(p)
    plan = gpu_lib.Plan(p, quad="vect", prec=prec, bound=b, deterministic=True)
    r1 = plan.run()
    r2 = plan.run()
    plan.close()
    r3 = gpu_lib.classic_run(p, quad="vect", prec=prec, bound=b, deterministic=True)
    atom = gpu_lib.classic_run(p, quad="vect", prec=prec, bound=b)
    for k in KEYS:
        np.testing.assert_array_equal(r1[k], r2[k], err_msg=k)
        np.testing.assert_array_equal(r1[k], r3[k], err_msg=k)
        assert rel_err(r1[k], atom[k]) < 1e-13, k
