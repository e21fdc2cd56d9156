"""GPU test of bsahip_classic_run_noalloc, the Classic run whose device buffers
all come from the library's static scratch array: two calls in a row through
the C ABI (the tower of configuration 2 on its 513^2 grid, then the smaller
20-node tower on 129^2) each equal bsahip_classic_run at 1e-12 (FP64; the
atomic summation order differs) and at 1e-12 in split-bf16x3 (the same
kernels); a problem beyond the array is refused with a message."""
import ctypes as C

import numpy as np
import pytest

from bsalib_amd import abi, synth
from tests.conftest import rel_err

pytestmark = pytest.mark.gpu


def _noalloc(prob, prec="fp64", quad="vect"):
    b = abi.Bound(prob)
    nm = b.nm
    out = {k: np.zeros(nm * nm) for k in ("m2mf", "m2mr", "m2o2mr")}
    out.update({k: np.zeros(nm ** 3) for k in ("m3mf", "m3mr")})
    flags = abi.PREC[prec] | abi.QUAD[quad]
    L = abi.lib()
    rc = L.bsahip_classic_run_noalloc(C.byref(b.struct), C.c_int32(prob.nfreqs),
                                      C.c_double(prob.df), C.c_int32(flags),
                                      *[abi._ptr(out[k], C.c_double)
                                        for k in ("m2mf", "m2mr", "m2o2mr", "m3mf", "m3mr")])
    return rc, out


@pytest.mark.parametrize("prec", ["fp64", "bf16x3"])
def test_noalloc_equals_classic_run_twice(gpu_lib, prec):
    for prob in (synth.config2(), synth.config1()):        # second call: a smaller input
        print(f"[noalloc] {prob.name} NN={prob.nn} NM={prob.nm} NF={prob.nf} {prec}", flush=True)
        rc, r = _noalloc(prob, prec)
        assert rc == 0, abi.lib().bsahip_last_error().decode()
        ref = gpu_lib.classic_run(prob, prec=prec)
        for k in ("m2mf", "m2mr", "m2o2mr", "m3mf", "m3mr"):
            assert rel_err(r[k], ref[k]) < 1e-12, (prob.name, k, rel_err(r[k], ref[k]))


def test_noalloc_refuses_a_problem_beyond_the_scratch(gpu_lib):
    rc, _ = _noalloc(synth.config3())                       # A table alone: 260 MB
    assert rc == abi.EINVAL
    msg = abi.lib().bsahip_last_error().decode()
    assert "static device scratch" in msg
    # the library still works afterwards, allocating and not
    rc, r = _noalloc(synth.config1())
    assert rc == 0
    ref = gpu_lib.classic_run(synth.config1())
    assert rel_err(r["m3mr"], ref["m3mr"]) < 1e-12
