"""The reference's standalone program (app/bsa.F90, unmodified, compiled against
this library's module BsaLib by oracle/ref/build_ref.sh into oracle/_ref/bsa_app)
reads bsa.bsadata / bsa.extdata written by bsalib_amd.inputs with its own
reader, runs bsa_Run on the GPU and exports its results.  The exported moments
equal the C ABI Classic run of the same problem, and the exported post-processing
(background/resonant split, skewness, the program's own nodal recombination and
peak factors) equals the oracle on those moments -- 1e-12."""
import subprocess

import numpy as np
import pytest

from bsalib_amd import abi, inputs
from oracle import oracle as O
from tests.cases import CASES
from tests.conftest import rel_err

pytestmark = pytest.mark.gpu

APP = O.HERE / "_ref" / "bsa_app"


def _vector(path, modal=False):
    v = path.read_text().split()
    if modal:
        nm = int(v[0])
        v = v[1 + nm:]
    n = int(v[0])
    return np.array([float(x) for x in v[1:1 + n]])


@pytest.mark.skipif(not APP.exists(), reason="oracle/_ref/bsa_app not built (oracle/ref/build_ref.sh)")
@pytest.mark.parametrize("name", ["tiny_u", "config1"])
def test_reference_program_runs_on_this_library(gpu_lib, tmp_path, name):
    p = CASES[name]()
    inputs.write_inputs(p, tmp_path, settings=dict(analysis=1, test=1))
    r = subprocess.run([str(APP)], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "BSA terminated correctly" in r.stdout, r.stdout[-2000:] + r.stderr
    want = abi.classic_run(p)
    nm = len(O.kept_modes(p))
    f = lambda s: tmp_path / f"export_{s}.txt"
    assert rel_err(_vector(f("m2_mf_cls_full")), want["m2mf"]) < 1e-12
    assert rel_err(_vector(f("m2_mr_cls_full")), want["m2mr"]) < 1e-12
    assert rel_err(_vector(f("m3_mf_cls_full")), want["m3mf"]) < 1e-12
    assert rel_err(_vector(f("m3_mr_cls_full")), want["m3mr"]) < 1e-12
    sk = O.skewness(want["m2mr"], want["m3mr"], nm)
    assert rel_err(_vector(f("sk_mr_cls_full"), modal=True), sk.reshape(-1, order="F")) < 1e-12
    # background / resonant split: modes, damping ratios, then (bkg, res) per mode
    v = f("m2_BR_decomp").read_text().split()
    k = int(v[0])
    br = np.array([float(x) for x in v[1 + 2 * k:]]).reshape(k, 2)
    bkg, res = O.br_decomp(p, want["m2mf"])
    assert rel_err(br[:, 0], bkg) < 1e-12 and rel_err(br[:, 1], res) < 1e-12
    # the program's own recombination over all DOFs and its peak factors
    modes = O.kept_modes(p)
    rc = O.modal_recombination(p.phi[:, modes - 1], want["m2mr"], want["m3mr"], want["m2o2mr"])
    assert rel_err(_vector(f("sk_r_full")), rc["sk_full"]) < 1e-12
    pf = O.peak_factors(rc["m2_full"], rc["m2o2_full"], 600.0, sk=rc["sk_full"])
    assert rel_err(_vector(f("peak_pos_r_fullF_g")), pf[0]) < 1e-12
    assert rel_err(_vector(f("peak_pos_r_fullF_ng")), pf[1]) < 1e-12
    assert rel_err(_vector(f("peak_neg_r_fullF_ng")), pf[2]) < 1e-12
