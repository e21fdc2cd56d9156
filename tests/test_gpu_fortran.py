"""GPU parity of the Fortran BsaLib front: a user-style program (setters,
bsa_Run, bsa_Finalise through module BsaLib) against the reference's golden
moments (1e-12 normwise relative, FP64; 1e-5 for split-bf16x3 third moments)
and, for the frequency-grid checks of computeFreqsVect_
(src/bsa/classic/BsaClassicImpl.F90:437-560), against the C ABI run on the
grid those checks must produce."""
import math

import numpy as np
import pytest

from oracle import oracle as O
from tests import fdriver
from tests.cases import CASES
from tests.conftest import load_golden, rel_err

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["tiny_u", "tiny_uvw", "tiny_dav", "config1", "tiny_2dof"])
def test_fortran_front_matches_reference(name):
    g = load_golden(name)
    p = CASES[name]()
    out, _ = fdriver.run(p)
    np.testing.assert_array_equal(out["modes"], O.kept_modes(p))
    for k in ("m2mf", "m2mr", "m3mf", "m3mr"):
        assert rel_err(out[k], g[k]) < 1e-12, (k, rel_err(out[k], g[k]))
    o = O.modal_psd(p)
    assert rel_err(out["m2o2mr"], o["m2o2mr"]) < 1e-12


def test_fortran_front_bf16x3():
    g = load_golden("config1")
    out, _ = fdriver.run(CASES["config1"](), "bf16x3")
    assert rel_err(out["m3mf"], g["m3mf"]) < 1e-5
    assert rel_err(out["m3mr"], g["m3mr"]) < 1e-5
    assert rel_err(out["m2mr"], g["m2mr"]) < 1e-12


@pytest.mark.parametrize("name", ["tiny_uvw", "config1"])
def test_fortran_front_scalar_version(gpu_lib, name):
    """The scalar Classic version integrates third moments by plain Riemann
    sums (BsaClassicImpl.F90:331,341); second moments keep the vectorised rule."""
    p = CASES[name]()
    out, r = fdriver.run(p, "scalar")
    assert "SCALAR" in r.stdout
    o = O.classic_factorised(p, rule="riemann", truncate=True, nthreads=8)
    assert rel_err(out["m3mf"], o["m3mf"]) < 1e-12
    assert rel_err(out["m3mr"], o["m3mr"]) < 1e-12
    g = load_golden(name)
    assert rel_err(out["m2mr"], g["m2mr"]) < 1e-12


def _expected_grid(p):
    """The two branches of the grid check, worked by hand for this tower:
    natural frequencies 0.2/0.6/1.2 Hz, damping ratio 0.01 -> the reference
    spacing is min(xsi f) / 5 = 0.002 / 5 = 4e-4 Hz."""
    xsi = np.diag(p.Cm) / (2.0 * np.sqrt(p.Km * p.Mm))
    df_ref = float(np.min(xsi * p.natf) / 5)
    assert abs(df_ref - 4e-4) < 1e-15
    nf1, df = p.nfreqs - 1, p.df
    fmax_ref = float(np.max(p.natf))
    if df * nf1 < fmax_ref:
        if df > df_ref:
            if df_ref * nf1 < fmax_ref:
                nf1 = math.ceil(fmax_ref / df_ref)
            df = df_ref
        else:
            nf1 = math.ceil(fmax_ref / df)
    return nf1 + 1, df


@pytest.mark.parametrize("nfreqs,df,branch", [(17, 0.03125, "df_too_big"),
                                              (17, 0.0002, "df_ok")])
def test_fortran_front_grid_checks(gpu_lib, nfreqs, df, branch):
    p = CASES["config1"]()
    p.nfreqs, p.df = nfreqs, df
    nfh, dfx = _expected_grid(p)
    assert (dfx == 4e-4) == (branch == "df_too_big")
    assert nfh > 2000
    out, r = fdriver.run(p, "fp64", "grid")
    assert f"nfreqs= {nfh}" in r.stdout, r.stdout
    # outside test mode a bispectrum above 1e9 elements runs the scalar version
    # (plain Riemann third moments): 12001^2 * 27 does, 6001^2 * 27 does not
    scalar = (2 * nfh - 1) ** 2 * 27 > 10 ** 9
    assert scalar == (branch == "df_ok")
    assert ("SCALAR" in r.stdout) == scalar
    q = CASES["config1"]()
    q.nfreqs, q.df = nfh, dfx
    want = gpu_lib.classic_run(q, quad="riemann" if scalar else "vect")
    for k in ("m2mf", "m2mr", "m2o2mr", "m3mf", "m3mr"):
        assert rel_err(out[k], want[k]) < 1e-13, k


@pytest.mark.parametrize("mode", ["mesher", "both"])
def test_fortran_front_mesher(gpu_lib, mode):
    """Analysis type 2 (Mesher) and 3 (both) through bsa_Run: the Mesher moments
    equal the C ABI Mesher with the same settings (bsa_setupMesher(0, 20, 2, 2, 2,
    ...), background width 1/max(L/U)) at the levels bsahip_mesher_adapt gives for
    the default target error 1e-3; bsa_setMesherTolerance(0) keeps the rule-based
    levels."""
    from bsalib_amd import abi, host
    p = CASES["config1"]()
    out, r = fdriver.run(p, mode)
    assert "Mesher on the GPU" in r.stdout
    M = abi.Mesher(p, bkg_base_rfmnt=20, tile_cells=32, bkg_peakw=host.bkg_peak_width(p))
    lv, _ = M.adapt(1e-3)
    want = M.run(levels=lv)
    rule = M.run()
    M.close()
    assert rel_err(out["m3mf_msh"], want["m3mf"]) < 1e-13
    assert rel_err(out["m3mr_msh"], want["m3mr"]) < 1e-13
    if mode == "mesher":
        out0, _ = fdriver.run(p, mode, "mtol=0")
        assert rel_err(out0["m3mr_msh"], rule["m3mr"]) < 1e-13
    if mode == "mesher":
        assert out["m3mr"] is None and out["m2mr"] is None
    else:
        g = load_golden("config1")
        assert rel_err(out["m3mr"], g["m3mr"]) < 1e-12


@pytest.mark.parametrize("opts,pod", [
    ((), dict()),
    (("podn=2",), dict(pod_nmodes=2)),
    (("podt=20",), dict(pod_trunc_lim=0.2)),
    (("podn=5",), dict()),                  # >= loaded nodes: ignored (BsaLibImpl.F90:483)
    (("podt=100",), dict()),                # 100 %: ignored (BsaLibImpl.F90:480)
])
def test_fortran_front_mesher_pod(gpu_lib, opts, pod):
    """bsa_setupMesher(isvd = 1, ...): the POD BFM reads coherence column 1 for every
    component (T4) and applies bsa_setPODNOfModesKept / bsa_setPODTruncationThreshold
    as validateAll_ validates them; equal to the C ABI Mesher with the same
    problem settings."""
    import copy
    from bsalib_amd import abi, host
    p = copy.deepcopy(CASES["tiny_uvw"]())
    c1 = p.nod_corr[:, 0].copy()
    p.nod_corr = np.asfortranarray(np.stack([c1, c1 ** 2, c1 ** 3], axis=1))
    out, r = fdriver.run(p, "mesher", "pod", *opts)
    assert "Mesher on the GPU" in r.stdout
    M = abi.Mesher(p, bkg_base_rfmnt=20, tile_cells=32, bkg_peakw=host.bkg_peak_width(p),
                   bound=abi.Bound(p, corr_col_mode=1, **pod))
    want = M.run(levels=M.adapt(1e-3)[0])
    M.close()
    assert rel_err(out["m3mf_msh"], want["m3mf"]) < 1e-13
    assert rel_err(out["m3mr_msh"], want["m3mr"]) < 1e-13
    if not opts:
        direct, _ = fdriver.run(p, "mesher")             # isvd = 0: column tc
        assert rel_err(direct["m3mr_msh"], out["m3mr_msh"]) > 1e-5
