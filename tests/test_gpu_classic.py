"""GPU parity of the Classic path (libbsahip, FP64) against the reference's golden
vectors and the oracle.  Tolerances: 1e-12 normwise relative on moments and
bispectra (max|ours - ref| / max|ref| per array)."""
import numpy as np
import pytest

from bsalib_amd import synth
from oracle import oracle as O
from tests.cases import CASES
from tests.conftest import load_golden, rel_err

pytestmark = pytest.mark.gpu

GOLDEN_CASES = ["tiny_u", "tiny_uvw", "tiny_dav", "config1", "tiny_2dof"]
TOL = 1e-12


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_classic_moments_match_reference(gpu_lib, name):
    g = load_golden(name)
    p = CASES[name]()
    r = gpu_lib.classic_run(p, quad="vect")
    for k in ("m2mf", "m2mr", "m3mf", "m3mr"):
        assert rel_err(r[k], g[k]) < TOL, (k, rel_err(r[k], g[k]))
    o = O.modal_psd(p)
    assert rel_err(r["m2o2mr"], o["m2o2mr"]) < TOL


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_classic_spectra_match_reference(gpu_lib, name):
    g = load_golden(name)
    p = CASES[name]()
    r = gpu_lib.classic_spectra(p)
    for k in ("psd_f", "psd_r"):
        assert rel_err(r[k], g[k]) < TOL, k
    if "bisp_f" in g:
        assert rel_err(r["bisp_f"], g["bisp_f"]) < TOL
        assert rel_err(r["bisp_r"], g["bisp_r"]) < TOL
    else:
        i1, i2 = g["sample_i1"], g["sample_i2"]
        assert rel_err(r["bisp_f"][i1, i2, :], g["bisp_f_sample"]) < TOL
        assert rel_err(r["bisp_r"][i1, i2, :], g["bisp_r_sample"]) < TOL


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_mesher_points_match_reference(gpu_lib, name):
    g = load_golden(name)
    p = CASES[name]()
    bfm, brm = gpu_lib.bfm_points(p, g["msh_fi"], g["msh_fj"])
    assert rel_err(bfm, g["bfm_msh"]) < TOL
    # response at the same points: FRF product of the oracle
    H1 = O.frf(p, g["msh_fi"])
    H2 = O.frf(p, g["msh_fj"])
    H3 = O.frf(p, g["msh_fi"] + g["msh_fj"])
    nm = H1.shape[1]
    fr = (H1[:, :, None, None] * H2[:, None, :, None] * np.conj(H3)[:, None, None, :]).real
    fr = fr.reshape(-1, nm ** 3, order="F").T
    assert rel_err(brm, g["bfm_msh"] * fr) < TOL


@pytest.mark.parametrize("rule", ["riemann", "trapezoid", "vect"])
def test_quadrature_rules_match_oracle(gpu_lib, rule):
    p = CASES["config1"]()
    r = gpu_lib.classic_run(p, quad=rule)
    o = O.classic_factorised(p, rule=rule, truncate=True, nthreads=8)
    assert rel_err(r["m3mf"], o["m3mf"]) < TOL
    assert rel_err(r["m3mr"], o["m3mr"]) < TOL


def test_exact_sum_spectra_match_mesher_oracle(gpu_lib):
    p = CASES["tiny_uvw"]()
    r = gpu_lib.classic_spectra(p, exact=True)
    o = O.classic_factorised(p, rule="vect", truncate=False, want_bisp=True)
    nf = p.nf
    B = r["bisp_f"].reshape(nf * nf, -1, order="F").T
    assert rel_err(B, o["bfm"]) < TOL


def test_config2_reduced_grid_matches_oracle(gpu_lib):
    """Config 2 structure (NN=200, u/v/w, NM=10) on a 129-point grid."""
    p = synth.config2()
    p.nfreqs, p.df = 65, 1.0 / 16.0
    r = gpu_lib.classic_run(p, quad="vect")
    o = O.classic_factorised(p, rule="vect", truncate=True, nthreads=16)
    assert rel_err(r["m3mf"], o["m3mf"]) < TOL
    assert rel_err(r["m3mr"], o["m3mr"]) < TOL
    m = O.modal_psd(p)
    for k in ("m2mf", "m2mr", "m2o2mr"):
        assert rel_err(r[k], m[k]) < TOL, k


def test_config2_full_grid_matches_oracle_fixture(gpu_lib):
    """Configuration 2 at its full size (NN=200, u/v/w, NM=10, 513^2 grid):
    moments and the bispectra of 27 entries at 300 sampled points (the four
    corners, |f1+f2| > fmax where the Classic S(f1+f2) is zero, the f1+f2 edge)
    against tests/golden/config2.npz (the factorised oracle, pinned to the
    reference harness on the small cases)."""
    g = load_golden("config2")
    p = synth.config2()
    assert p.nf == 513
    r = gpu_lib.classic_run(p, quad="vect")
    for k in ("m2mf", "m2mr", "m2o2mr", "m3mf", "m3mr"):
        assert rel_err(r[k], g[k]) < TOL, (k, rel_err(r[k], g[k]))
    s = gpu_lib.classic_bispectra_sel(p, g["sel"])
    i1, i2 = g["sample_i1"], g["sample_i2"]
    assert rel_err(s["bisp_f"][i1, i2, :], g["bisp_f_sample"]) < TOL
    assert rel_err(s["bisp_r"][i1, i2, :], g["bisp_r_sample"]) < TOL
    # entries with |f1+f2| > fmax are present and differ from the exact-sum values
    c = (p.nf - 1) // 2
    beyond = np.abs(i1 + i2 - 2 * c) > c
    assert beyond.sum() >= 70
    e = gpu_lib.classic_bispectra_sel(p, g["sel"][:3], exact=True)
    assert np.abs(e["bisp_f"][i1[beyond], i2[beyond], :] -
                  s["bisp_f"][i1[beyond], i2[beyond], :3]).max() > 0.0


def test_mode_subset_consistency_full_size(gpu_lib):
    """Deck of configuration 3 at full size (NN=1000, NM=30, NTC=2, 1025^2):
    a mode triple's third moment does not depend on the other modes, so a run
    restricted to modes {a,b,c} must reproduce those entries of the full run."""
    p = synth.config3()
    full = gpu_lib.classic_run(p, quad="vect")
    nm = 30
    sel = np.array([2, 11, 29], dtype=np.int32)          # 1-based kept-mode ids
    sub = gpu_lib.classic_run(p, quad="vect", bound=gpu_lib.Bound(p, modes=sel))
    F3 = full["m3mr"].reshape(nm, nm, nm, order="F")
    S3 = sub["m3mr"].reshape(3, 3, 3, order="F")
    idx = sel - 1
    want = F3[np.ix_(idx, idx, idx)]
    assert rel_err(S3, want) < TOL
    Ff = full["m3mf"].reshape(nm, nm, nm, order="F")
    Sf = sub["m3mf"].reshape(3, 3, 3, order="F")
    assert rel_err(Sf, Ff[np.ix_(idx, idx, idx)]) < TOL
    # force third moment is symmetric under (n,p) swap: B_mnp(f1,f2) = B_mpn(f2,f1)
    assert rel_err(Ff, np.transpose(Ff, (0, 2, 1))) < TOL
    assert np.all(np.isfinite(full["m3mr"])) and np.all(np.isfinite(full["m2mr"]))


@pytest.mark.parametrize("prec", ["fp64", "bf16x3"])
def test_plan_runs_repeat_and_match_one_shot(gpu_lib, prec):
    p = CASES["config1"]()
    one = gpu_lib.classic_run(p, quad="vect", prec=prec)
    plan = gpu_lib.Plan(p, quad="vect", prec=prec)
    a = plan.run()
    b = plan.run()
    plan.close()
    for k in ("m2mf", "m2mr", "m2o2mr", "m3mf", "m3mr"):
        assert rel_err(a[k], one[k]) < 1e-14, k
        assert rel_err(b[k], a[k]) < 1e-14, k


@pytest.mark.parametrize("cfg,prec,exact,quad", [
    ("config2", "fp64", False, "vect"), ("config2", "bf16x3", True, "vect"),
    ("config3", "fp64", False, "vect"), ("config3", "bf16x3", False, "vect"),
    ("config2", "fp64", False, "riemann"), ("config2", "bf16x3", False, "trapezoid")])
def test_specialised_epilogue_equals_generic(gpu_lib, cfg, prec, exact, quad):
    """Full-size grids (513^2, 1025^2): the epilogue's specialised paths for
    interior blocks, for blocks beyond the f1+f2 edge and for the band that
    straddles it (whose f1+f2 weight is the rule's edge weight at n) give the
    generic epilogue's moments to rounding (the same G values, other summation),
    under each quadrature rule."""
    p = getattr(synth, cfg)()
    b = gpu_lib.Bound(p)
    fast = gpu_lib.classic_run(p, quad=quad, prec=prec, exact=exact, bound=b, psd=False)
    slow = gpu_lib.classic_run(p, quad=quad, prec=prec, exact=exact, bound=b, psd=False,
                               epi_generic=True)
    for k in ("m3mf", "m3mr"):
        assert rel_err(fast[k], slow[k]) < 1e-13, k


def test_more_than_64_modes_matches_oracle(gpu_lib):
    """NM = 72 kept modes (the projection runs in chunks of 64 modes): moments
    against the factorised oracle on a small tower."""
    p = synth.tower(nn=8, nm=72, comps=(1, 3), natf=np.geomspace(0.1, 3.0, 72),
                    xi=np.full(72, 0.01), nfreqs=9, df=0.25, seed=11)
    r = gpu_lib.classic_run(p, quad="vect")
    assert r["m3mr"].size == 72 ** 3
    o = O.classic_factorised(p, rule="vect", truncate=True, nthreads=16)
    assert rel_err(r["m3mf"], o["m3mf"]) < TOL
    assert rel_err(r["m3mr"], o["m3mr"]) < TOL
    m = O.modal_psd(p)
    assert rel_err(r["m2mr"], m["m2mr"]) < TOL
