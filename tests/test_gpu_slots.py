"""GPU parity of the operator-slot entry points against the reference's golden
spectra: bsahip_frf_apply (the response operators getRM_full_vect_cls_ /
getRM_full_scalar_cls_ / getBRM_msh, functionsImpl.F90:1031-1115, 1598-1738,
2329-2450) maps the reference's force spectra onto its response spectra, and
bsahip_bfm_points_cls (the scalar Classic BFM, functionsImpl.F90:2122-2318)
reproduces the reference's grid bispectra point by point, truncation
included.  The Fortran slots bound by bsa_Run (BsaLibData.F90:166-310) are
checked through the test driver against the same library calls.  The same for
the ONLY_DIAG model against the reference's golden diagonal spectra
(getFM_diag_tnlm_vect_cls_ / getRM_diag_vect_cls_), whose point form is also
held to the oracle's diagonal point bispectra (msh_points_diag).
Tolerance: 1e-12 normwise relative (FP64)."""
import numpy as np
import pytest

from bsalib_amd import abi
from oracle import oracle as O
from tests import fdriver
from tests.cases import CASES
from tests.conftest import load_golden, rel_err

pytestmark = pytest.mark.gpu

FULL_GRID = ["tiny_u", "tiny_uvw", "tiny_dav", "tiny_2dof"]
TOL = 1e-12


@pytest.mark.parametrize("name", FULL_GRID + ["config1"])
def test_frf_psd_maps_reference_force_psd_to_response(gpu_lib, name):
    g = load_golden(name)
    p = CASES[name]()
    f = p.freqs()
    out = gpu_lib.frf_apply(p, abi.FRF_PSD, g["psd_f"], f)
    assert rel_err(out, g["psd_r"]) < TOL


@pytest.mark.parametrize("name", FULL_GRID)
def test_frf_grid_maps_reference_bfm_to_brm(gpu_lib, name):
    g = load_golden(name)
    p = CASES[name]()
    f = p.freqs()
    out = gpu_lib.frf_apply(p, abi.FRF_GRID, g["bisp_f"], f, f)
    assert rel_err(out, g["bisp_r"]) < TOL


def test_frf_points_maps_reference_sample(gpu_lib):
    """config1's golden holds 300 sampled grid points of its bispectra."""
    g = load_golden("config1")
    p = CASES["config1"]()
    f = p.freqs()
    fi, fj = f[g["sample_i1"]], f[g["sample_i2"]]
    out = gpu_lib.frf_apply(p, abi.FRF_POINTS, g["bisp_f_sample"].T, fi, fj)
    assert rel_err(out, g["bisp_r_sample"].T) < TOL


@pytest.mark.parametrize("name", FULL_GRID)
def test_bfm_points_cls_equals_reference_grid(gpu_lib, name):
    """Every point of the grid, the corners where |f1 + f2| passes fmax (the
    Classic truncation) included."""
    g = load_golden(name)
    p = CASES[name]()
    f = p.freqs()
    nf = f.size
    I, J = np.meshgrid(np.arange(nf), np.arange(nf), indexing="ij")
    fi, fj = f[I.ravel(order="F")], f[J.ravel(order="F")]
    bfm, brm = gpu_lib.bfm_points_cls(p, fi, fj)
    nm3 = bfm.shape[0]
    want_f = g["bisp_f"].reshape(nf * nf, nm3, order="F").T
    want_r = g["bisp_r"].reshape(nf * nf, nm3, order="F").T
    assert rel_err(bfm, want_f) < TOL
    assert rel_err(brm, want_r) < TOL
    # outside the truncated corners the exact-sum points agree too
    inside = np.abs(fi + fj) <= f[-1] * (1 + 1e-12)
    assert (~inside).any()
    ex, _ = gpu_lib.bfm_points(p, fi[inside], fj[inside], with_brm=False)
    assert rel_err(ex, want_f[:, inside]) < TOL


def test_bfm_points_cls_sample_config1(gpu_lib):
    g = load_golden("config1")
    p = CASES["config1"]()
    f = p.freqs()
    bfm, brm = gpu_lib.bfm_points_cls(p, f[g["sample_i1"]], f[g["sample_i2"]])
    assert rel_err(bfm, g["bisp_f_sample"].T) < TOL
    assert rel_err(brm, g["bisp_r_sample"].T) < TOL


def test_frf_points_agrees_with_point_brm(gpu_lib):
    """FRF_POINTS over bsahip_bfm_points' BFM is its own BRM (off-grid points)."""
    p = CASES["tiny_uvw"]()
    rng = np.random.default_rng(3)
    fi = rng.uniform(-0.6, 0.6, 40)
    fj = rng.uniform(-0.6, 0.6, 40)
    bfm, brm = gpu_lib.bfm_points(p, fi, fj)
    assert rel_err(gpu_lib.frf_apply(p, abi.FRF_POINTS, bfm, fi, fj), brm) < TOL


def test_frf_apply_errors(gpu_lib):
    p = CASES["tiny_uvw"]()
    f = p.freqs()
    b = abi.Bound(p)
    nm = b.nm
    with pytest.raises(abi.BsaHipError) as e:
        gpu_lib.frf_apply(p, abi.FRF_PSD, np.zeros((f.size, nm * nm)), f, flags=abi.SUM_EXACT, bound=b)
    assert e.value.code == abi.EINVAL
    with pytest.raises(ValueError):        # diagonal model: NM columns, not NM^2
        gpu_lib.frf_apply(p, abi.FRF_PSD, np.zeros((f.size, nm * nm)), f, flags=abi.ONLY_DIAG, bound=b)
    with pytest.raises(abi.BsaHipError) as e:
        gpu_lib.frf_apply(p, 7, np.zeros(3), f, bound=b)
    assert e.value.code == abi.EINVAL
    # the full-model spectra entry refuses the diagonal flag instead of writing NM^3 entries
    nf = p.nf
    buf = np.zeros(nf * nf * nm ** 3)
    rc = abi.lib().bsahip_classic_spectra(abi.C.byref(b.struct), abi.C.c_int32(p.nfreqs),
                                          abi.C.c_double(p.df), abi.C.c_int32(abi.ONLY_DIAG),
                                          None, None, abi._ptr(buf, abi.C.c_double), None)
    assert rc == abi.EUNSUPPORTED
    # an empty list is a no-op
    out = gpu_lib.frf_apply(p, abi.FRF_POINTS, np.zeros((nm ** 3, 0)), np.zeros(0), np.zeros(0), bound=b)
    assert out.size == 0


@pytest.mark.parametrize("name", ["tiny_uvw", "config1"])
def test_fortran_operator_slots(gpu_lib, name):
    """bsa_Run binds the six slots; their vectorised outputs are the library's
    Classic spectra, the scalar and Mesher operators agree with them point by
    point (SLOT lines printed by the driver)."""
    p = CASES[name]()
    out, r = fdriver.run(p, "both", "slots")
    slot = {ln.split()[1]: float(ln.split()[2]) for ln in r.stdout.splitlines()
            if ln.startswith("SLOT")}
    for k in ("bfm_scalar", "brm_scalar", "psdf_scalar", "psdr_scalar", "bfm_msh", "brm_msh"):
        assert slot[k] < TOL, (k, slot[k])
    assert slot["msh_points"] > 0
    s = gpu_lib.classic_spectra(p)
    for k in ("psd_f", "psd_r", "bisp_f", "bisp_r"):
        assert rel_err(out["slot_" + k], s[k].ravel(order="F")) < TOL, k


# ---- the ONLY_DIAG model ----

@pytest.mark.parametrize("name", FULL_GRID + ["config1"])
def test_diag_psd_and_frf_psd_match_reference(gpu_lib, name):
    g = load_golden(name)
    p = CASES[name]()
    f = p.freqs()
    psd = gpu_lib.diag_psd_points(p, f)
    assert rel_err(psd.T, g["psd_f_diag"]) < TOL
    out = gpu_lib.frf_apply(p, abi.FRF_PSD, g["psd_f_diag"], f, flags=abi.ONLY_DIAG)
    assert rel_err(out, g["psd_r_diag"]) < TOL


@pytest.mark.parametrize("name", FULL_GRID)
def test_diag_grid_frf_and_points_match_reference(gpu_lib, name):
    g = load_golden(name)
    p = CASES[name]()
    f = p.freqs()
    nf = f.size
    out = gpu_lib.frf_apply(p, abi.FRF_GRID, g["bisp_f_diag"], f, f, flags=abi.ONLY_DIAG)
    assert rel_err(out, g["bisp_r_diag"]) < TOL
    I, J = np.meshgrid(np.arange(nf), np.arange(nf), indexing="ij")
    fi, fj = f[I.ravel(order="F")], f[J.ravel(order="F")]
    bfm, brm = gpu_lib.bfm_points_cls(p, fi, fj, only_diag=True)
    nm = bfm.shape[0]
    assert rel_err(bfm, g["bisp_f_diag"].reshape(nf * nf, nm, order="F").T) < TOL
    assert rel_err(brm, g["bisp_r_diag"].reshape(nf * nf, nm, order="F").T) < TOL


def test_diag_points_sample_config1(gpu_lib):
    g = load_golden("config1")
    p = CASES["config1"]()
    f = p.freqs()
    fi, fj = f[g["sample_i1"]], f[g["sample_i2"]]
    bfm, brm = gpu_lib.bfm_points_cls(p, fi, fj, only_diag=True)
    assert rel_err(bfm, g["bisp_f_diag_sample"].T) < TOL
    assert rel_err(brm, g["bisp_r_diag_sample"].T) < TOL
    out = gpu_lib.frf_apply(p, abi.FRF_POINTS, g["bisp_f_diag_sample"].T, fi, fj, flags=abi.ONLY_DIAG)
    assert rel_err(out, g["bisp_r_diag_sample"].T) < TOL


@pytest.mark.parametrize("name", ["tiny_uvw", "tiny_dav", "config1"])
def test_diag_points_exact_sum_match_oracle(gpu_lib, name):
    """bsahip_bfm_points with BSAHIP_ONLY_DIAG (S(f1+f2) evaluated beyond the
    grid) against the oracle's diagonal point bispectra."""
    p = CASES[name]()
    rng = np.random.default_rng(5)
    fmax = p.freqs()[-1]
    fi = rng.uniform(-fmax, fmax, 50)
    fj = rng.uniform(-fmax, fmax, 50)
    bfm, brm = gpu_lib.bfm_points(p, fi, fj, only_diag=True)
    want = O.msh_points_diag(p, fi, fj)
    assert rel_err(bfm, want) < TOL
    assert rel_err(gpu_lib.frf_apply(p, abi.FRF_POINTS, bfm, fi, fj, flags=abi.ONLY_DIAG), brm) < TOL


@pytest.mark.parametrize("name", ["tiny_uvw", "config1"])
def test_fortran_operator_slots_diag(gpu_lib, name):
    """The slots of the ONLY_DIAG model: vectorised outputs equal the reference's
    golden diagonal spectra, scalar and Mesher operators agree with them."""
    g = load_golden(name)
    p = CASES[name]()
    out, r = fdriver.run(p, "both", "diag", "slots")
    slot = {ln.split()[1]: float(ln.split()[2]) for ln in r.stdout.splitlines()
            if ln.startswith("SLOT")}
    for k in ("bfm_scalar", "brm_scalar", "psdf_scalar", "psdr_scalar", "bfm_msh", "brm_msh"):
        assert slot[k] < TOL, (k, slot[k])
    assert slot["msh_points"] > 0
    assert rel_err(out["slot_psd_f"], g["psd_f_diag"].ravel(order="F")) < TOL
    assert rel_err(out["slot_psd_r"], g["psd_r_diag"].ravel(order="F")) < TOL
    if "bisp_f_diag" in g:
        assert rel_err(out["slot_bisp_f"], g["bisp_f_diag"].ravel(order="F")) < TOL
        assert rel_err(out["slot_bisp_r"], g["bisp_r_diag"].ravel(order="F")) < TOL
    else:
        nf = p.nf
        bf = out["slot_bisp_f"].reshape(nf, nf, -1, order="F")
        assert rel_err(bf[g["sample_i1"], g["sample_i2"], :], g["bisp_f_diag_sample"]) < TOL
