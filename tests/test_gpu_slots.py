"""GPU parity of the operator-slot entry points against the reference's golden
spectra: bsahip_frf_apply (the response operators getRM_full_vect_cls_ /
getRM_full_scalar_cls_ / getBRM_msh, functionsImpl.F90:1031-1115, 1598-1738,
2329-2450) maps the reference's force spectra onto its response spectra, and
bsahip_bfm_points_cls (the scalar Classic BFM, functionsImpl.F90:2122-2318)
reproduces the reference's grid bispectra point by point, truncation
included.  The Fortran slots bound by bsa_Run (BsaLibData.F90:166-310) are
checked through the test driver against the same library calls.
Tolerance: 1e-12 normwise relative (FP64)."""
import numpy as np
import pytest

from bsalib_amd import abi
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
        gpu_lib.frf_apply(p, abi.FRF_PSD, np.zeros((f.size, nm * nm)), f, flags=abi.ONLY_DIAG, bound=b)
    assert e.value.code == abi.EUNSUPPORTED
    with pytest.raises(abi.BsaHipError) as e:
        gpu_lib.frf_apply(p, 7, np.zeros(3), f, bound=b)
    assert e.value.code == abi.EINVAL
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
