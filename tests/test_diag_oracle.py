"""The oracle's ONLY_DIAG restatement (oracle.classic_diag) against the
reference's own getFM_diag_tnlm_vect_cls_ / getRM_diag_vect_cls_ outputs
(golden *_diag arrays written by the reference harness) and their quadrature."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.cases import CASES
from tests.conftest import load_golden, rel_err


@pytest.mark.parametrize("name", list(CASES))
def test_classic_diag_matches_reference(name):
    p = CASES[name]()
    g = load_golden(name)
    o = O.classic_diag(p)
    assert rel_err(o["psd_f"], g["psd_f_diag"]) < 1e-13
    assert rel_err(o["psd_r"], g["psd_r_diag"]) < 1e-13
    if "bisp_f_diag" in g:
        assert rel_err(o["bisp_f"], g["bisp_f_diag"]) < 1e-13
        assert rel_err(o["bisp_r"], g["bisp_r_diag"]) < 1e-13
    else:
        i1, i2 = g["sample_i1"], g["sample_i2"]
        assert rel_err(o["bisp_f"][i1, i2, :], g["bisp_f_diag_sample"]) < 1e-13
        assert rel_err(o["bisp_r"][i1, i2, :], g["bisp_r_diag_sample"]) < 1e-13
    for k in ("m2mf", "m2mr", "m3mf", "m3mr"):
        assert rel_err(o[k], g[f"{k}_diag"]) < 1e-13, k


def test_diag_model_is_not_the_diagonal_of_the_full_model():
    """Coherence between nodes is dropped: with positively correlated nodes the
    full model's diagonal exceeds the diagonal model (which sums the nodes as
    independent)."""
    p = CASES["tiny_u"]()
    g = load_golden("tiny_u")
    nm = g["m2mf_diag"].size
    full_diag = g["m2mf"].reshape(nm, nm, order="F").diagonal()
    assert np.all(full_diag > g["m2mf_diag"])


@pytest.mark.parametrize("name", ["tiny_u", "tiny_uvw", "tiny_dav"])
def test_msh_points_diag_equals_reference_classic_diag(name):
    """The point form of the diagonal model (the Mesher's BFM) equals the
    reference's own Classic diagonal bispectrum at every grid point where
    f1 + f2 stays on the grid (there the Classic zero padding does not act)."""
    p = CASES[name]()
    g = load_golden(name)
    f = p.freqs()
    nf = f.size
    c = nf // 2
    ii, jj = np.meshgrid(np.arange(nf), np.arange(nf), indexing="ij")
    inside = np.abs(ii + jj - 2 * c) <= c
    b = O.msh_points_diag(p, f[ii[inside]], f[jj[inside]])          # (nm, npts)
    assert rel_err(b.T, g["bisp_f_diag"][inside]) < 1e-13
