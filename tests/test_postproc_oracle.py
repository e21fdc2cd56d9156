"""Pins of the oracle's post-processing restatements:

* br_sfm (getBR_SFm_val_) against the reference's own modal force PSD on the
  Classic grid (golden psd_f, written by the reference harness), and against the
  factorised projection sum_j A(m, j; f) Psi_a(m, j);
* modal_recombination against moments of samples: for samples q of modal
  coordinates and x = Phi q, the complete cubic combination of the sample's
  modal moments equals the sample moments of x exactly (up to rounding); with
  Phi = I and with uncorrelated modes it reduces to the modal diagonal;
* peak_factors against Davenport's Gaussian peak factor, its Gaussian limit
  (sk = 0: all three equal) and the sk -> -sk mirror symmetry.
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.cases import CASES
from tests.conftest import load_golden


@pytest.mark.parametrize("name", list(CASES))
def test_br_sfm_equals_reference_modal_psd(name):
    p = CASES[name]()
    g = load_golden(name)
    modes = O.kept_modes(p)
    nm = len(modes)
    nf = g["psd_f"].shape[0]
    f = (np.arange(nf) - (nf - 1) // 2) * p.df
    scale = np.abs(g["psd_f"]).max()
    A = O.a_table(p, f)                         # (nm, nnl, ntc, nf)
    P = O.psi(p)
    for q in (0, nf // 3, (nf - 1) // 2, nf - 1):
        for im, m in enumerate(modes):
            v = O.br_sfm(p, f[q], int(m))
            assert abs(v - g["psd_f"][q, im * nm + im]) < 1e-13 * scale
            fact = sum(float(A[im, :, c, q] @ P[im, :, t - 1]) for c, t in enumerate(p.tc))
            assert abs(v - fact) < 1e-13 * scale


def _sample_moments(rng, n, nm):
    """Skewed correlated samples of modal coordinates and their exact sample moments."""
    z = rng.gamma(2.0, 1.0, size=(n, nm)) - 2.0
    q = z @ rng.normal(size=(nm, nm))
    m2 = np.einsum("si,sj->ij", q, q) / n
    m3 = np.einsum("si,sj,sk->ijk", q, q, q) / n
    return q, m2, m3


def test_complete_combination_equals_sample_moments():
    rng = np.random.default_rng(4)
    nm, nd = 4, 7
    q, m2, m3 = _sample_moments(rng, 2000, nm)
    phi = rng.normal(size=(nd, nm))
    x = q @ phi.T
    o = O.modal_recombination(phi, m2.reshape(-1, order="F"), m3.reshape(-1, order="F"),
                              m2.reshape(-1, order="F"))
    assert np.allclose(o["m2_full"], (x * x).mean(0), rtol=1e-12)
    assert np.allclose(o["m3_full"], (x ** 3).mean(0), rtol=1e-11, atol=1e-12 * np.abs(x ** 3).mean())
    assert np.allclose(o["sk_full"], (x ** 3).mean(0) / (x * x).mean(0) ** 1.5, rtol=1e-11)
    assert np.allclose(o["m2o2_full"], o["m2_full"], rtol=1e-14)


def test_identity_and_uncorrelated_modes_reduce_to_the_diagonal():
    rng = np.random.default_rng(5)
    nm = 3
    _, m2, m3 = _sample_moments(rng, 500, nm)
    o = O.modal_recombination(np.eye(nm), m2.reshape(-1, order="F"), m3.reshape(-1, order="F"))
    for k in ("m2_diag", "m2_full"):
        assert np.allclose(o[k], np.diag(m2), rtol=1e-14)
    for k in ("m3_diag", "m3_full"):
        assert np.allclose(o[k], [m3[i, i, i] for i in range(nm)], rtol=1e-14)
    # uncorrelated modes: only diagonal modal moments -> unilateral == complete
    m2d, m3d = np.diag(np.diag(m2)), np.zeros_like(m3)
    for i in range(nm):
        m3d[i, i, i] = m3[i, i, i]
    phi = rng.normal(size=(5, nm))
    o = O.modal_recombination(phi, m2d.reshape(-1, order="F"), m3d.reshape(-1, order="F"))
    assert np.allclose(o["m2_diag"], o["m2_full"], rtol=1e-13)
    assert np.allclose(o["m3_diag"], o["m3_full"], rtol=1e-12)


def test_peak_factors_gaussian_limit_davenport_and_symmetry():
    m2 = np.array([1.0, 4.0, 0.25])
    m2o2 = np.array([40.0, 90.0, 30.0])
    T = 600.0
    pg, pp, pn = O.peak_factors(m2, m2o2, T, sk=np.zeros(3))
    nu = np.sqrt(m2o2 / m2) / (2 * np.pi)                 # up-crossing rate
    b = np.sqrt(2 * np.log(nu * T))
    assert np.allclose(pg, b + 0.5772 / b, rtol=1e-15)       # Davenport (1964)
    assert np.allclose(pp, pg, rtol=1e-15) and np.allclose(pn, pg, rtol=1e-15)
    sk = np.array([0.3, -0.5, 1.2])
    _, p1, n1 = O.peak_factors(m2, m2o2, T, sk=sk)
    _, p2, n2 = O.peak_factors(m2, m2o2, T, sk=-sk)
    assert np.allclose(p1, n2, rtol=1e-14) and np.allclose(n1, p2, rtol=1e-14)
    # the tail on the side of the skewness is the heavier one
    assert np.all((p1 > n1) == (sk > 0))
    # early returns of the reference
    assert O.peak_factors(m2, np.zeros(3), T) is None
    assert O.peak_factors(m2, m2o2, 1.0) is None            # beta < 1: too short
