"""CPU tests: the oracle against the reference's golden vectors, and the checks
that pin each oracle function independently of its own formula."""
import math

import numpy as np
import pytest
from scipy import integrate, special

from bsalib_amd import synth
from oracle import oracle as O
from tests.cases import CASES
from tests.conftest import load_golden, rel_err

GOLDEN_CASES = ["tiny_u", "tiny_uvw", "tiny_dav", "config1", "tiny_2dof"]


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_generators_reproduce_golden_inputs(name):
    g = load_golden(name)
    p = CASES[name]()
    np.testing.assert_array_equal(p.phi, g["phi"])
    np.testing.assert_array_equal(p.natf, g["natf"])
    assert float(p.nod_corr.sum()) == float(g["nod_corr_sum"])


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_psd_table_matches_reference(name):
    g = load_golden(name)
    p = CASES[name]()
    assert rel_err(O.suvw_table(p, p.freqs()), g["suvw"]) < 1e-14


@pytest.mark.parametrize("name", ["tiny_u", "tiny_uvw", "tiny_dav"])
def test_classic_vect_matches_reference_full_arrays(name):
    g = load_golden(name)
    r = O.classic_vect(CASES[name]())
    for k in ("psd_f", "psd_r", "bisp_f", "bisp_r", "m2mf", "m2mr", "m3mf", "m3mr"):
        assert rel_err(r[k], g[k]) < 1e-13, k


def test_config1_classic_matches_reference():
    g = load_golden("config1")
    r = O.classic_vect(CASES["config1"](), nthreads=8)
    i1, i2 = g["sample_i1"], g["sample_i2"]
    assert rel_err(r["bisp_f"][i1, i2, :], g["bisp_f_sample"]) < 1e-13
    assert rel_err(r["bisp_r"][i1, i2, :], g["bisp_r_sample"]) < 1e-13
    for k in ("psd_f", "psd_r", "m2mf", "m2mr", "m3mf", "m3mr"):
        assert rel_err(r[k], g[k]) < 1e-13, k


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_factorised_oracle_matches_reference_moments(name):
    g = load_golden(name)
    fz = O.classic_factorised(CASES[name](), rule="vect", truncate=True, nthreads=8)
    assert rel_err(fz["m3mf"], g["m3mf"]) < 1e-12
    assert rel_err(fz["m3mr"], g["m3mr"]) < 1e-12
    mp = O.modal_psd(CASES[name]())
    for k in ("m2mf", "m2mr", "psd_f", "psd_r"):
        assert rel_err(mp[k], g[k]) < 1e-13, k


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_mesher_point_bfm_matches_reference(name):
    g = load_golden(name)
    bm = O.msh_points(CASES[name](), g["msh_fi"], g["msh_fj"])
    assert rel_err(bm, g["bfm_msh"]) < 1e-13


# ---------------------------------------------------------------------------
# independent pins
# ---------------------------------------------------------------------------
def test_corr_index_is_the_packed_upper_triangle():
    nn = 7
    seen = {}
    for i in range(1, nn + 1):
        for j in range(i, nn + 1):
            seen[(i, j)] = int(synth.corr_index(i, j, nn))
            assert int(synth.corr_index(j, i, nn)) == seen[(i, j)]
            assert O.lib().orc_corr_index(i, j, nn) == seen[(i, j)]
    # bijection onto 1..nn(nn+1)/2, row-major over the upper triangle
    assert sorted(seen.values()) == list(range(1, nn * (nn + 1) // 2 + 1))
    assert seen[(1, 1)] == 1 and seen[(1, nn)] == nn and seen[(2, 2)] == nn + 1


def _single_node(psd_id, comp=1):
    return synth.tower(nn=1, nm=1, comps=(comp,), natf=[0.5], xi=[0.02], nfreqs=9, df=0.05,
                       psd_id=psd_id, seed=0)


def test_von_karman_u_integrates_to_variance():
    # int_0^inf 4 s^2 (L/U) (1 + 70.7 (f L/U)^2)^(-5/6) df
    #   = s^2 * 4/sqrt(70.7) * (sqrt(pi)/2) Gamma(1/3)/Gamma(5/6)
    p = _single_node(1)
    sig = p.turb_std[0, 0]
    val, _ = integrate.quad(lambda f: O.eval_psd(p, [f], 1, pulsation=False)[0, 0], 0, np.inf,
                            limit=400)
    closed = sig ** 2 * 4 / math.sqrt(70.7) * math.sqrt(math.pi) / 2 * \
        special.gamma(1 / 3) / special.gamma(5 / 6)
    assert abs(val / closed - 1) < 1e-7
    assert abs(closed / sig ** 2 - 1) < 1e-3      # the 70.7 constant normalises to ~sigma^2
    # pulsation convention divides by 4 pi
    assert O.eval_psd(p, [0.3], 1)[0, 0] * 4 * math.pi == pytest.approx(
        O.eval_psd(p, [0.3], 1, pulsation=False)[0, 0], rel=1e-15)


def test_von_karman_vw_closed_form_at_zero_and_variance():
    p = _single_node(1, comp=3)
    sig, L, U = p.turb_std[2, 0], p.turb_scales[2, 0, 0], p.u_node[0]
    assert O.eval_psd(p, [0.0], 3, pulsation=False)[0, 0] == pytest.approx(4 * sig ** 2 * L / U, rel=1e-14)
    val, _ = integrate.quad(lambda f: O.eval_psd(p, [f], 3, pulsation=False)[0, 0], 0, np.inf,
                            limit=400)
    assert abs(val / sig ** 2 - 1) < 2e-2


def test_davenport_uliege_integrates_to_variance_exactly():
    # int_0^inf (2/3) f (L/U)^2 s^2 (1 + (f L/U)^2)^(-4/3) df = s^2 exactly
    p = _single_node(5)
    sig = p.turb_std[0, 0]
    val, _ = integrate.quad(lambda f: O.eval_psd(p, [f], 1, pulsation=False)[0, 0], 0, np.inf,
                            limit=400)
    assert val == pytest.approx(sig ** 2, rel=1e-8)


def test_frf_closed_forms():
    p = _single_node(1)
    H0 = O.frf(p, [0.0])[0, 0]
    assert H0 == pytest.approx(1 / p.Km[0], rel=1e-15)
    fn = p.natf[0]
    Hn = O.frf(p, [fn])[0, 0]
    wn = 2 * math.pi * fn
    assert Hn == pytest.approx(-1j / (p.Cm[0, 0] * wn), rel=1e-12)     # K = w^2 M at resonance
    assert O.frf(p, [-0.3])[0, 0] == pytest.approx(np.conj(O.frf(p, [0.3])[0, 0]), rel=1e-15)


def test_quadrature_weights_by_hand():
    nf, d = 7, 2 * math.pi * 0.1
    f = np.arange(-(nf // 2), nf // 2 + 1) * 0.1
    one2 = np.ones((nf, 1), order="F")
    one3 = np.ones((nf, nf, 1), order="F")
    m2 = np.zeros(1)
    m3 = np.zeros(1)
    O._intg(f, one2, m2, one3, m3)
    assert m2[0] == pytest.approx(d * (nf - 1), rel=1e-14)
    # interior (nf-2)^2, 4(nf-2) edges at 1/2, 4 corners at 1 - 3/2 = -1/2
    assert m3[0] == pytest.approx(d * d * ((nf - 2) ** 2 + 2 * (nf - 2) - 2), rel=1e-14)
    # a corner delta exposes the corner weight alone
    b = np.zeros((nf, nf, 1), order="F")
    b[0, 0, 0] = 1.0
    O._intg(f, one2, m2, b, m3)
    assert m3[0] == pytest.approx(-0.5 * d * d, rel=1e-14)
    w = O.intg_weights_2d(nf, d, "vect")
    assert w.sum() == pytest.approx(d * d * ((nf - 2) ** 2 + 2 * (nf - 2) - 2), rel=1e-14)


def test_single_node_single_mode_closed_form():
    p = _single_node(1)
    r = O.classic_vect(p)
    P = O.psi(p)
    pa, pb = P[0, 0, 0], P[0, 0, 3]
    f = p.freqs()
    S = O.eval_psd(p, f, 1)[:, 0]
    nf = f.size
    c = nf // 2
    Sp = np.zeros(3 * nf)
    Sp[nf:2 * nf] = S
    for a in (0, 3, c, nf - 1):
        for b in (0, 2, c, nf - 1):
            s3 = Sp[nf + a + b - c]                    # zero beyond the grid (Classic)
            want = 2 * pb * pa * pa * (S[a] * S[b] + S[b] * s3 + S[a] * s3)
            assert r["bisp_f"][a, b, 0] == pytest.approx(want, rel=1e-14, abs=0)


def test_gaussian_limit_gives_zero_third_moment():
    p = synth.tower(nn=5, nm=2, comps=(1,), nfreqs=7, df=0.15, seed=3)
    p.wfc[:, 3:, :] = 0.0
    r = O.classic_vect(p)
    assert np.abs(r["bisp_f"]).max() == 0.0
    assert np.abs(r["m3mr"]).max() == 0.0


def test_bispectrum_swap_and_even_symmetry():
    p = synth.tower(nn=5, nm=3, comps=(1, 2), nfreqs=9, df=0.1, seed=4)
    r = O.classic_vect(p)
    B = r["bisp_f"].reshape(p.nf, p.nf, 3, 3, 3, order="F")
    c = p.nf // 2
    for a in range(p.nf):
        for b in range(p.nf):
            # B_mnp(f1,f2) = B_mpn(f2,f1) everywhere (the three node sums of
            # functionsImpl.F90:1523-1528 map onto each other under this swap)
            np.testing.assert_allclose(B[a, b], np.transpose(B[b, a], (0, 2, 1)), rtol=1e-13, atol=0)
            # B(-f1,-f2) = B(f1,f2) everywhere (the truncation is symmetric too)
            np.testing.assert_allclose(B[a, b], B[2 * c - a, 2 * c - b], rtol=1e-13, atol=0)


def test_full_coherence_gives_rank_one_projection():
    p = synth.tower(nn=6, nm=3, comps=(1,), nfreqs=5, df=0.2, seed=2)
    p.nod_corr[:] = 1.0
    A = O.a_table(p, p.freqs())[:, :, 0, :]             # (nm, nnl, nf)
    P = O.psi(p)
    for q in range(p.nf):
        s = np.sqrt(O.eval_psd(p, [p.freqs()[q]], 1)[0])
        want = np.outer(P[:, :, 0] @ s, s)            # rho = 1: S = s s^T
        np.testing.assert_allclose(A[:, :, q], want, rtol=1e-13)


def test_bsacl_semantics_equal_exact_riemann_for_one_zone():
    p = synth.tower(nn=5, nm=3, comps=(1,), nfreqs=8, df=0.12, seed=5, psd_id=5,
                    zone_per_node=False)
    fx = O.classic_factorised(p, rule="riemann", truncate=False)
    hb = O.bsacl_run(p, p.freqs(), p.freqs())
    assert rel_err(hb, fx["m3mr"]) < 1e-13


def test_skewness_definition():
    nm = 2
    m2 = np.array([4.0, 0.3, 0.3, 9.0])
    m3 = np.arange(1, 9, dtype=float)
    sk = O.skewness(m2, m3, nm).reshape(-1, order="F")
    # sk(m,n,p) = m3 / (s_m s_n s_p), s = (2, 3)
    s = [2.0, 3.0]
    k = 0
    for pp in range(nm):
        for n in range(nm):
            for m in range(nm):
                assert sk[k] == pytest.approx(m3[k] / (s[m] * s[n] * s[pp]), rel=1e-15)
                k += 1
