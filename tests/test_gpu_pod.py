"""GPU parity of the Mesher's POD BFM (getFM_full_tm_scalar_msh_POD_,
functionsImpl.F90:474-1019) against the oracle's literal restatement
(oracle.msh_points_pod, eigendecompositions by numpy's LAPACK):

* without truncation the POD BFM is the direct point BFM with coherence column 1
  for every component (corr_col_mode 1) -- 1e-12;
* truncated by count or by threshold, the device path (rocSOLVER syevd + two
  FP64 GEMMs per matrix) -- 1e-10: the rank-K part of S(f) is defined up to the
  eigensolvers' rounding, which the gap lambda_K - lambda_K+1 amplifies; the
  test points are checked for a gap of at least 0.5 % of lambda_1 first;
* the Mesher run through the truncated POD equals the oracle's zone quadrature
  over the same tiles -- 1e-10."""
import copy
import functools

import numpy as np
import pytest

from bsalib_amd import abi
from oracle import oracle as O
from bsalib_amd import synth
from tests.conftest import rel_err

pytestmark = pytest.mark.gpu

FI = np.array([0.1, -0.3, 0.0, 0.25, 0.6, 0.45, -0.05])
FJ = np.array([0.2, 0.15, 0.4, -0.25, -0.35, 0.45, 0.3])


def _problem():
    p = synth.tower(nn=12, nm=3, comps=(1, 3), seed=8)
    c1 = p.nod_corr[:, 0].copy()
    p.nod_corr = np.asfortranarray(np.stack([c1, c1 ** 2, c1 ** 3], axis=1))   # T4 visible
    return p


def _min_gap(p, k_of, freqs):
    g = np.inf
    for f in freqs:
        for t in p.tc:
            lam = np.linalg.eigvalsh(O.nodal_psd_matrix(p, f, t))[::-1]
            k = k_of(lam)
            # a cut inside a cluster of ~zero eigenvalues (S(0) has rank 1) is harmless
            if k < lam.size and lam[k - 1] > 1e-10 * lam[0]:
                g = min(g, (lam[k - 1] - lam[k]) / lam[0])
    return g


def test_untruncated_pod_is_column_one_point_bfm(gpu_lib):
    p = _problem()
    bfm, _ = abi.bfm_points(p, FI, FJ, with_brm=False, bound=abi.Bound(p, corr_col_mode=1))
    assert rel_err(bfm, O.msh_points_pod(p, FI, FJ)) < 1e-12
    direct, _ = abi.bfm_points(p, FI, FJ, with_brm=False)          # column tc per component
    assert rel_err(direct, O.msh_points(p, FI, FJ)) < 1e-12
    assert rel_err(bfm, direct) > 1e-5                              # T4 is visible


@pytest.mark.parametrize("nmodes", [1, 2, 5])
def test_pod_truncated_by_count(gpu_lib, nmodes):
    p = _problem()
    freqs = np.abs(np.concatenate([FI, FJ, FI + FJ]))
    assert _min_gap(p, lambda lam: nmodes, freqs) > 5e-3
    b = abi.Bound(p, corr_col_mode=1, pod_nmodes=nmodes)
    bfm, brm = abi.bfm_points(p, FI, FJ, bound=b)
    ref = O.msh_points_pod(p, FI, FJ, nmodes=nmodes)
    assert rel_err(bfm, ref) < 1e-10
    assert rel_err(bfm, O.msh_points_pod(p, FI, FJ)) > 1e-4       # truncation is visible
    # BRM = BFM x Re[H_m(f1) H_n(f2) conj(H_p(f1+f2))]
    H1, H2, H3 = O.frf(p, FI), O.frf(p, FJ), O.frf(p, FI + FJ)
    R = np.einsum("qm,qn,qp->mnpq", H1, H2, np.conj(H3)).real.reshape(-1, FI.size, order="F")
    assert rel_err(brm, ref * R) < 1e-10


@pytest.mark.parametrize("lim", [0.05, 0.2])
def test_pod_truncated_by_threshold(gpu_lib, lim):
    p = _problem()
    freqs = np.abs(np.concatenate([FI, FJ, FI + FJ]))
    assert _min_gap(p, lambda lam: O.pod_threshold_count(lam, lim), freqs) > 5e-3
    b = abi.Bound(p, corr_col_mode=1, pod_trunc_lim=lim, pod_nmodes=3)   # threshold wins
    bfm, _ = abi.bfm_points(p, FI, FJ, with_brm=False, bound=b)
    assert rel_err(bfm, O.msh_points_pod(p, FI, FJ, lim=lim)) < 1e-10


def test_pod_truncation_needs_column_one(gpu_lib):
    p = _problem()
    with pytest.raises(abi.BsaHipError):
        abi.bfm_points(p, FI, FJ, bound=abi.Bound(p, corr_col_mode=0, pod_nmodes=2))
    with pytest.raises(abi.BsaHipError):
        abi.bfm_points(p, FI, FJ, bound=abi.Bound(p, corr_col_mode=1, pod_trunc_lim=-1.0))


def test_mesher_through_truncated_pod_matches_oracle(gpu_lib):
    p = _problem()
    b = abi.Bound(p, corr_col_mode=1, pod_nmodes=3)
    M = abi.Mesher(p, bkg_base_rfmnt=5, tile_cells=2, bkg_peakw=0.25, bound=b)
    geom, _ = M.tiles()
    lv = np.random.default_rng(5).choice([1, 2], size=M.ntiles).astype(np.int32)
    r = M.run(levels=lv)
    M.close()
    o = O.mesher_integrate(p, geom, lv, points=functools.partial(O.msh_points_pod, nmodes=3))
    assert rel_err(r["m3mf"], o["m3mf"]) < 1e-10
    assert rel_err(r["m3mr"], o["m3mr"]) < 1e-10
    # the same Mesher without truncation differs
    M2 = abi.Mesher(p, bkg_base_rfmnt=5, tile_cells=2, bkg_peakw=0.25,
                    bound=abi.Bound(p, corr_col_mode=1))
    r2 = M2.run(levels=lv)
    M2.close()
    assert rel_err(r2["m3mr"], r["m3mr"]) > 1e-4
