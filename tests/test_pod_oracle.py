"""Pins of the oracle's POD restatement (oracle.msh_points_pod,
getFM_full_tm_scalar_msh_POD_, functionsImpl.F90:474-1019) against facts it
must satisfy independently of its own formula:

* without truncation the eigenpair sums rebuild S(f), so the POD BFM equals the
  non-POD literal restatement (msh_points, a C transcription checked against
  the reference's golden vectors) once every component reads coherence
  column 1 (T4: the POD path passes idir = 1, functionsImpl.F90:628-629, 770);
* with full coherence S(f) = s s^T has rank one, so keeping one mode loses
  nothing;
* with truncation the blocks 5-2-2 / 2-2-5 / 2-5-2 equal the factorised
  three-term sum with S(f) replaced by its rank-K part V_K diag(lam_K) V_K^T
  (written here from the algebra, not from the oracle's loops);
* the threshold count follows getNPODModesByThreshold_ on hand-worked lists.
"""
import copy

import numpy as np
import pytest

from oracle import oracle as O
from bsalib_amd import synth

FI = np.array([0.1, -0.3, 0.0, 0.25, 0.6])
FJ = np.array([0.2, 0.15, 0.4, -0.25, -0.35])


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def _distinct_columns(p):
    """Coherence columns 2 and 3 differ from column 1 (rho^2, rho^3)."""
    q = copy.deepcopy(p)
    c1 = q.nod_corr[:, 0].copy()
    q.nod_corr = np.asfortranarray(np.stack([c1, c1 ** 2, c1 ** 3], axis=1))
    return q


def test_threshold_count_hand_worked():
    assert O.pod_threshold_count([4.0, 2.0, 1.0, 0.5], 0.3) == 2      # 2 >= 1.2 > 1
    assert O.pod_threshold_count([4.0, 2.0, 1.0, 0.5], 0.1) == 4      # all >= 0.4
    assert O.pod_threshold_count([4.0, 2.0, 1.0, 0.5], 0.5) == 2      # 2 >= 2
    assert O.pod_threshold_count([3.0, 3.0, 3.0], 0.5) == 1           # all equal
    assert O.pod_threshold_count([4.0, 2.0, 1.0], 1.5) == 1           # above the largest
    assert O.pod_threshold_count([4.0, 1.0, 1.0, 1.0], 0.25) == 4     # ties at the limit


def test_untruncated_pod_equals_nonpod_single_component():
    p = synth.tower(nn=10, nm=3, comps=(1,), seed=2)
    assert _rel(O.msh_points_pod(p, FI, FJ), O.msh_points(p, FI, FJ)) < 1e-12


def test_untruncated_pod_reads_column_one_for_every_component():
    p = _distinct_columns(synth.tower(nn=8, nm=3, comps=(1, 2, 3), seed=4))
    col1 = copy.deepcopy(p)
    col1.nod_corr = np.asfortranarray(np.stack([p.nod_corr[:, 0]] * 3, axis=1))
    pod = O.msh_points_pod(p, FI, FJ)
    assert _rel(pod, O.msh_points(col1, FI, FJ)) < 1e-12
    # and differs from the per-component columns the non-POD path reads
    assert _rel(pod, O.msh_points(p, FI, FJ)) > 1e-3


def test_rank_one_coherence_one_mode_is_exact():
    p = synth.tower(nn=9, nm=3, comps=(1, 3), seed=6)
    p.nod_corr[:] = 1.0
    full = O.msh_points(p, FI, FJ)
    assert _rel(O.msh_points_pod(p, FI, FJ, nmodes=1), full) < 1e-12
    assert _rel(O.msh_points_pod(p, FI, FJ, lim=0.5), full) < 1e-12


def _factorised_rank_k(p, fi, fj, rank):
    """B(m,n,o) = 2 sum_tc sum_i [Pq(m,i) A1(n,i) A2(o,i) + A3(m,i) Pq(n,i) A2(o,i)
    + A3(m,i) A1(n,i) Pq(o,i)], A = Pl S_K, S_K the rank-K part of S; rank is a
    number or a function of the decreasing eigenvalues."""
    P = O.psi(p)
    nm = P.shape[0]
    out = np.zeros((nm ** 3, fi.size))
    for q in range(fi.size):
        B = np.zeros((nm, nm, nm))
        for t in p.tc:
            Pl, Pq = P[:, :, t - 1], P[:, :, t + 2]
            A = []
            for f in (fi[q], fj[q], fi[q] + fj[q]):
                lam, V = np.linalg.eigh(O.nodal_psd_matrix(p, f, t))
                k = rank(lam[::-1]) if callable(rank) else rank
                Vk, lk = V[:, lam.size - k:], lam[lam.size - k:]
                A.append(Pl @ (Vk * lk) @ Vk.T)
            A1, A2, A3 = A
            B += 2.0 * (np.einsum("mi,ni,oi->mno", Pq, A1, A2)
                        + np.einsum("mi,ni,oi->mno", A3, Pq, A2)
                        + np.einsum("mi,ni,oi->mno", A3, A1, Pq))
        out[:, q] = B.reshape(-1, order="F")
    return out


@pytest.mark.parametrize("rank", [1, 3, 6])
def test_truncated_pod_equals_rank_k_factorised(rank):
    p = synth.tower(nn=10, nm=3, comps=(1, 3), seed=8)
    lit = O.msh_points_pod(p, FI, FJ, nmodes=rank)
    assert _rel(lit, _factorised_rank_k(p, FI, FJ, rank)) < 1e-12
    if rank < 6:
        assert _rel(lit, O.msh_points_pod(p, FI, FJ)) > 1e-4    # truncation is visible


@pytest.mark.parametrize("lim", [0.05, 0.2])
def test_threshold_truncation_keeps_per_matrix_counts(lim):
    p = synth.tower(nn=10, nm=3, comps=(1, 3), seed=9)
    counts = {O.pod_threshold_count(np.linalg.eigvalsh(O.nodal_psd_matrix(p, f, 1))[::-1], lim)
              for f in np.concatenate([FI, FJ, FI + FJ])}
    assert len(counts) > 1                        # counts differ between frequencies
    ref = _factorised_rank_k(p, FI, FJ, lambda d: O.pod_threshold_count(d, lim))
    assert _rel(O.msh_points_pod(p, FI, FJ, lim=lim), ref) < 1e-12
