"""GPU test of the library's FP64 strided-batched GEMM (k_dgemm_sb, the
kernel of the ONLY_DIAG model's T_m and of the truncated-POD tables) against
numpy on shapes that are not multiples of its 64 x 64 tile and 16-deep K
chunks, every transpose combination, a shared A (stride 0), beta != 0 and
leading dimensions above the row counts.  Tolerance 1e-13 normwise."""
import ctypes as C

import numpy as np
import pytest

from bsalib_amd import abi
from tests.conftest import rel_err

pytestmark = pytest.mark.gpu


def _gemm(ta, tb, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, Cm, ldc, sC, batch):
    L = abi.lib()
    D = C.POINTER(C.c_double)
    rc = L.bsahip_dgemm_strided_batched(
        C.c_int32(ta), C.c_int32(tb), C.c_int32(M), C.c_int32(N), C.c_int32(K), C.c_double(alpha),
        A.ctypes.data_as(D), C.c_int64(lda), C.c_int64(sA), B.ctypes.data_as(D), C.c_int64(ldb),
        C.c_int64(sB), C.c_double(beta), Cm.ctypes.data_as(D), C.c_int64(ldc), C.c_int64(sC),
        C.c_int32(batch))
    assert rc == 0, L.bsahip_last_error().decode()


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("M,N,K", [(70, 45, 37), (130, 129, 200), (1, 64, 17), (64, 64, 0)])
def test_dgemm_matches_numpy(gpu_lib, ta, tb, M, N, K):
    rng = np.random.default_rng(M * 1000 + N * 10 + K + 7 * ta + 3 * tb)
    batch = 3
    ra, ca = (K, M) if ta else (M, K)
    rb, cb = (N, K) if tb else (K, N)
    lda, ldb, ldc = ra + 3, rb + 1, M + 2
    sA = 0                                               # one A for the whole batch
    sB, sC = ldb * cb + 5, ldc * N + 7
    A = rng.standard_normal(lda * max(ca, 1))
    B = rng.standard_normal(sB * (batch - 1) + ldb * max(cb, 1))
    C0 = rng.standard_normal(sC * (batch - 1) + ldc * N)
    Cg = C0.copy()
    alpha, beta = 1.5, -0.5
    _gemm(ta, tb, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, Cg, ldc, sC, batch)
    Am = A[:lda * ca].reshape(ca, lda).T[:ra, :] if ca else np.zeros((ra, 0))
    opA = Am.T if ta else Am
    for b in range(batch):
        Bm = B[b * sB:b * sB + ldb * cb].reshape(cb, ldb).T[:rb, :] if cb else np.zeros((rb, 0))
        opB = Bm.T if tb else Bm
        c0 = C0[b * sC:b * sC + ldc * N].reshape(N, ldc).T[:M, :]
        want = alpha * (opA @ opB) + beta * c0
        got = Cg[b * sC:b * sC + ldc * N].reshape(N, ldc).T[:M, :]
        assert rel_err(got, want) < 1e-13, (b, rel_err(got, want))
        # entries outside the M x N block stay untouched
        full0 = C0[b * sC:b * sC + ldc * N].reshape(N, ldc).T
        fullg = Cg[b * sC:b * sC + ldc * N].reshape(N, ldc).T
        np.testing.assert_array_equal(fullg[M:, :], full0[M:, :])
