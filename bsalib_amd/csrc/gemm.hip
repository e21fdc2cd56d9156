// FP64 strided-batched GEMM on v_mfma_f64_16x16x4f64 for the dense
// contractions outside the G kernels: the ONLY_DIAG model's per-mode
// T_m = S W_m S^T (diag.hip) and the truncated-POD tables (Psi_a V) diag(w) V^T
// (pod.hip).
//
//   C_b = alpha op(A_b) op(B_b) + beta C_b,   b = blockIdx.z,
//   op(X) = X or X^T, column-major, A_b = A + b sA (sA = 0: one A for the batch).
//
// Workgroup tile 64 x 64 of C, 4 waves of 32 x 32 (2 x 2 blocks of 16 x 16),
// K staged in LDS in chunks of 16 (both operands as [k][row] so a lane's
// fragment reads are row-contiguous), single-buffered with the next chunk
// held in registers while the current one is multiplied.  Edges (M, N, K not
// multiples of the tile) are zero-padded in the staging.  One workgroup per
// (tile row, tile column, batch): grid (ceil(M/64), ceil(N/64), batch).
// v_mfma_f64_16x16x4f64: A lane l holds (row l%16, k l/16), B (k l/16, col
// l%16), D register t of lane l holds (row l/16 + 4t, col l%16).

#define GK 16

__global__ __launch_bounds__(256) void k_dgemm_sb(int transa, int transb, int M, int N, int K,
                                                   double alpha, const double *__restrict__ A,
                                                   long lda, long sA, const double *__restrict__ B,
                                                   long ldb, long sB, double beta,
                                                   double *__restrict__ C, long ldc, long sC)
{
   __shared__ double As[GK][64 + 1], Bs[GK][64 + 1];
   const int b = blockIdx.z;
   const int m0 = blockIdx.x * 64, n0 = blockIdx.y * 64;
   const double *Ab = A + (size_t)b * sA;
   const double *Bb = B + (size_t)b * sB;
   double *Cb = C + (size_t)b * sC;
   const int tid = threadIdx.x, lane = tid & 63;
   const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
   const int wr = w >> 1, wc = w & 1;

   // staging: 64 x GK elements per operand, 4 per thread; op(A)(i, k) with
   // i fastest when A is not transposed (coalesced), k fastest otherwise
   auto ldA = [&](int k0, double (&r)[4]) {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
         const int t = tid + 256 * e;
         int i, k;
         if (!transa) { i = t & 63; k = t >> 6; } else { k = t % GK; i = t / GK; }
         const int gi = m0 + i, gk = k0 + k;
         double v = 0.0;
         if (gi < M && gk < K) v = transa ? Ab[(size_t)gi * lda + gk] : Ab[(size_t)gk * lda + gi];
         r[e] = v;
      }
   };
   auto ldB = [&](int k0, double (&r)[4]) {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
         const int t = tid + 256 * e;
         int j, k;
         if (transb) { j = t & 63; k = t >> 6; } else { k = t % GK; j = t / GK; }
         const int gj = n0 + j, gk = k0 + k;
         double v = 0.0;
         if (gj < N && gk < K) v = transb ? Bb[(size_t)gk * ldb + gj] : Bb[(size_t)gj * ldb + gk];
         r[e] = v;
      }
   };
   auto stA = [&](const double (&r)[4]) {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
         const int t = tid + 256 * e;
         if (!transa) As[t >> 6][t & 63] = r[e];
         else As[t % GK][t / GK] = r[e];
      }
   };
   auto stB = [&](const double (&r)[4]) {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
         const int t = tid + 256 * e;
         if (transb) Bs[t >> 6][t & 63] = r[e];
         else Bs[t % GK][t / GK] = r[e];
      }
   };

   d4 acc[2][2];
#pragma unroll
   for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = (d4){0.0, 0.0, 0.0, 0.0};
   double ra[4], rb[4];
   ldA(0, ra);
   ldB(0, rb);
   for (int k0 = 0; k0 < K; k0 += GK) {
      __syncthreads();
      stA(ra);
      stB(rb);
      __syncthreads();
      if (k0 + GK < K) {
         ldA(k0 + GK, ra);
         ldB(k0 + GK, rb);
      }
#pragma unroll
      for (int kk = 0; kk < GK; kk += 4) {
         const int kr = kk + (lane >> 4);
         const double a0 = As[kr][wr * 32 + (lane & 15)];
         const double a1 = As[kr][wr * 32 + 16 + (lane & 15)];
         const double b0 = Bs[kr][wc * 32 + (lane & 15)];
         const double b1 = Bs[kr][wc * 32 + 16 + (lane & 15)];
         acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
         acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
         acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
         acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
      }
   }
#pragma unroll
   for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
         const int col = n0 + wc * 32 + 16 * j + (lane & 15);
         if (col >= N) continue;
#pragma unroll
         for (int t = 0; t < 4; ++t) {
            const int row = m0 + wr * 32 + 16 * i + (lane >> 4) + 4 * t;
            if (row >= M) continue;
            double *c = Cb + (size_t)col * ldc + row;
            const double v = alpha * acc[i][j][t];
            *c = (beta == 0.0) ? v : v + beta * *c;
         }
      }
}
