// Truncated POD of the nodal wind PSD matrices for the point BFM.
//
// The reference's Mesher evaluates the force bispectrum through the
// eigendecomposition S(f) = V diag(lambda) V^T of the loaded-node PSD matrix
// S_ij(f) = sqrt(S_i S_j) rho_ij^|f| (getFullNodalPSD, src/wind/WindPSDImpl.F90:76-108,
// coherence column 1) at f1, f2 and f1+f2, summing over the kept eigenpairs
// (getFM_full_tm_scalar_msh_POD_, src/bsa/functions/functionsImpl.F90:474-1019).
// Each of its three blocks ("5-2-2", "2-2-5", "2-5-2", :840-997) is one term of
//   B(m,n,p) = 2 sum_i [Psi_b(m,i) A1(n,i) A2(p,i) + A3(m,i) Psi_b(n,i) A2(p,i)
//                       + A3(m,i) A1(n,i) Psi_b(p,i)]
// with A(f) = Psi_a S_K(f) and S_K = sum over kept p of lambda_p v_p v_p^T, so
// truncation only changes the projection tables: this file builds
// A_K(y, i; f) = (Psi_a V) diag(w) V^T (w = lambda on kept pairs, 0 elsewhere)
// for a list of |f| values, in the layout of the projected tables
// (U[(c*nnlp + i)*ldU + y*rows + q]); the point kernels are unchanged.
//
// Pipeline per chunk of (component, frequency) matrices: k_pod_smat builds
// the matrices, rocsolver_dsyevd_strided_batched decomposes them (ascending
// eigenvalues, orthonormal vectors), k_pod_weights applies the truncation
// rule, two strided-batched FP64 GEMMs form A_K, k_pod_scatter writes the
// table.

// S_b (column-major, nnl x nnl) for b = b0 + blockIdx.y, b = c*nfl + q.
// grid (ceil(nnl^2/256), nb), block 256
__global__ void k_pod_smat(const double *__restrict__ lnr,   // [c][nnl(j)][nnl(i)] ln rho
                           const double *__restrict__ sq,    // [c][nfl][nnl] sqrt(S_i)
                           const double *__restrict__ fabs_list, int nfl, int nnl, long b0,
                           double *__restrict__ S)
{
   const size_t nn2 = (size_t)nnl * nnl;
   const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
   if (e >= nn2) return;
   const long b = b0 + blockIdx.y;
   const int c = (int)(b / nfl), q = (int)(b % nfl);
   const int i = (int)(e % nnl), j = (int)(e / nnl);
   const double af = fabs_list[q];
   const double *s = sq + ((size_t)c * nfl + q) * nnl;
   // rho^0 == 1, also for rho == 0 (Fortran 0**0)
   const double coh = (af == 0.0) ? 1.0 : exp(af * lnr[((size_t)c * nnl + j) * nnl + i]);
   S[(size_t)blockIdx.y * nn2 + e] = coh * (s[i] * s[j]);
}

// Truncation weights.  lam[bl*n + p]: eigenvalues in ascending order.
// w[bl*n + p] = lam_p for the K largest, 0 for the others; kept[bl] = K.
//   lim > 0 : getNPODModesByThreshold_ (functionsImpl.F90:203-230) on the
//             values in decreasing order: K = 1 when all values are equal,
//             else 1 + the number of values after the largest that are
//             >= lim * largest (a leading run, since the values are sorted)
//   lim <= 0: K = min(nkeep, n)   (nmodes_POD_; the choice between the two
//             rules is functionsImpl.F90:803-823)
// Info from the eigensolver (info[bl] != 0) is passed on as kept[bl] = -1.
// grid (nb), block 64
__global__ void k_pod_weights(const double *__restrict__ lam, const int *__restrict__ info, int n,
                              int nkeep, double lim, double *__restrict__ w, int *__restrict__ kept)
{
   const int bl = blockIdx.x, lane = threadIdx.x;
   const double *l = lam + (size_t)bl * n;
   double *wb = w + (size_t)bl * n;
   int K = min(nkeep, n);
   if (lim > 0.0) {
      const double top = l[n - 1];
      const double limval = lim * top;
      int cnt = 0;
      for (int p = lane; p < n - 1; p += 64) cnt += (l[p] >= limval) ? 1 : 0;
#pragma unroll
      for (int o = 32; o >= 1; o >>= 1) cnt += __shfl_xor(cnt, o, 64);
      K = (l[0] == top) ? 1 : 1 + cnt;
   }
   for (int p = lane; p < n; p += 64) wb[p] = (p >= n - K) ? l[p] : 0.0;
   if (lane == 0) kept[bl] = (info[bl] != 0) ? -1 : K;
}

// T(y, p) *= w(p), T column-major nm x n per matrix.  grid (ceil(nm*n/256), nb)
__global__ void k_pod_scale(double *__restrict__ T, const double *__restrict__ w, int nm, int n)
{
   const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
   if (e >= (size_t)nm * n) return;
   const size_t o = (size_t)blockIdx.y * nm * n + e;
   T[o] *= w[(size_t)blockIdx.y * n + e / nm];
}

// A_b (column-major nm x nnl) into U[(c*nnlp + i)*ldU + y*rows + q].
// grid (ceil(nm*nnl/256), nb)
__global__ void k_pod_scatter(const double *__restrict__ A, int nm, int nnl, int nfl, long b0,
                              int nnlp, int rows, double *__restrict__ U, long ldU)
{
   const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
   if (e >= (size_t)nm * nnl) return;
   const long b = b0 + blockIdx.y;
   const int c = (int)(b / nfl), q = (int)(b % nfl);
   const int y = (int)(e % nm), i = (int)(e / nm);
   U[((size_t)c * nnlp + i) * ldU + (size_t)y * rows + q] = A[(size_t)blockIdx.y * nm * nnl + e];
}
