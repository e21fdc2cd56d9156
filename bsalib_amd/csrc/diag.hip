// The ONLY_DIAG Classic model on the device (getFM_diag_tnlm_vect_cls_,
// src/bsa/functions/functionsImpl.F90:1747-1970; getRM_diag_vect_cls_,
// :1975-2101): loaded nodes without coherence between them and only the
// diagonal modal entries.  Per mode m, with W_m(k) = Psi_q Psi_l^2 of the K
// index k = (component, node) and S_k the nodal PSD,
//   B_m(f1, f2) = 2 sum_k W_m(k) [S_k(f1) S_k(f2) + S_k(f1+f2) S_k(f2) + S_k(f1+f2) S_k(f1)]
//               = 2 [T_m(|f1|, |f2|) + T_m(|f1+f2|, |f2|) + T_m(|f1+f2|, |f1|)],
//   T_m(u, v)   = sum_k S_k(u) W_m(k) S_k(v)     (one FP64 GEMM per mode),
// with S(f1+f2) = 0 beyond +-fmax (the reference's zero padding), and
// psd_m(f) = sum_k Psi_l^2 S_k(f).  The moments integrate B_m and psd_m with
// the response factors Re[H_m(f1) H_m(f2) conj H_m(f1+f2)] and |H_m(f)|^2.

// X[k*nh + u] = S_k(u df); Y[(m*K + k)*nh + u] = W_m(k) S_k(u df).
// grid (ceil(nh/64), K), block 64
__global__ void k_diag_tables(const double *__restrict__ sq, int nh, int nnl, int kv,
                              const double *__restrict__ pa, const double *__restrict__ pb, int K,
                              int nm, double *__restrict__ X, double *__restrict__ Y)
{
   const int u = blockIdx.x * 64 + threadIdx.x;
   const int k = blockIdx.y;
   if (u >= nh) return;
   const int c = k / nnl, n = k % nnl;                // rows k >= kv are padding
   const double s = (k < kv) ? sq[((size_t)c * nh + u) * nnl + n] : 0.0;
   const double S = s * s;
   X[(size_t)k * nh + u] = S;
   for (int m = 0; m < nm; ++m) {
      const double a = pa[(size_t)m * K + k], q = pb[(size_t)m * K + k];
      Y[((size_t)m * K + k) * nh + u] = q * a * a * S;
   }
}

// Third moments of mode m = blockIdx.y over grid row i = blockIdx.x (threads
// over the columns j), quadrature weight qweight(rule); sums added to m3f[m],
// m3r[m].  T[m][u + v*nh] symmetric.  grid (nf, nm), block 256
__global__ __launch_bounds__(256) void k_diag_m3(const double *__restrict__ T, int nh,
                                                 const double2 *__restrict__ H, int nft, int rule,
                                                 double d2, double *__restrict__ m3f,
                                                 double *__restrict__ m3r)
{
   __shared__ double red[2][4];
   const int m = blockIdx.y;
   const int n = nh - 1;
   const int F1 = (int)blockIdx.x - n;
   const int a1 = abs(F1);
   const double *Tm = T + (size_t)m * nh * nh;
   const double2 h1 = hget(H, nft, m, F1);
   double sf = 0.0, sr = 0.0;
   for (int j = threadIdx.x; j < 2 * nh - 1; j += 256) {
      const int F2 = j - n, a2 = abs(F2), s = F1 + F2, as = abs(s);
      double b = Tm[a1 + (size_t)a2 * nh];
      if (as <= n) b += Tm[as + (size_t)a2 * nh] + Tm[as + (size_t)a1 * nh];
      b *= 2.0;
      const double w = qweight(F1, F2, n, rule, d2);
      const double r = re3(h1, hget(H, nft, m, F2), hget(H, nft, m, s));
      sf += w * b;
      sr += w * b * r;
   }
   sf = wave_sum(sf);
   sr = wave_sum(sr);
   const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
   if (lane == 0) {
      red[0][wv] = sf;
      red[1][wv] = sr;
   }
   __syncthreads();
   if (threadIdx.x == 0) {
      atomicAdd(m3f + m, red[0][0] + red[0][1] + red[0][2] + red[0][3]);
      atomicAdd(m3r + m, red[1][0] + red[1][1] + red[1][2] + red[1][3]);
   }
}

// Second moments of mode m = blockIdx.y (vectorised rule: every point d, the
// two ends d/2; +u and -u), psd_m(u) = sum_k Psi_l(m,k)^2 S_k(u).
// grid (ceil(nh/64), nm), block 64
__global__ void k_diag_m2(const double *__restrict__ X, int nh, const double *__restrict__ pa,
                          int K, const double2 *__restrict__ H, int nft, double d, double om_step,
                          double *__restrict__ m2f, double *__restrict__ m2r,
                          double *__restrict__ m2o)
{
   const int m = blockIdx.y;
   const int u = blockIdx.x * 64 + threadIdx.x;
   double sf = 0.0, sr = 0.0, so = 0.0;
   if (u < nh) {
      double v = 0.0;
      for (int k = 0; k < K; ++k) {
         const double a = pa[(size_t)m * K + k];
         v = fma(a * a, X[(size_t)k * nh + u], v);
      }
      double w = (u == nh - 1) ? 0.5 * d : d;
      if (u > 0) w *= 2.0;
      const double2 h = H[(size_t)m * nft + u];
      const double fr = h.x * h.x + h.y * h.y;
      const double om = om_step * u;
      sf = w * v;
      sr = w * v * fr;
      so = w * v * fr * om * om;
   }
   sf = wave_sum(sf);
   sr = wave_sum(sr);
   so = wave_sum(so);
   if (threadIdx.x == 0) {
      if (m2f) atomicAdd(m2f + m, sf);
      if (m2r) atomicAdd(m2r + m, sr);
      if (m2o) atomicAdd(m2o + m, so);
   }
}

// Diagonal bispectra of selected modes on the Classic grid from T, the terms
// of k_diag_m3: out[q1 + nf*q2 + nf*nf*s] for mode sel[s] (the reference's
// bisp(i, j, m) of getFM_diag_tnlm_vect_cls_ / getRM_diag_vect_cls_).
// grid (ceil(nf*nf/256), nsel), block 256
__global__ __launch_bounds__(256) void k_diag_bisp_sel(const double *__restrict__ T, int nh,
                                                       const double2 *__restrict__ H, int nft,
                                                       const int *__restrict__ sel,
                                                       double *__restrict__ bfm,
                                                       double *__restrict__ brm)
{
   const int n = nh - 1, nf = 2 * nh - 1;
   const long npts = (long)nf * nf;
   const long pt = (long)blockIdx.x * blockDim.x + threadIdx.x;
   if (pt >= npts) return;
   const int m = sel[blockIdx.y];
   const double *Tm = T + (size_t)m * nh * nh;
   const int F1 = (int)(pt % nf) - n, F2 = (int)(pt / nf) - n, s = F1 + F2;
   const int a1 = abs(F1), a2 = abs(F2), as = abs(s);
   double b = Tm[a1 + (size_t)a2 * nh];
   if (as <= n) b += Tm[as + (size_t)a2 * nh] + Tm[as + (size_t)a1 * nh];
   b *= 2.0;
   const size_t o = (size_t)blockIdx.y * npts + pt;
   if (bfm) bfm[o] = b;
   if (brm) brm[o] = b * re3(hget(H, nft, m, F1), hget(H, nft, m, F2), hget(H, nft, m, s));
}

// Diagonal-model force bispectrum at points, the Mesher's ONLY_DIAG BFM
// (getFM_diag_tnm_scalar_msh_, functionsImpl.F90:1123-1216, with S(f1+f2) in
// its second and third products as the Classic model has them):
//   out[slot[q]*nm + m] = 2 sum_k W_m(k) [S_k(f1) S_k(f2) + S_k(f1+f2) (S_k(f2) + S_k(f1))],
// W_m(k) = Psi_q Psi_l^2, S_k = sq^2 from the table sq[(c*nu + u)*nnl + n] at rows
// i1, i2, i3 (|f1|, |f2|, |f1+f2|).  grid (ceil(npts*nm/256)), block 256
__global__ __launch_bounds__(256) void k_diag_points(
   const double *__restrict__ sq, int nu, int nnl, int kv, const double *__restrict__ pa,
   const double *__restrict__ pb, int K, int nm, int npts, const int *__restrict__ i1,
   const int *__restrict__ i2, const int *__restrict__ i3, const long *__restrict__ slot,
   double *__restrict__ out)
{
   const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
   if (gid >= (long)npts * nm) return;
   const int q = (int)(gid / nm), m = (int)(gid % nm);
   const int a1 = i1[q], a2 = i2[q], a3 = i3[q];
   double s = 0.0;
   for (int k = 0; k < kv; ++k) {                      // rows k >= kv are padding
      const int c = k / nnl, n = k % nnl;
      const double *col = sq + (size_t)c * nu * nnl + n;
      const double r1 = col[(size_t)a1 * nnl], r2 = col[(size_t)a2 * nnl];
      const double r3 = a3 >= 0 ? col[(size_t)a3 * nnl] : 0.0;
      const double S1 = r1 * r1, S2 = r2 * r2, S3 = r3 * r3;
      const double a = pa[(size_t)m * K + k], b = pb[(size_t)m * K + k];
      s += b * a * a * (S1 * S2 + S3 * S2 + S3 * S1);
   }
   out[(size_t)slot[q] * nm + m] = 2.0 * s;
}
