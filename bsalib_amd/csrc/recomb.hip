// Post-processing of the modal moments on the device: the modal force PSD at
// arbitrary frequencies (for the background/resonant decomposition) and the
// recombination of modal second and third moments into nodal ones.

// ---------------------------------------------------------------------------
// Modal force PSD S_Q(n, p; f_q) = sum_k A(n, k; f_q) Psi_a(p, k) from the
// projected table U[k*ldU + n*nfl + q] (rows per mode = nfl).  The diagonal at
// f = f_n is getBR_SFm_val_ (src/bsa/functions/functionsImpl.F90:2650-2743).
// grid (ceil(nfl/64), nm*nm), block 64
// ---------------------------------------------------------------------------
__global__ void k_psd_points(const double *__restrict__ U, long ldU, const double *__restrict__ pa,
                             int K, int nm, int nfl, double *__restrict__ psd)
{
   const int q = blockIdx.x * 64 + threadIdx.x;
   const int np = blockIdx.y;
   if (q >= nfl) return;
   const int n = np % nm, p = np / nm;
   double s = 0.0;
   for (int k = 0; k < K; ++k) s = fma(U[(size_t)k * ldU + (size_t)n * nfl + q], pa[(size_t)p * K + k], s);
   psd[(size_t)q * nm * nm + np] = s;
}

// ---------------------------------------------------------------------------
// Nodal moments from modal ones, one thread per DOF d (modalRecombination,
// app/bsa.F90:935-1053), phi_m = Phi(d, m):
//   ccc = 0 (unilateral): m2 = sum_m m2(m,m) phi_m^2, m3 = sum_m m3(m,m,m) phi_m^3
//   ccc = 1 (complete cubic combination): m2 = sum_{j,k} m2(j,k) phi_j phi_k,
//           m3 = sum_{i,j,k} m3(i,j,k) phi_i phi_j phi_k
//   m2o2 as m2 (when given), sk = m3 / m2^(3/2).
// The products and the summation order are the reference's (k, j, i from the
// outside in; ((m3 phi_i) phi_j) phi_k), so the sums agree to rounding.  The
// thread's row of Phi is cached in LDS ([m][64]); the modal moments are
// wave-uniform reads.  grid ceil(ndofs/64), block 64, dynamic LDS nm*64*8.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_recomb(int ndofs, int nm, const double *__restrict__ phi,
                                                const double *__restrict__ m2,
                                                const double *__restrict__ m3,
                                                const double *__restrict__ m2o2, int ccc,
                                                double *__restrict__ m2_r, double *__restrict__ m3_r,
                                                double *__restrict__ m2o2_r,
                                                double *__restrict__ sk_r)
{
   extern __shared__ double ph[];          // [m][64]
   const int tid = threadIdx.x;
   const int d = blockIdx.x * 64 + tid;
   const bool ok = d < ndofs;
   for (int m = 0; m < nm; ++m) ph[m * 64 + tid] = ok ? phi[(size_t)m * ndofs + d] : 0.0;
   // each thread reads only its own LDS column: no barrier needed
   if (!ok) return;
   double s2 = 0.0, s3 = 0.0, s22 = 0.0;
   const size_t nm2 = (size_t)nm * nm;
   if (ccc == 0) {
      for (int m = 0; m < nm; ++m) {
         const double p1 = ph[m * 64 + tid];
         double r = p1 * p1;
         s2 += m2[(size_t)m * nm + m] * r;
         if (m2o2) s22 += m2o2[(size_t)m * nm + m] * r;
         r = r * p1;
         s3 += m3[(size_t)m * nm2 + (size_t)m * nm + m] * r;
      }
   } else {
      for (int k = 0; k < nm; ++k) {
         const double p3 = ph[k * 64 + tid];
         for (int j = 0; j < nm; ++j) {
            const double p2 = ph[j * 64 + tid];
            const size_t pm2 = (size_t)k * nm + j;
            s2 += m2[pm2] * p2 * p3;
            if (m2o2) s22 += m2o2[pm2] * p2 * p3;
            const double *m3jk = m3 + (size_t)k * nm2 + (size_t)j * nm;
            for (int i = 0; i < nm; ++i) s3 += m3jk[i] * ph[i * 64 + tid] * p2 * p3;
         }
      }
   }
   m2_r[d] = s2;
   m3_r[d] = s3;
   if (m2o2_r) m2o2_r[d] = s22;
   if (sk_r) sk_r[d] = s3 / pow(s2, 1.5);
}

extern "C" {

int bsahip_modal_psd_points(const bsahip_problem *p, int32_t npts, const double *f, double *psd)
{
   std::lock_guard<std::recursive_mutex> lk(g.mu);
   CHK(ensure_init());
   if (npts < 0 || (npts > 0 && (!f || !psd))) return fail(BSAHIP_EINVAL, "npts < 0 or NULL arrays");
   if (npts == 0) return 0;
   if (npts > 65535) return fail(BSAHIP_EUNSUPPORTED, "more than 65535 frequencies in one call");
   Bufs &B = oneshot_bufs();
   Prob D;
   CHK(load_problem(p, 0, B, D));
   std::vector<double> fl(npts);
   for (int q = 0; q < npts; ++q) fl[q] = std::fabs(f[q]);
   const long ldU = (long)D.nm * npts;
   double *dfl, *U, *out;
   CHK(upload(B, "fl", fl, &dfl));
   CHK(B.get("U", (size_t)D.K * ldU, &U));
   CHK(build_table(B, D, dfl, npts, npts, U, ldU));
   const size_t n = (size_t)D.nm * D.nm * npts;
   CHK(B.get("psd_pts", n, &out));
   dim3 gq((unsigned)((npts + 63) / 64), (unsigned)(D.nm * D.nm));
   k_psd_points<<<gq, 64, 0, g.s>>>(U, ldU, D.pa_k, D.K, D.nm, npts, out);
   HIPC(hipGetLastError());
   HIPC(hipMemcpyAsync(psd, out, n * sizeof(double), hipMemcpyDeviceToHost, g.s));
   HIPC(hipStreamSynchronize(g.s));
   return 0;
}

int bsahip_modal_recombination(int32_t ndofs, int32_t nm, const double *phi, const double *m2,
                               const double *m3, const double *m2o2, int32_t ccc, double *m2_r,
                               double *m3_r, double *m2o2_r, double *sk_r)
{
   std::lock_guard<std::recursive_mutex> lk(g.mu);
   CHK(ensure_init());
   if (ndofs < 0 || nm < 1 || nm > 64 || (ccc != 0 && ccc != 1))
      return fail(BSAHIP_EINVAL, "need ndofs >= 0, nm in 1..64 and ccc 0 or 1");
   if (ndofs == 0) return 0;
   if (!phi || !m2 || !m3 || !m2_r || !m3_r || (m2o2_r && !m2o2))
      return fail(BSAHIP_EINVAL, "NULL recombination array");
   Bufs &B = oneshot_bufs();
   const size_t nm2 = (size_t)nm * nm, nm3 = nm2 * nm;
   double *dphi, *dm2, *dm3, *dm22 = nullptr, *out;
   CHK(B.get("rc_phi", (size_t)ndofs * nm, &dphi));
   CHK(B.get("rc_m2", nm2, &dm2));
   CHK(B.get("rc_m3", nm3, &dm3));
   CHK(B.get("rc_out", 4 * (size_t)ndofs, &out));
   HIPC(hipMemcpyAsync(dphi, phi, (size_t)ndofs * nm * sizeof(double), hipMemcpyHostToDevice, g.s));
   HIPC(hipMemcpyAsync(dm2, m2, nm2 * sizeof(double), hipMemcpyHostToDevice, g.s));
   HIPC(hipMemcpyAsync(dm3, m3, nm3 * sizeof(double), hipMemcpyHostToDevice, g.s));
   if (m2o2) {
      CHK(B.get("rc_m2o2", nm2, &dm22));
      HIPC(hipMemcpyAsync(dm22, m2o2, nm2 * sizeof(double), hipMemcpyHostToDevice, g.s));
   }
   double *o2 = out, *o3 = out + ndofs, *o22 = out + 2 * (size_t)ndofs, *osk = out + 3 * (size_t)ndofs;
   k_recomb<<<(unsigned)((ndofs + 63) / 64), 64, (size_t)nm * 64 * sizeof(double), g.s>>>(
      ndofs, nm, dphi, dm2, dm3, dm22, ccc, o2, o3, dm22 ? o22 : nullptr, osk);
   HIPC(hipGetLastError());
   HIPC(hipMemcpyAsync(m2_r, o2, (size_t)ndofs * sizeof(double), hipMemcpyDeviceToHost, g.s));
   HIPC(hipMemcpyAsync(m3_r, o3, (size_t)ndofs * sizeof(double), hipMemcpyDeviceToHost, g.s));
   if (m2o2_r) HIPC(hipMemcpyAsync(m2o2_r, o22, (size_t)ndofs * sizeof(double), hipMemcpyDeviceToHost, g.s));
   if (sk_r) HIPC(hipMemcpyAsync(sk_r, osk, (size_t)ndofs * sizeof(double), hipMemcpyDeviceToHost, g.s));
   HIPC(hipStreamSynchronize(g.s));
   return 0;
}

}  // extern "C"
