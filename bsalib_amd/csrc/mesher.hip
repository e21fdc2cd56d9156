// Zone Mesher on the device: adaptive (f1,f2) quadrature of the modal force and
// response bispectra over rectangular tiles, with all zone data resident in HBM.
//
// The reference's Mesher (src/bsa/meshing/BsaMesherImpl.F90, zones/*, policy/MPolicy.F90)
// samples the smooth modal force bispectrum (BFM) on coarse zone grids, refines
// the sampling where its policy asks for it, and integrates the response
// bispectrum (BRM = BFM * FRF product) on a finer grid where the BFM is
// interpolated bilinearly and the FRF evaluated exactly, with trapezoid weights
// (vertex 1/4, edge 1/2, interior 1; zones/_shared_poly2d.fi:827-832).  Here:
//
//  * 1D zone limits per axis as prefetchZoneLimits_ (BsaMesherImpl.F90:1638-1859):
//    background half-width b = bkg_peakw * bkg_area_ext / 2, peak zones
//    f_n +- xsi_n f_n peak_area_ext (getActualPeakZoneExtensionLimits_, :1864-1890),
//    pre-peak gaps between them, external padding up to max f_n * max_area_ext
//    (getMaxSpaceExtension_, :1626-1631); mirrored to negative frequencies.
//  * Rectangular zones are the tensor products of those 1D intervals, cut into
//    tiles of at most tile_cells x tile_cells level-0 cells.  Level-0 spacing is
//    the background spacing df_ref = 2b/(bkg_base_rfmnt-1) times the interval's
//    delta factor (BKG 1, PEAK 1, PRE_PEAK 2, PAD 4).
//  * A tile's BFM is sampled at refinement level l in {1,2,4,8} (l cells per
//    level-0 cell and axis).  The BRM grid of a tile does not depend on l: its
//    spacing follows the FRF (msh_brm_sub: 16 BRM cells per peak width at a
//    resonance line, growing with the distance to it), so the level only
//    decides how densely the smooth BFM is sampled.
//  * The rule-based level (the reference's delta validation,
//    _shared_poly2d.fi:735-795) asks for BFM spacing <= w/fct per axis
//    (w = narrowest interest-mode peak width 2 xsi f peak_area_ext in a peak
//    interval, bkg_peakw elsewhere; fct = the policy's BFM factor), capped at 8.
//
// Each tile's BFM points are evaluated by k_bfm_batch (FP64 MFMA, S(f1+f2)
// evaluated exactly, the semantics of getFM_full_tnm_scalar_msh_,
// functionsImpl.F90:247-465); k_msh_integrate interpolates, applies the FRF
// and the trapezoid weights and adds into m3mf, m3mr.

// ---------------------------------------------------------------------------
// V table for point evaluation: V[(u*K + k)*NMP + y] = A(y, k; |f_u|), zero for
// y >= nm.  Built from the U table U[k*ldU + y*nu + u].
// grid (ceil(nu/64), K), block 256
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_vtable(const double *__restrict__ U, long ldU, int nu,
                                                int nm, int K, int nmp, double *__restrict__ V)
{
   __shared__ double s[64][65];
   const int k = blockIdx.y;
   const int u0 = blockIdx.x * 64;
   const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
   for (int y0 = 0; y0 < nmp; y0 += 64) {
      for (int y = y0 + ty; y < y0 + 64; y += 4) {
         const int u = u0 + tx;
         double v = 0.0;
         if (y < nm && u < nu) v = U[(size_t)k * ldU + (size_t)y * nu + u];
         s[y - y0][tx] = v;
      }
      __syncthreads();
      const int ny = min(64, nmp - y0);
      for (int t = threadIdx.x; t < 64 * ny; t += 256) {
         const int i = t / ny, y = t % ny;
         const int u = u0 + i;
         if (u < nu) V[((size_t)u * K + k) * nmp + y0 + y] = s[y][i];
      }
      __syncthreads();
   }
}

// Psi_b transposed and padded: PB[k*NMP + y]
__global__ void k_pbt(const double *__restrict__ pb, int nm, int K, int nmp, double *__restrict__ PB)
{
   const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
   if (gid >= (long)K * nmp) return;
   const int k = (int)(gid / nmp), y = (int)(gid % nmp);
   PB[gid] = (y < nm) ? pb[(size_t)y * K + k] : 0.0;
}

// ---------------------------------------------------------------------------
// BFM at a batch of points on FP64 MFMA (v_mfma_f64_16x16x4f64).
// For one point, with A1 = A(f1), A2 = A(f2), A3 = A(f1+f2):
//   B(m,n,p)/2 = sum_k Psi_b(m,k) A1(n,k) A2(p,k)
//              + sum_k A3(m,k) [Psi_b(n,k) A2(p,k) + Psi_b(p,k) A1(n,k)]
// i.e. one NM x NM^2 product with a 2K-long contraction whose right operand
// is formed on the fly.  Rows m are padded to MT*16, columns (n,p) (col =
// n + p*nm) tiled by 16 and dealt to the 8 waves (TPW tiles each).
// One workgroup per point; out[slot[pt]*nm3 + m + col*nm] = B.
// ---------------------------------------------------------------------------
#define BKC 16   // K chunk staged in LDS

template <int MT, int TPW>
__global__ __launch_bounds__(512) void k_bfm_batch(
   const double *__restrict__ V, const double *__restrict__ PB, int K, int nm, int npts,
   const int *__restrict__ ia1, const int *__restrict__ ia2, const int *__restrict__ ia3,
   const long *__restrict__ slot, double *__restrict__ out)
{
   constexpr int NMP = MT * 16;
   __shared__ double s1[BKC][NMP], s2[BKC][NMP], s3[BKC][NMP], sp[BKC][NMP];
   const int pt = blockIdx.x;
   if (pt >= npts) return;
   const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
   const int ncol = nm * nm, ct = (ncol + 15) / 16;
   const double *v1 = V + (size_t)ia1[pt] * K * NMP;
   const double *v2 = V + (size_t)ia2[pt] * K * NMP;
   const double *v3 = V + (size_t)ia3[pt] * K * NMP;

   // columns beyond nm^2 compute garbage that is never stored (MFMA columns
   // are independent), so they just read mode 0
   int cn[TPW], cp[TPW];
#pragma unroll
   for (int j = 0; j < TPW; ++j) {
      const int col = (wv + 8 * j) * 16 + (lane & 15);
      const bool ok = col < ncol;
      cn[j] = ok ? col % nm : 0;
      cp[j] = ok ? col / nm : 0;
   }
   d4 acc[MT][TPW];
#pragma unroll
   for (int r = 0; r < MT; ++r)
#pragma unroll
      for (int j = 0; j < TPW; ++j) acc[r][j] = (d4){0.0, 0.0, 0.0, 0.0};

   const int kr = lane >> 4, rr = lane & 15;
   for (int k0 = 0; k0 < K; k0 += BKC) {
      for (int t = threadIdx.x; t < BKC * NMP; t += 512) {
         const int kk = t / NMP, y = t % NMP;
         const size_t o = (size_t)(k0 + kk) * NMP + y;
         s1[kk][y] = v1[o];
         s2[kk][y] = v2[o];
         s3[kk][y] = v3[o];
         sp[kk][y] = PB[o];
      }
      __syncthreads();
#pragma unroll
      for (int k4 = 0; k4 < BKC; k4 += 4) {
         const int kk = k4 + kr;
         double la[MT], lb[MT];
#pragma unroll
         for (int r = 0; r < MT; ++r) {
            la[r] = sp[kk][r * 16 + rr];
            lb[r] = s3[kk][r * 16 + rr];
         }
#pragma unroll
         for (int j = 0; j < TPW; ++j) {
            if (wv + 8 * j < ct) {
               const int n = cn[j], p = cp[j];
               const double a1n = s1[kk][n], a2p = s2[kk][p];
               const double pbn = sp[kk][n], pbp = sp[kk][p];
               const double ra = a1n * a2p;
               const double rb = pbn * a2p + pbp * a1n;
#pragma unroll
               for (int r = 0; r < MT; ++r) {
                  acc[r][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(la[r], ra, acc[r][j], 0, 0, 0);
                  acc[r][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(lb[r], rb, acc[r][j], 0, 0, 0);
               }
            }
         }
      }
      __syncthreads();
   }
   const long nm3 = (long)nm * ncol;
   double *o = out + slot[pt] * nm3;
#pragma unroll
   for (int j = 0; j < TPW; ++j) {
      const int col = (wv + 8 * j) * 16 + (lane & 15);
      if (wv + 8 * j >= ct || col >= ncol) continue;
#pragma unroll
      for (int r = 0; r < MT; ++r)
#pragma unroll
         for (int t = 0; t < 4; ++t) {
            const int m = r * 16 + (lane >> 4) + 4 * t;
            if (m < nm) o[m + (size_t)col * nm] = 2.0 * acc[r][j][t];
         }
   }
}

// FRF table H[i*nm + m] = 1/(K_m - w^2 M_m + i w C_mm), w = 2 pi (f0 + i h)
__global__ void k_frf_axis(double f0, double h, int n, const double *__restrict__ Mm,
                           const double *__restrict__ Km, const double *__restrict__ Cm, int nm,
                           double2 *__restrict__ H)
{
   const int t = blockIdx.x * blockDim.x + threadIdx.x;
   if (t >= n * nm) return;
   const int i = t / nm, m = t % nm;
   const double w = 2.0 * M_PI * (f0 + i * h);
   const double r = Km[m] - w * w * Mm[m], im = w * Cm[m], d = r * r + im * im;
   H[t] = make_double2(r / d, -im / d);
}

// Entry e of the per-point bispectrum vector: the triple (m, n, p) of
// m + n*nm + p*nm^2 (full model), or (e, e, e) (ONLY_DIAG model, nm entries).
__device__ __forceinline__ void msh_entry(int e, int nm, int diag, int &m, int &n, int &p)
{
   if (diag) {
      m = n = p = e;
   } else {
      m = e % nm;
      n = (e / nm) % nm;
      p = e / (nm * nm);
   }
}

// ---------------------------------------------------------------------------
// The post-mesh of many tiles in one launch each (bsahip_mesher_run):
// k_frf_segs fills the FRF tables of every tile, k_msh_copy_pts places the
// resident level-0 BFM into the fine grids of refined tiles, k_msh_integrate
// runs the quadrature over a list of (tile, cell range) items.
// ---------------------------------------------------------------------------
// FRF segment s: n points f0 + i h, out H[off + i*nm + m]; segment of thread t
// by binary search over the entry offsets eoff (n*nm per segment).
struct MshSeg {
   double f0, h;
   long off;
   const double *Mm, *Km, *Cm;   // modal mass, stiffness, damping of the segment's structure
   int nm;
};

__global__ void k_frf_segs(const MshSeg *__restrict__ seg, const long *__restrict__ eoff, int nseg,
                           double2 *__restrict__ H)
{
   const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
   if (t >= eoff[nseg]) return;
   int lo = 0, hi = nseg - 1;                      // eoff[lo] <= t < eoff[lo + 1]
   while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (eoff[mid] <= t) lo = mid;
      else hi = mid - 1;
   }
   const MshSeg S = seg[lo];
   const long e = t - eoff[lo];
   const long i = e / S.nm;
   const int m = (int)(e % S.nm);
   const double w = 2.0 * M_PI * (S.f0 + (double)i * S.h);
   const double r = S.Km[m] - w * w * S.Mm[m], im = w * S.Cm[m], d = r * r + im * im;
   H[S.off + e] = make_double2(r / d, -im / d);
}

// fine[dst[q]*nm3 + e] = lv0[src[q]*nm3 + e]; grid (ceil(nm3/256), npairs)
__global__ void k_msh_copy_pts(const double *__restrict__ lv0, const long *__restrict__ src,
                               const long *__restrict__ dst, int nm3, double *__restrict__ fine)
{
   const int e = blockIdx.x * blockDim.x + threadIdx.x;
   if (e >= nm3) return;
   const long q = blockIdx.y;
   fine[dst[q] * nm3 + e] = lv0[src[q] * nm3 + e];
}

// One tile of a batched post-mesh (tiles of several structures may share a
// launch: the modal data and accumulators are per tile).
struct MshTileRun {
   const double *bf;           // the tile's BFM grid (lv0 or fine)
   double *m3f, *m3r;          // force accumulator of its structure, response of this tile
   const double *Mm, *Km, *Cm; // modal mass, stiffness, damping of its structure
   long ldq;
   long h1, h2, h3;            // FRF table offsets (h3 < 0: f1 + f2 evaluated here)
   int nm, nb1, nb2, s1, s2;
   double f3a, dh1, dh2, ws;
};

// Tile quadrature: bilinear BFM on the BRM grid, exact FRF, trapezoid
// weights.  A tile's BFM grid is nb1 x nb2 (point q = i + j*ldq, values
// bf[q*nm3 + mnp]); each BFM cell holds s1 x s2 BRM cells.  BRM point (a, b):
// f1 = f1a + a*dh1, f2 = f2a + b*dh2.  Adds sum w*B into m3f and
// sum w*B*Re[H_m(f1)H_n(f2)H_p*(f1+f2)] into the tile's m3r.
// grid (nitems * bx), block 256: block -> (item blockIdx.x / bx, entries
// (blockIdx.x % bx) * 256 ..); item q covers cells [c0, c1) of tile it[q]
__global__ __launch_bounds__(256) void k_msh_integrate(
   const MshTileRun *__restrict__ tr, const int *__restrict__ it, const int2 *__restrict__ cr,
   int bx, int diag, const double2 *__restrict__ H)
{
   const long q = blockIdx.x / bx;
   const MshTileRun T = tr[it[q]];
   const int nm = T.nm;
   const int nm3 = diag ? nm : nm * nm * nm;
   const int mnp = (int)(blockIdx.x % bx) * blockDim.x + threadIdx.x;
   if (mnp >= nm3) return;
   const int c0 = cr[q].x, c1 = cr[q].y;
   int m, n, p;
   msh_entry(mnp, nm, diag, m, n, p);
   const int nc1 = T.nb1 - 1, nc2 = T.nb2 - 1;
   const int NA = nc1 * T.s1, NB = nc2 * T.s2;
   const double Mp = T.Mm[p], Kp = T.Km[p], Cp = T.Cm[p];
   const double inv1 = 1.0 / T.s1, inv2 = 1.0 / T.s2;
   const double2 *H1 = H + T.h1, *H2 = H + T.h2, *H3 = T.h3 >= 0 ? H + T.h3 : nullptr;
   double sf = 0.0, sr = 0.0;
   for (int c = c0; c < c1; ++c) {
      const int ci = c % nc1, cj = c / nc1;
      const size_t q00 = (size_t)ci + (size_t)cj * T.ldq;
      const double b00 = T.bf[q00 * nm3 + mnp], b10 = T.bf[(q00 + 1) * nm3 + mnp];
      const double b01 = T.bf[(q00 + T.ldq) * nm3 + mnp], b11 = T.bf[(q00 + T.ldq + 1) * nm3 + mnp];
      const int ea = (ci == nc1 - 1) ? T.s1 : T.s1 - 1;   // owned BRM points per axis
      const int eb = (cj == nc2 - 1) ? T.s2 : T.s2 - 1;
      for (int bb = 0; bb <= eb; ++bb) {
         const int b = cj * T.s2 + bb;
         const double wb = (b == 0 || b == NB) ? 0.5 : 1.0;
         const double sv = bb * inv2;
         const double2 h2v = H2[(size_t)b * nm + n];
         const double e0 = (1.0 - sv) * b00 + sv * b01, e1 = (1.0 - sv) * b10 + sv * b11;
         for (int aa = 0; aa <= ea; ++aa) {
            const int a = ci * T.s1 + aa;
            const double wa = (a == 0 || a == NA) ? 0.5 : 1.0;
            const double t = aa * inv1;
            const double bv = (1.0 - t) * e0 + t * e1;
            const double2 h1v = H1[(size_t)a * nm + m];
            double2 h3v;
            if (H3) {
               h3v = H3[(size_t)(a + b) * nm + p];
            } else {
               const double w3 = 2.0 * M_PI * (T.f3a + a * T.dh1 + b * T.dh2);
               const double r = Kp - w3 * w3 * Mp, im = w3 * Cp, d = r * r + im * im;
               h3v = make_double2(r / d, -im / d);
            }
            const double w = wa * wb;
            sf += w * bv;
            sr += w * bv * re3(h1v, h2v, h3v);
         }
      }
   }
   atomicAdd(T.m3f + mnp, sf * T.ws);
   atomicAdd(T.m3r + mnp, sr * T.ws);
}

// The same quadrature for the full model (nm^3 entries), one thread per
// (m, p) and MSH_NQ consecutive n: Re[H_m(f1) H_n(f2) H_p*(f1+f2)] is
// Re[H_n(f2) Q] with Q = H_m(f1) conj H_p(f1+f2) shared by the thread's n, and
// the sum over a BRM row (fixed b) is taken on Q before the H_n(f2) factor:
//   sum_a w_a w_b B Re[H_n Q] = w_b Re[H_n sum_a w_a B Q].
// Along the row B is linear in t = (a - a0) / s1, B = (1 - t) e0 + t e1 with
// e0, e1 per n, so sum_a w_a B Q = e0 sum_a w_a (1 - t) Q + e1 sum_a w_a t Q:
// the point loop forms the two n-independent sums, each n uses them once per
// row.  grid (nitems * bx), bx covering nm^2 ceil(nm / MSH_NQ) threads of the
// item's structure, block 256.
#ifndef MSH_NQ
#define MSH_NQ 8
#endif

__global__ __launch_bounds__(256) void k_msh_integrate_full(
   const MshTileRun *__restrict__ tr, const int *__restrict__ it, const int2 *__restrict__ cr,
   int bx, const double2 *__restrict__ H)
{
   const long q = blockIdx.x / bx;
   const MshTileRun T = tr[it[q]];
   const int nm = T.nm;
   const int nm2 = nm * nm, nm3 = nm2 * nm;
   const int nthr = nm2 * ((nm + MSH_NQ - 1) / MSH_NQ);
   const int th = (int)(blockIdx.x % bx) * blockDim.x + threadIdx.x;
   if (th >= nthr) return;
   const int m = th % nm, p = (th / nm) % nm, n0 = (th / nm2) * MSH_NQ;
   const int nn = min(MSH_NQ, nm - n0);
   const int c0 = cr[q].x, c1 = cr[q].y;
   const int nc1 = T.nb1 - 1, nc2 = T.nb2 - 1;
   const int NA = nc1 * T.s1, NB = nc2 * T.s2;
   const double Mp = T.Mm[p], Kp = T.Km[p], Cp = T.Cm[p];
   const double inv1 = 1.0 / T.s1, inv2 = 1.0 / T.s2;
   const double2 *H1 = H + T.h1, *H2 = H + T.h2, *H3 = T.h3 >= 0 ? H + T.h3 : nullptr;
   int e[MSH_NQ];
#pragma unroll
   for (int j = 0; j < MSH_NQ; ++j) e[j] = m + min(n0 + j, nm - 1) * nm + p * nm2;
   double sf[MSH_NQ], sr[MSH_NQ];
#pragma unroll
   for (int j = 0; j < MSH_NQ; ++j) sf[j] = sr[j] = 0.0;
   for (int c = c0; c < c1; ++c) {
      const int ci = c % nc1, cj = c / nc1;
      const size_t q00 = (size_t)ci + (size_t)cj * T.ldq;
      double b00[MSH_NQ], b10[MSH_NQ], b01[MSH_NQ], b11[MSH_NQ];
#pragma unroll
      for (int j = 0; j < MSH_NQ; ++j) {
         b00[j] = T.bf[q00 * nm3 + e[j]];
         b10[j] = T.bf[(q00 + 1) * nm3 + e[j]];
         b01[j] = T.bf[(q00 + T.ldq) * nm3 + e[j]];
         b11[j] = T.bf[(q00 + T.ldq + 1) * nm3 + e[j]];
      }
      const int ea = (ci == nc1 - 1) ? T.s1 : T.s1 - 1;   // owned BRM points per axis
      const int eb = (cj == nc2 - 1) ? T.s2 : T.s2 - 1;
      for (int bb = 0; bb <= eb; ++bb) {
         const int b = cj * T.s2 + bb;
         const double wb = (b == 0 || b == NB) ? 0.5 : 1.0;
         const double sv = bb * inv2;
         // along the row the BFM is e0 + t (e1 - e0): the sums over a of
         // w_a (1 - t) Q and w_a t Q do not depend on n, so they are taken
         // once per point and combined with each n's e0, e1 once per row
         double W0 = 0.0, W1 = 0.0, A0r = 0.0, A0i = 0.0, A1r = 0.0, A1i = 0.0;
         for (int aa = 0; aa <= ea; ++aa) {
            const int a = ci * T.s1 + aa;
            const double wa = (a == 0 || a == NA) ? 0.5 : 1.0;
            const double t = aa * inv1;
            const double2 h1v = H1[(size_t)a * nm + m];
            double2 h3v;
            if (H3) {
               h3v = H3[(size_t)(a + b) * nm + p];
            } else {
               const double w3 = 2.0 * M_PI * (T.f3a + a * T.dh1 + b * T.dh2);
               const double r = Kp - w3 * w3 * Mp, im = w3 * Cp, d = r * r + im * im;
               h3v = make_double2(r / d, -im / d);
            }
            // Q = h1 conj(h3)
            const double Qr = h1v.x * h3v.x + h1v.y * h3v.y;
            const double Qi = h1v.y * h3v.x - h1v.x * h3v.y;
            const double w0 = wa * (1.0 - t), w1 = wa * t;
            W0 += w0;
            W1 += w1;
            A0r += w0 * Qr;
            A0i += w0 * Qi;
            A1r += w1 * Qr;
            A1i += w1 * Qi;
         }
#pragma unroll
         for (int j = 0; j < MSH_NQ; ++j) {
            const double e0 = (1.0 - sv) * b00[j] + sv * b01[j];
            const double e1 = (1.0 - sv) * b10[j] + sv * b11[j];
            const double Dr = e0 * A0r + e1 * A1r, Di = e0 * A0i + e1 * A1i;
            const double2 h2v = H2[(size_t)b * nm + min(n0 + j, nm - 1)];
            sf[j] += wb * (e0 * W0 + e1 * W1);
            sr[j] += wb * (h2v.x * Dr - h2v.y * Di);
         }
      }
   }
#pragma unroll
   for (int j = 0; j < MSH_NQ; ++j)
      if (j < nn) {
         atomicAdd(T.m3f + e[j], sf[j] * T.ws);
         atomicAdd(T.m3r + e[j], sr[j] * T.ws);
      }
}

// ---------------------------------------------------------------------------
// Tile export: the integrand of k_msh_integrate without the weights at every
// BRM point (a, b) of the tile, pt = a + b*(NA+1).  Column c of val
// (val[pt*ncol + c]) is BRM entry sel[c] when sel[c] >= 0, else
// sum_e coef[c*nc + e] BRM_e over the nc entries.  grid (ceil(npts/256)), block 256
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_msh_export(
   const double *__restrict__ bf, long ldq, int nm, int nc, int diag, int nb1, int nb2, int s1,
   int s2, const double2 *__restrict__ H1, const double2 *__restrict__ H2,
   const double2 *__restrict__ H3, double f3a, double h1, double h2,
   const double *__restrict__ Mm, const double *__restrict__ Km, const double *__restrict__ Cm,
   int ncol, const int *__restrict__ sel, const double *__restrict__ coef,
   double *__restrict__ val)
{
   const int nm3 = nc;
   const int nc1 = nb1 - 1, nc2 = nb2 - 1;
   const long NA1 = (long)nc1 * s1 + 1, NB1 = (long)nc2 * s2 + 1;
   const long pt = (long)blockIdx.x * blockDim.x + threadIdx.x;
   if (pt >= NA1 * NB1) return;
   const int a = (int)(pt % NA1), b = (int)(pt / NA1);
   const int ci = min(a / s1, nc1 - 1), cj = min(b / s2, nc2 - 1);
   const int aa = a - ci * s1, bb = b - cj * s2;
   const size_t q00 = (size_t)ci + (size_t)cj * ldq;
   const double inv1 = 1.0 / s1, inv2 = 1.0 / s2;
   const double sv = bb * inv2, t = aa * inv1;
   auto brm = [&](int mnp) {
      int m, n, p;
      msh_entry(mnp, nm, diag, m, n, p);
      const double b00 = bf[q00 * nm3 + mnp], b10 = bf[(q00 + 1) * nm3 + mnp];
      const double b01 = bf[(q00 + ldq) * nm3 + mnp], b11 = bf[(q00 + ldq + 1) * nm3 + mnp];
      const double e0 = (1.0 - sv) * b00 + sv * b01, e1 = (1.0 - sv) * b10 + sv * b11;
      const double bv = (1.0 - t) * e0 + t * e1;
      const double2 h1v = H1[(size_t)a * nm + m], h2v = H2[(size_t)b * nm + n];
      double2 h3v;
      if (H3) {
         h3v = H3[(size_t)(a + b) * nm + p];
      } else {
         const double w3 = 2.0 * M_PI * (f3a + a * h1 + b * h2);
         const double r = Km[p] - w3 * w3 * Mm[p], im = w3 * Cm[p], d = r * r + im * im;
         h3v = make_double2(r / d, -im / d);
      }
      return bv * re3(h1v, h2v, h3v);
   };
   for (int c = 0; c < ncol; ++c) {
      double v;
      if (sel[c] >= 0) {
         v = brm(sel[c]);
      } else {
         v = 0.0;
         const double *cc = coef + (size_t)c * nm3;
         for (int mnp = 0; mnp < nm3; ++mnp)
            if (cc[mnp] != 0.0) v += cc[mnp] * brm(mnp);
      }
      val[(size_t)pt * ncol + c] = v;
   }
}

// Feature stencil of a tile's level-0 BFM: out[s] = sqrt(sum_mnp B^2) at the
// level-0 point nearest to stencil node s (ns x ns nodes over the tile).
// All tiles in one launch: tile k (blockIdx.y) has its level-0 values at
// lv0 + t0[k] * nm3 (row stride ldq) on an nb1[k] x nb2[k] grid; out[k*ns*ns + s].
__global__ void k_bfm_stencil(const double *__restrict__ lv0, const long *__restrict__ t0,
                              const int2 *__restrict__ nb, long ldq, int nm3, int ns,
                              double *__restrict__ out)
{
   const int s = blockIdx.x;
   const long k = blockIdx.y;
   const int si = s % ns, sj = s / ns;
   const int nb1 = nb[k].x, nb2 = nb[k].y;
   const int i = (int)lround((double)si * (nb1 - 1) / (ns - 1));
   const int j = (int)lround((double)sj * (nb2 - 1) / (ns - 1));
   const double *b = lv0 + ((size_t)t0[k] + (size_t)i + (size_t)j * ldq) * nm3;
   double acc = 0.0;
   for (int t = threadIdx.x; t < nm3; t += 64) acc += b[t] * b[t];
   acc = wave_sum(acc);
   if (threadIdx.x == 0) out[k * ns * ns + s] = sqrt(acc);
}

// ===========================================================================
// host side
// ===========================================================================
namespace {

// policy ids of the reference (src/bsa/meshing/policy/MPolicy.F90:10-19)
enum { POL_PRE_PEAK_1 = 3, POL_PRE_PEAK_2 = 4, POL_PEAK = 5, POL_PAD_EXTERN = 9, POL_BKG = 100 };

struct MshInterval {
   double a, b;      // [a, b] in Hz
   int pol;
   double w;         // narrowest interest-mode peak width of a PEAK interval, else
                     // bkg_peakw (pre-peak gaps validate against the background
                     // width, _shared_poly2d.fi:748-750)
};

// BFM factor of a policy (MPolicy_fromID bfm_pol_%i_fct_; the background zone
// carries MPolicy_PEAK, BsaMesherImpl.F90:390)
int msh_bfm_fct(int pol)
{
   switch (pol) {
   case POL_PRE_PEAK_1: return 4;
   case POL_PRE_PEAK_2: return 2;
   case POL_PEAK: return 4;
   case POL_PAD_EXTERN: return 1;
   default: return 4;   // background
   }
}

// level-0 spacing factor of an interval relative to the background spacing
double msh_delta_fct(int pol)
{
   switch (pol) {
   case POL_PRE_PEAK_1:
   case POL_PRE_PEAK_2: return 2.0;
   case POL_PAD_EXTERN: return 4.0;
   default: return 1.0;
   }
}

// Positive-axis limits as prefetchZoneLimits_ (BsaMesherImpl.F90:1638-1859) for
// modes sorted by natural frequency; returns intervals from b up to the
// maximum extension.
int msh_axis(const bsahip_mesher_cfg *c, int nm, std::vector<MshInterval> &out, double &b,
             double &fmax)
{
   std::vector<int> ord(nm);
   for (int i = 0; i < nm; ++i) ord[i] = i;
   std::sort(ord.begin(), ord.end(), [&](int x, int y) { return c->natf[x] < c->natf[y]; });
   b = 0.5 * c->bkg_peakw * c->bkg_area_ext;
   double fn_max = 0.0;
   for (int i = 0; i < nm; ++i) fn_max = std::max(fn_max, c->natf[i]);
   fmax = fn_max * c->max_area_ext;
   out.clear();
   std::vector<double> lo(nm), hi(nm), wd(nm);
   for (int q = 0; q < nm; ++q) {
      const int i = ord[q];
      const double e = c->xsi[i] * c->natf[i] * c->peak_area_ext;
      lo[q] = c->natf[i] - e;
      hi[q] = c->natf[i] + e;
      wd[q] = 2.0 * e;
   }
   int q = 0;
   while (q < nm && hi[q] <= b) ++q;              // peaks entirely inside the background zone
   if (q < nm) {
      double cur_hi;
      if (b < lo[q]) {
         out.push_back({b, lo[q], POL_PRE_PEAK_1, c->bkg_peakw});
         out.push_back({lo[q], hi[q], POL_PEAK, wd[q]});
      } else {
         out.push_back({b, hi[q], POL_PEAK, wd[q]});
      }
      cur_hi = hi[q];
      for (++q; q < nm; ++q) {
         if (lo[q] < cur_hi) {                     // overlap: extend the current peak zone
            MshInterval &pk = out.back();
            pk.b = std::max(pk.b, hi[q]);
            pk.w = std::min(pk.w, wd[q]);
            cur_hi = pk.b;
         } else {
            out.push_back({cur_hi, lo[q], POL_PRE_PEAK_2, c->bkg_peakw});
            out.push_back({lo[q], hi[q], POL_PEAK, wd[q]});
            cur_hi = hi[q];
         }
      }
      if (fmax > cur_hi) out.push_back({cur_hi, fmax, POL_PAD_EXTERN, c->bkg_peakw});
   } else if (fmax > b) {
      out.push_back({b, fmax, POL_PAD_EXTERN, c->bkg_peakw});
   }
   for (auto &iv : out)
      if (!(iv.b > iv.a)) return fail(BSAHIP_EINVAL, "degenerate Mesher zone limits");
   return 0;
}

struct MshAxisCell {   // one interval of the full axis, cut into level-0 cells
   double a, b;
   int pol, ncell, sub;  // sub = BRM cells per level-0 cell
   double w;
};

int msh_full_axis(const bsahip_mesher_cfg *c, int nm, std::vector<MshAxisCell> &ax, double &dfref)
{
   std::vector<MshInterval> pos;
   double b, fmax;
   CHK(msh_axis(c, nm, pos, b, fmax));
   if (c->bkg_base_rfmnt < 2) return fail(BSAHIP_EINVAL, "bkg_base_rfmnt must be >= 2");
   dfref = 2.0 * b / (c->bkg_base_rfmnt - 1);
   auto cells = [&](double a, double bb, int pol, double w) {
      const double h = dfref * msh_delta_fct(pol);
      const int n = std::max(1, (int)std::ceil((bb - a) / h - 1e-9));
      return MshAxisCell{a, bb, pol, n, 8, w};
   };
   ax.clear();
   for (int i = (int)pos.size() - 1; i >= 0; --i)
      ax.push_back(cells(-pos[i].b, -pos[i].a, pos[i].pol, pos[i].w));
   ax.push_back(MshAxisCell{-b, b, POL_BKG, c->bkg_base_rfmnt - 1, 8, c->bkg_peakw});
   for (auto &iv : pos) ax.push_back(cells(iv.a, iv.b, iv.pol, iv.w));
   return 0;
}

constexpr int MSH_MAX_SUB = 256;   // BRM cells per level-0 cell, at most

int msh_check_cfg(const bsahip_mesher_cfg *cfg)
{
   if (!(cfg->bkg_peakw > 0.0) || !(cfg->bkg_area_ext > 0.0) || !(cfg->peak_area_ext > 0.0) ||
       !(cfg->max_area_ext > 0.0) || cfg->tile_cells < 1 || cfg->tile_cells > 128)
      return fail(BSAHIP_EINVAL, "Mesher extensions and bkg_peakw must be > 0, tile_cells in 1..128");
   const int c = cfg->brm_cells;
   if (c < 0 || c > MSH_MAX_SUB || (c & (c - 1)) != 0)
      return fail(BSAHIP_EINVAL, "brm_cells must be 0 or a power of two up to 256");
   return 0;
}

struct MshTile {
   double a1, b1, a2, b2;   // f1 in [a1,b1], f2 in [a2,b2]
   int n1, n2;              // level-0 cells
   int pol1, pol2, sub1, sub2;
   int rule;                // rule-based level
   int g1, g2;              // first level-0 node of the tile on each axis in the global grid
};

// rule-based level on one axis: smallest l in {1,2,4,8} with h0/l <= w/fct
int msh_rule_level(double h0, double w, int pol)
{
   const double lim = w / msh_bfm_fct(pol) * (1.0 + 1e-9);   // spacing ties count as met
   int l = 1;
   while (l < 8 && h0 / l > lim) l *= 2;
   return l;
}

// BRM cells per level-0 cell of a tile along each axis: a power of two >= 8
// (every refinement level divides it) such that the BRM spacing is at most
// max(w_n/16, d_n/16) for every mode n, where d_n is the distance from the
// tile's range on that axis to +-f_n (resonance of H_m(f1) or H_n(f2)) and,
// with diag_crest, also from its range of f1+f2 (resonance of H_p(f1+f2),
// the reference's diagonal crest, BSA_PREMESH_TYPE_DIAG_CREST_YES).  w_n is the
// peak width 2 xsi_n f_n peak_area_ext: 16 BRM cells per peak width, fewer as
// the FRF tail flattens with distance.
void msh_brm_sub(MshTile &t, const bsahip_mesher_cfg *c, int nm, double dfref)
{
   auto dist = [](double a, double b, double f) { return f < a ? a - f : (f > b ? f - b : 0.0); };
   double r1 = 1e300, r2 = 1e300;
   for (int i = 0; i < nm; ++i) {
      const double w = 2.0 * c->xsi[i] * c->natf[i] * c->peak_area_ext;
      for (int sg = -1; sg <= 1; sg += 2) {
         const double f = sg * c->natf[i];
         r1 = std::min(r1, std::max(w, dist(t.a1, t.b1, f)) / 16.0);
         r2 = std::min(r2, std::max(w, dist(t.a2, t.b2, f)) / 16.0);
         if (c->diag_crest) {
            const double rd = std::max(w, dist(t.a1 + t.a2, t.b1 + t.b2, f)) / 16.0;
            r1 = std::min(r1, rd);
            r2 = std::min(r2, rd);
         }
      }
   }
   const double h1 = (t.b1 - t.a1) / t.n1, h2 = (t.b2 - t.a2) / t.n2;
   if (c->brm_cells > 0) {
      // uniform BRM lattice: spacing df_ref / brm_cells on every tile
      const double hb = dfref / c->brm_cells * (1.0 + 1e-9);
      t.sub1 = 1;
      while (h1 / t.sub1 > hb && t.sub1 < MSH_MAX_SUB) t.sub1 *= 2;
      t.sub2 = 1;
      while (h2 / t.sub2 > hb && t.sub2 < MSH_MAX_SUB) t.sub2 *= 2;
      return;
   }
   t.sub1 = 8;
   while (h1 / t.sub1 > r1 * (1.0 + 1e-9) && t.sub1 < MSH_MAX_SUB) t.sub1 *= 2;
   t.sub2 = 8;
   while (h2 / t.sub2 > r2 * (1.0 + 1e-9) && t.sub2 < MSH_MAX_SUB) t.sub2 *= 2;
}

void msh_tiles(const std::vector<MshAxisCell> &ax, const bsahip_mesher_cfg *cfg, int nm,
               double dfref, std::vector<MshTile> &tiles)
{
   const int tile_cells = cfg->tile_cells;
   struct Seg { double a, b; int n, pol, sub; double w; int g; };
   std::vector<Seg> segs;
   int g = 0;
   for (const auto &c : ax) {
      const int nt = (c.ncell + tile_cells - 1) / tile_cells;
      const double h = (c.b - c.a) / c.ncell;
      int done = 0;
      for (int t = 0; t < nt; ++t) {
         const int n = (c.ncell - done + (nt - t) - 1) / (nt - t);   // balanced split
         const double a = c.a + done * h, bb = (done + n == c.ncell) ? c.b : c.a + (done + n) * h;
         segs.push_back({a, bb, n, c.pol, c.sub, c.w, g});
         done += n;
         g += n;
      }
   }
   tiles.clear();
   for (const auto &s2 : segs)
      for (const auto &s1 : segs) {
         MshTile t;
         t.a1 = s1.a; t.b1 = s1.b; t.a2 = s2.a; t.b2 = s2.b;
         t.n1 = s1.n; t.n2 = s2.n;
         t.g1 = s1.g; t.g2 = s2.g;
         t.pol1 = s1.pol; t.pol2 = s2.pol; t.sub1 = s1.sub; t.sub2 = s2.sub;
         t.rule = std::max(msh_rule_level((s1.b - s1.a) / s1.n, s1.w, s1.pol),
                           msh_rule_level((s2.b - s2.a) / s2.n, s2.w, s2.pol));
         msh_brm_sub(t, cfg, nm, dfref);
         tiles.push_back(t);
      }
}

// Level-0 nodes of the whole axis (the tiles' level-0 grids are sub-grids of
// the tensor product of these nodes).
std::vector<double> msh_axis_nodes(const std::vector<MshTile> &tiles)
{
   std::vector<double> x;
   for (const auto &t : tiles) {
      if (t.a2 != tiles[0].a2) break;               // the first row of tiles spans axis 1
      const double h = (t.b1 - t.a1) / t.n1;
      if (x.empty()) x.push_back(t.a1);
      for (int i = 1; i <= t.n1; ++i) x.push_back(i == t.n1 ? t.b1 : t.a1 + i * h);
   }
   return x;
}

}  // namespace

struct bsahip_mesher {
   Bufs B;
   Prob D;
   int nmp = 0;
   std::vector<MshTile> tiles;
   double dfref = 0.0;
   int nc = 0;                // bispectrum entries per point: NM^3, or NM (ONLY_DIAG)
   bool diag = false;         // the ONLY_DIAG model
   double *lv0 = nullptr;     // resident level-0 BFM on the global grid: [i + j*nx][nc]
   long nx = 0;               // level-0 nodes per axis
   double *PB = nullptr;      // [K][NMP]
   long evals_lv0 = 0;
};

namespace {

// first point of a tile's level-0 grid in the global grid
size_t msh_lv0_at(const bsahip_mesher *M, const MshTile &t)
{
   return (size_t)t.g1 + (size_t)t.g2 * (size_t)M->nx;
}

// BFM at the points (f1[q], f2[q]) into out + slot[q]*nm3, in chunks of at
// most MSH_CHUNK points so the A tables stay small (3 distinct |f| per point
// at most).
constexpr int MSH_CHUNK = 8192;

int msh_eval_chunk(bsahip_mesher *M, int npts, const double *f1, const double *f2,
                   const long *slot_p, double *out);

int msh_eval(bsahip_mesher *M, const std::vector<double> &f1, const std::vector<double> &f2,
             const std::vector<long> &slot, double *out)
{
   for (size_t q0 = 0; q0 < f1.size(); q0 += MSH_CHUNK) {
      const int n = (int)std::min<size_t>(MSH_CHUNK, f1.size() - q0);
      CHK(msh_eval_chunk(M, n, f1.data() + q0, f2.data() + q0, slot.data() + q0, out));
   }
   return 0;
}

int msh_eval_chunk(bsahip_mesher *M, int npts, const double *f1p, const double *f2p,
                   const long *slot_p, double *out)
{
   if (npts == 0) return 0;
   Prob &D = M->D;
   std::vector<double> uq;
   std::vector<int> i1, i2, i3;
   const std::vector<long> slot(slot_p, slot_p + npts);
   point_tables(npts, f1p, f2p, uq, i1, i2, i3);
   const int nu = (int)uq.size();
   if (M->diag) {
      double *fl, *sq;
      int *d1, *d2, *d3;
      long *ds;
      CHK(upload(M->B, "m_fl", uq, &fl));
      CHK(M->B.get("m_sq", (size_t)D.ntc * nu * D.nnl, &sq));
      dim3 gp((D.nnl + 63) / 64, nu, D.ntc);
      k_psd_sqrt<<<gp, 64, 0, g.s>>>(fl, nu, D.prm, D.formula, D.nnl, D.ntc, D.pulsation, sq);
      HIPC(hipGetLastError());
      CHK(upload(M->B, "m_i1", i1, &d1));
      CHK(upload(M->B, "m_i2", i2, &d2));
      CHK(upload(M->B, "m_i3", i3, &d3));
      CHK(upload(M->B, "m_slot", slot, &ds));
      const long tot = (long)npts * D.nm;
      k_diag_points<<<(unsigned)((tot + 255) / 256), 256, 0, g.s>>>(
         sq, nu, D.nnl, D.kv, D.pa_k, D.pb_k, D.K, D.nm, npts, d1, d2, d3, ds, out);
      HIPC(hipGetLastError());
      return 0;
   }
   const long ldU = (long)D.nm * nu;
   double *fl, *U, *V;
   CHK(upload(M->B, "m_fl", uq, &fl));
   CHK(M->B.get("m_U", (size_t)D.K * ldU, &U));
   CHK(build_table(M->B, D, fl, nu, nu, U, ldU));
   CHK(M->B.get("m_V", (size_t)nu * D.K * M->nmp, &V));
   dim3 gv((nu + 63) / 64, D.K);
   k_vtable<<<gv, 256, 0, g.s>>>(U, ldU, nu, D.nm, D.K, M->nmp, V);
   HIPC(hipGetLastError());
   int *d1, *d2, *d3;
   long *ds;
   CHK(upload(M->B, "m_i1", i1, &d1));
   CHK(upload(M->B, "m_i2", i2, &d2));
   CHK(upload(M->B, "m_i3", i3, &d3));
   CHK(upload(M->B, "m_slot", slot, &ds));
   const int ncol = D.nm * D.nm, ct = (ncol + 15) / 16, tpw = (ct + 7) / 8;
#define LAUNCH_BATCH(MT, TPW)                                                                 \
   k_bfm_batch<MT, TPW><<<npts, 512, 0, g.s>>>(V, M->PB, D.K, D.nm, npts, d1, d2, d3, ds, out)
   if (M->nmp == 16) {
      if (tpw <= 1) LAUNCH_BATCH(1, 1);
      else LAUNCH_BATCH(1, 2);
   } else if (M->nmp == 32) {
      if (tpw <= 2) LAUNCH_BATCH(2, 2);
      else if (tpw <= 4) LAUNCH_BATCH(2, 4);
      else LAUNCH_BATCH(2, 8);
   } else {
      return fail(BSAHIP_EUNSUPPORTED, "point evaluation supports at most 32 modes");
   }
#undef LAUNCH_BATCH
   HIPC(hipGetLastError());
   return 0;
}

// Points of a tile's BFM grid at level l (q = i + j*nb1).
void msh_grid(const MshTile &t, int l, std::vector<double> &f1, std::vector<double> &f2,
              int &nb1, int &nb2)
{
   nb1 = t.n1 * l + 1;
   nb2 = t.n2 * l + 1;
   const double h1 = (t.b1 - t.a1) / (nb1 - 1), h2 = (t.b2 - t.a2) / (nb2 - 1);
   f1.resize((size_t)nb1 * nb2);
   f2.resize((size_t)nb1 * nb2);
   for (int j = 0; j < nb2; ++j)
      for (int i = 0; i < nb1; ++i) {
         f1[i + (size_t)j * nb1] = (i == nb1 - 1) ? t.b1 : t.a1 + i * h1;
         f2[i + (size_t)j * nb1] = (j == nb2 - 1) ? t.b2 : t.a2 + j * h2;
      }
}

// BRM grid of a tile at BFM level l: s1 x s2 BRM cells per BFM cell, NA x NB
// cells, spacings h1, h2 (independent of l up to the cap s >= 1).
struct MshBrmGrid {
   int s1, s2, NA, NB;
   double h1, h2;
   double2 *H1 = nullptr, *H2 = nullptr, *H3 = nullptr;
};

MshBrmGrid msh_brm_grid(const MshTile &t, int l)
{
   MshBrmGrid G;
   G.s1 = std::max(1, t.sub1 / l);
   G.s2 = std::max(1, t.sub2 / l);
   G.NA = t.n1 * l * G.s1;
   G.NB = t.n2 * l * G.s2;
   G.h1 = (t.b1 - t.a1) / G.NA;
   G.h2 = (t.b2 - t.a2) / G.NB;
   return G;
}

// FRF tables along the two axes and, with equal spacings, along f1 + f2.
int msh_frf_tables(bsahip_mesher *M, const MshTile &t, MshBrmGrid &G)
{
   Prob &D = M->D;
   const int NA = G.NA, NB = G.NB;
   const bool diag = std::fabs(G.h1 - G.h2) <= 1e-13 * std::max(std::fabs(G.h1), std::fabs(G.h2));
   const size_t n3 = diag ? (size_t)(NA + NB + 1) : 0;
   double2 *H;
   CHK(M->B.get("m_H", ((size_t)(NA + 1) + (NB + 1) + n3) * D.nm, &H));
   G.H1 = H;
   G.H2 = H + (size_t)(NA + 1) * D.nm;
   G.H3 = diag ? G.H2 + (size_t)(NB + 1) * D.nm : nullptr;
   auto table = [&](double f0, double h, int n, double2 *out) -> int {
      const int tot = n * D.nm;
      k_frf_axis<<<(tot + 255) / 256, 256, 0, g.s>>>(f0, h, n, D.Mm, D.Km, D.Cm, D.nm, out);
      HIPC(hipGetLastError());
      return 0;
   };
   CHK(table(t.a1, G.h1, NA + 1, G.H1));
   CHK(table(t.a2, G.h2, NB + 1, G.H2));
   if (diag) CHK(table(t.a1 + t.a2, G.h1, NA + NB + 1, G.H3));
   return 0;
}

// Per-tile visitor of the post-mesh: the tile's index, its level and its BFM
// grid (nb1 x nb2 points, bf[(i + j*ldq)*nm3 + mnp], on the device).
using MshVisit = std::function<int(size_t, const MshTile &, int, const double *, long, int, int)>;

// One visited tile whose quadrature is still to be launched.
struct MshRef {
   size_t k;
   int l, nb1, nb2;
   const double *bf;
   long ldq;
};

// The visited tiles of one structure awaiting their quadrature: its m3f
// accumulator and each tile's m3r accumulator (m3r_of(tile index)).
struct MshJob {
   bsahip_mesher *M = nullptr;
   std::vector<MshRef> refs;
   double *m3f = nullptr;
   std::function<double *(size_t)> m3r_of;
   long brm = 0;                  // BRM points integrated
};

// The quadrature of the visited tiles of one or more structures (all of the
// full model, or all ONLY_DIAG) in two launches: FRF tables of every tile
// (k_frf_segs), then k_msh_integrate_full / k_msh_integrate over (tile, cell
// range) items.  Scratch buffers come from the first structure's buffers.
int msh_integrate_jobs(const std::vector<MshJob *> &jobs)
{
   size_t ntr = 0;
   for (MshJob *J : jobs) ntr += J->refs.size();
   if (ntr == 0) return 0;
   const bool diag = jobs[0]->M->diag;
   int bx = 1;
   std::vector<MshSeg> seg;
   std::vector<long> eoff(1, 0);
   std::vector<MshTileRun> tr;
   std::vector<long> ncell_of;
   tr.reserve(ntr);
   long ncells = 0;
   for (MshJob *J : jobs) {
      bsahip_mesher *M = J->M;
      Prob &D = M->D;
      if (M->diag != diag) return fail(BSAHIP_EINVAL, "a Mesher batch mixes the ONLY_DIAG and full models");
      // threads per item: one per entry (ONLY_DIAG), or one per (m, p) and
      // MSH_NQ n (full model)
      const int nthr = diag ? M->nc : D.nm * D.nm * ((D.nm + MSH_NQ - 1) / MSH_NQ);
      bx = std::max(bx, (nthr + 255) / 256);
      auto add_seg = [&](double f0, double h, int n) -> long {
         const long off = eoff.back();
         seg.push_back(MshSeg{f0, h, off, D.Mm, D.Km, D.Cm, D.nm});
         eoff.push_back(off + (long)n * D.nm);
         return off;
      };
      for (const MshRef &R : J->refs) {
         const MshTile &t = M->tiles[R.k];
         const MshBrmGrid G = msh_brm_grid(t, R.l);
         const bool eq = std::fabs(G.h1 - G.h2) <= 1e-13 * std::max(std::fabs(G.h1), std::fabs(G.h2));
         MshTileRun T;
         T.bf = R.bf;
         T.m3f = J->m3f;
         T.m3r = J->m3r_of(R.k);
         T.Mm = D.Mm;
         T.Km = D.Km;
         T.Cm = D.Cm;
         T.ldq = R.ldq;
         T.h1 = add_seg(t.a1, G.h1, G.NA + 1);
         T.h2 = add_seg(t.a2, G.h2, G.NB + 1);
         T.h3 = eq ? add_seg(t.a1 + t.a2, G.h1, G.NA + G.NB + 1) : -1;
         T.nm = D.nm;
         T.nb1 = R.nb1;
         T.nb2 = R.nb2;
         T.s1 = G.s1;
         T.s2 = G.s2;
         T.f3a = t.a1 + t.a2;
         T.dh1 = G.h1;
         T.dh2 = G.h2;
         T.ws = (2.0 * M_PI * G.h1) * (2.0 * M_PI * G.h2);
         tr.push_back(T);
         ncell_of.push_back((long)(R.nb1 - 1) * (R.nb2 - 1));
         ncells += ncell_of.back();
         J->brm += (long)(G.NA + 1) * (G.NB + 1);
      }
   }
   // items: each tile's cells cut in ranges of at most `per` cells, so that
   // the launch holds some 8192 (item, thread block) workgroups or more
   const long per = std::max<long>(1, ncells * bx / 8192);
   std::vector<int> it;
   std::vector<int2> cr;
   for (size_t r = 0; r < tr.size(); ++r)
      for (long c0 = 0; c0 < ncell_of[r]; c0 += per) {
         it.push_back((int)r);
         cr.push_back(make_int2((int)c0, (int)std::min<long>(ncell_of[r], c0 + per)));
      }
   if ((long)it.size() * bx > 0x7fffffffL)
      return fail(BSAHIP_EUNSUPPORTED, "Mesher post-mesh launch above 2^31 workgroups");
   Bufs &B = jobs[0]->M->B;
   MshSeg *dseg;
   long *doff;
   double2 *H;
   MshTileRun *dtr;
   int *dit;
   int2 *dcr;
   CHK(upload(B, "m_seg", seg, &dseg));
   CHK(upload(B, "m_segoff", eoff, &doff));
   CHK(B.get("m_H", (size_t)eoff.back(), &H));
   CHK(upload(B, "m_tiles_run", tr, &dtr));
   CHK(upload(B, "m_items", it, &dit));
   CHK(upload(B, "m_ranges", cr, &dcr));
   k_frf_segs<<<(unsigned)((eoff.back() + 255) / 256), 256, 0, g.s>>>(dseg, doff, (int)seg.size(), H);
   HIPC(hipGetLastError());
   if (diag)
      k_msh_integrate<<<(unsigned)(it.size() * bx), 256, 0, g.s>>>(dtr, dit, dcr, bx, 1, H);
   else
      k_msh_integrate_full<<<(unsigned)(it.size() * bx), 256, 0, g.s>>>(dtr, dit, dcr, bx, H);
   HIPC(hipGetLastError());
   for (MshJob *J : jobs) J->refs.clear();
   return 0;
}

// Post-mesh traversal shared by the run and the export: level-1 tiles read the
// resident level-0 BFM in place; refined tiles are evaluated in groups whose
// fine grids fit the staging buffer.  `done(last)`, when given, is called once
// the tiles visited so far may be consumed: after each group, before its
// staging buffer is reused (last = false), and at the end (last = true; those
// tiles stay valid until this Mesher's next post-mesh).  Returns the
// refinement evaluations.
int msh_post(bsahip_mesher *m, const int32_t *levels, const MshVisit &visit, long &evals_post,
             const std::function<int(bool)> &done = nullptr)
{
   const size_t nm3 = (size_t)m->nc, nt = m->tiles.size();
   const size_t cap = std::max<size_t>(66049, ((size_t)1 << 29) / nm3);   // points (4 GiB)
   std::vector<double> f1, f2, e1, e2;
   std::vector<long> slot;
   std::vector<size_t> grp, goff;
   size_t used = 0;
   evals_post = 0;
   std::vector<long> csrc, cdst;
   auto flush = [&](bool last) -> int {
      if (grp.empty()) return 0;
      double *fine;
      CHK(m->B.get("m_fine", used * nm3, &fine));
      CHK(msh_eval(m, e1, e2, slot, fine));
      evals_post += (long)e1.size();
      // the level-0 points of every refined tile into its fine grid, one launch
      csrc.clear();
      cdst.clear();
      for (size_t k = 0; k < grp.size(); ++k) {
         const MshTile &t = m->tiles[grp[k]];
         const int l = levels ? levels[grp[k]] : t.rule;
         const int nb1 = t.n1 * l + 1;
         const long base = msh_lv0_at(m, t);
         for (int j = 0; j <= t.n2; ++j)
            for (int i = 0; i <= t.n1; ++i) {
               csrc.push_back(base + i + (long)j * m->nx);
               cdst.push_back((long)goff[k] + (long)i * l + (long)j * l * nb1);
            }
      }
      if (csrc.size() > 65535u * 65535u) return fail(BSAHIP_EUNSUPPORTED, "too many level-0 points");
      long *dsrc, *ddst;
      CHK(upload(m->B, "m_csrc", csrc, &dsrc));
      CHK(upload(m->B, "m_cdst", cdst, &ddst));
      for (size_t q0 = 0; q0 < csrc.size(); q0 += 65535) {
         const unsigned nq = (unsigned)std::min<size_t>(65535, csrc.size() - q0);
         dim3 gc((unsigned)((nm3 + 255) / 256), nq);
         k_msh_copy_pts<<<gc, 256, 0, g.s>>>(m->lv0, dsrc + q0, ddst + q0, (int)nm3, fine);
         HIPC(hipGetLastError());
      }
      for (size_t k = 0; k < grp.size(); ++k) {
         const MshTile &t = m->tiles[grp[k]];
         const int l = levels ? levels[grp[k]] : t.rule;
         const int nb1 = t.n1 * l + 1, nb2 = t.n2 * l + 1;
         CHK(visit(grp[k], t, l, fine + goff[k] * nm3, nb1, nb1, nb2));
      }
      if (done && !last) CHK(done(false));
      HIPC(hipStreamSynchronize(g.s));
      grp.clear();
      goff.clear();
      e1.clear();
      e2.clear();
      slot.clear();
      used = 0;
      return 0;
   };
   for (size_t i = 0; i < nt; ++i) {
      const MshTile &t = m->tiles[i];
      const int l = levels ? levels[i] : t.rule;
      if (l == 1) {
         CHK(visit(i, t, 1, m->lv0 + msh_lv0_at(m, t) * nm3, m->nx, t.n1 + 1, t.n2 + 1));
         continue;
      }
      const size_t np = (size_t)(t.n1 * l + 1) * (t.n2 * l + 1);
      if (!grp.empty() && used + np > cap) CHK(flush(false));
      int nb1, nb2;
      msh_grid(t, l, f1, f2, nb1, nb2);
      for (int j = 0; j < nb2; ++j)
         for (int ii = 0; ii < nb1; ++ii)
            if ((ii % l) || (j % l)) {
               const size_t q = ii + (size_t)j * nb1;
               e1.push_back(f1[q]);
               e2.push_back(f2[q]);
               slot.push_back((long)(used + q));
            }
      grp.push_back(i);
      goff.push_back(used);
      used += np;
   }
   CHK(flush(true));   // the last group's tiles are consumed with the others by done(true)
   if (done) CHK(done(true));
   return 0;
}

int msh_check_levels(const bsahip_mesher *m, const int32_t *levels)
{
   for (size_t i = 0; i < m->tiles.size(); ++i) {
      const int l = levels ? levels[i] : m->tiles[i].rule;
      if (!(l == 1 || l == 2 || l == 4 || l == 8))
         return fail(BSAHIP_EINVAL, "tile levels must be 1, 2, 4 or 8");
   }
   return 0;
}

// Everything of bsahip_mesher_create but the pre-mesh evaluation: problem
// tables, tile layout, the level-0 buffer and its axis nodes x.
int msh_setup(const bsahip_problem *p, const bsahip_mesher_cfg *cfg, int32_t flags,
              std::unique_ptr<bsahip_mesher> &M, std::vector<double> &x)
{
   if (!cfg || !cfg->natf || !cfg->xsi) return fail(BSAHIP_EINVAL, "NULL Mesher argument");
   CHK(ensure_init());
   if ((flags & BSAHIP_PREC_MASK) != BSAHIP_PREC_FP64)
      return fail(BSAHIP_EUNSUPPORTED, "the Mesher runs in FP64");
   CHK(msh_check_cfg(cfg));
   M.reset(new bsahip_mesher());
   CHK(load_problem(p, 0, M->B, M->D));
   // the ONLY_DIAG model has no coherence between nodes and no POD
   // (BsaMesherImpl.F90:68: the POD only with i_only_diag_ == 0)
   if (!(flags & BSAHIP_ONLY_DIAG)) CHK(set_pod(p, M->D));
   Prob &D = M->D;
   for (int i = 0; i < D.nm; ++i)
      if (!(cfg->natf[i] > 0.0) || !(cfg->xsi[i] > 0.0))
         return fail(BSAHIP_EINVAL, "natural frequencies and damping ratios must be > 0");
   if (D.nm > 32) return fail(BSAHIP_EUNSUPPORTED, "the Mesher supports at most 32 modes");
   M->nmp = D.nm <= 16 ? 16 : 32;
   std::vector<MshAxisCell> ax;
   CHK(msh_full_axis(cfg, D.nm, ax, M->dfref));
   msh_tiles(ax, cfg, D.nm, M->dfref, M->tiles);
   CHK(M->B.get("m_PB", (size_t)D.K * M->nmp, &M->PB));
   k_pbt<<<(unsigned)((D.K * M->nmp + 255) / 256), 256, 0, g.s>>>(D.pb_k, D.nm, D.K, M->nmp, M->PB);
   HIPC(hipGetLastError());
   M->diag = (flags & BSAHIP_ONLY_DIAG) != 0;
   M->nc = M->diag ? D.nm : D.nm * D.nm * D.nm;
   const size_t nm3 = (size_t)M->nc;
   x = msh_axis_nodes(M->tiles);
   M->nx = (long)x.size();
   CHK(M->B.get("m_lv0", x.size() * x.size() * nm3, &M->lv0));
   return 0;
}

void msh_geom(const bsahip_mesher *m, double *geom, int32_t *rule_level)
{
   for (size_t i = 0; i < m->tiles.size(); ++i) {
      const MshTile &t = m->tiles[i];
      if (geom) {
         double *gq = geom + 10 * i;
         gq[0] = t.a1; gq[1] = t.b1; gq[2] = t.a2; gq[3] = t.b2;
         gq[4] = t.n1; gq[5] = t.n2; gq[6] = t.pol1; gq[7] = t.pol2;
         gq[8] = t.sub1; gq[9] = t.sub2;
      }
      if (rule_level) rule_level[i] = t.rule;
   }
}

// checkpoint file of a Mesher's resident zones
const char MSH_MAGIC[16] = "BSAHIP-MSH-DUMP";

}  // namespace

extern "C" {

int bsahip_mesher_create(const bsahip_problem *p, const bsahip_mesher_cfg *cfg, int32_t flags,
                         bsahip_mesher **out)
{
   std::lock_guard<std::recursive_mutex> lk(g.mu);
   if (!out) return fail(BSAHIP_EINVAL, "NULL Mesher argument");
   *out = nullptr;
   std::unique_ptr<bsahip_mesher> M;
   std::vector<double> x;
   CHK(msh_setup(p, cfg, flags, M, x));
   // pre-mesh: the level-0 BFM on the global grid (every tile's level-0 points
   // once, shared edges and corners included once), kept resident
   const size_t npts = x.size() * x.size();
   std::vector<double> f1, f2;
   std::vector<long> slot;
   for (size_t j = 0; j < x.size(); ++j)
      for (size_t i = 0; i < x.size(); ++i) {
         f1.push_back(x[i]);
         f2.push_back(x[j]);
         slot.push_back((long)(i + j * x.size()));
      }
   CHK(msh_eval(M.get(), f1, f2, slot, M->lv0));
   HIPC(hipStreamSynchronize(g.s));
   M->evals_lv0 = (long)npts;
   *out = M.release();
   return 0;
}

int bsahip_mesher_layout(const bsahip_mesher_cfg *cfg, int32_t nm, int32_t max_tiles,
                         double *geom, int32_t *rule_level, int32_t *ntiles)
{
   if (!cfg || !cfg->natf || !cfg->xsi || !ntiles || nm < 1)
      return fail(BSAHIP_EINVAL, "NULL Mesher argument or nm < 1");
   CHK(msh_check_cfg(cfg));
   std::vector<MshAxisCell> ax;
   double dfref;
   CHK(msh_full_axis(cfg, nm, ax, dfref));
   std::vector<MshTile> tiles;
   msh_tiles(ax, cfg, nm, dfref, tiles);
   *ntiles = (int32_t)tiles.size();
   if (!geom && !rule_level) return 0;
   if ((int32_t)tiles.size() > max_tiles) return fail(BSAHIP_EINVAL, "max_tiles too small");
   for (size_t i = 0; i < tiles.size(); ++i) {
      const MshTile &t = tiles[i];
      if (geom) {
         double *gq = geom + 10 * i;
         gq[0] = t.a1; gq[1] = t.b1; gq[2] = t.a2; gq[3] = t.b2;
         gq[4] = t.n1; gq[5] = t.n2; gq[6] = t.pol1; gq[7] = t.pol2;
         gq[8] = t.sub1; gq[9] = t.sub2;
      }
      if (rule_level) rule_level[i] = t.rule;
   }
   return 0;
}

int bsahip_mesher_ntiles(const bsahip_mesher *m, int32_t *ntiles)
{
   if (!m || !ntiles) return fail(BSAHIP_EINVAL, "NULL argument");
   *ntiles = (int32_t)m->tiles.size();
   return 0;
}

int bsahip_mesher_tiles(const bsahip_mesher *m, double *geom, int32_t *rule_level)
{
   if (!m) return fail(BSAHIP_EINVAL, "NULL Mesher");
   msh_geom(m, geom, rule_level);
   return 0;
}

int bsahip_mesher_features(bsahip_mesher *m, int32_t ns, double *stencil)
{
   std::lock_guard<std::recursive_mutex> lk(g.mu);
   if (!m || !stencil || ns < 2) return fail(BSAHIP_EINVAL, "need a Mesher, ns >= 2 and an output");
   const int nm3 = m->nc;
   const size_t nt = m->tiles.size(), per = (size_t)ns * ns;
   double *d;
   CHK(m->B.get("m_stencil", nt * per, &d));
   std::vector<long> t0(nt);
   std::vector<int2> nb(nt);
   for (size_t i = 0; i < nt; ++i) {
      const MshTile &t = m->tiles[i];
      t0[i] = msh_lv0_at(m, t);
      nb[i] = make_int2(t.n1 + 1, t.n2 + 1);
   }
   long *dt0;
   int2 *dnb;
   CHK(upload(m->B, "m_st_t0", t0, &dt0));
   CHK(upload(m->B, "m_st_nb", nb, &dnb));
   for (size_t k0 = 0; k0 < nt; k0 += 65535) {
      const unsigned nk = (unsigned)std::min<size_t>(65535, nt - k0);
      k_bfm_stencil<<<dim3((unsigned)per, nk), 64, 0, g.s>>>(m->lv0, dt0 + k0, dnb + k0, m->nx, nm3,
                                                            ns, d + k0 * per);
      HIPC(hipGetLastError());
   }
   HIPC(hipMemcpyAsync(stencil, d, nt * per * sizeof(double), hipMemcpyDeviceToHost, g.s));
   HIPC(hipStreamSynchronize(g.s));
   return 0;
}

int bsahip_mesher_run_tiles(bsahip_mesher *m, const int32_t *levels, double *tile_m3mr,
                            double *m3mf, double *m3mr, int64_t *counts)
{
   std::lock_guard<std::recursive_mutex> lk(g.mu);
   if (!m) return fail(BSAHIP_EINVAL, "NULL Mesher");
   const size_t nm3 = (size_t)m->nc, nt = m->tiles.size();
   CHK(msh_check_levels(m, levels));
   // accumulators: m3f, m3r, then one m3r slot per tile when asked for
   double *acc;
   const size_t nacc = 2 * nm3 + (tile_m3mr ? nt * nm3 : 0);
   CHK(m->B.get("m_acc", nacc, &acc));
   CHK(fill(acc, nacc, 0.0));
   long evals_post = 0;
   // tiles are collected as visited and integrated together (msh_integrate_jobs)
   MshJob job;
   job.M = m;
   job.m3f = acc;
   job.m3r_of = [&](size_t k) { return tile_m3mr ? acc + 2 * nm3 + k * nm3 : acc + nm3; };
   CHK(msh_post(
      m, levels,
      [&](size_t k, const MshTile &, int l, const double *bf, long ldq, int nb1, int nb2) {
         job.refs.push_back(MshRef{k, l, nb1, nb2, bf, ldq});
         return 0;
      },
      evals_post, [&](bool) { return msh_integrate_jobs({&job}); }));
   const long brm = job.brm;
   if (tile_m3mr) {
      HIPC(hipMemcpyAsync(tile_m3mr, acc + 2 * nm3, nt * nm3 * sizeof(double),
                          hipMemcpyDeviceToHost, g.s));
   }
   if (m3mf) HIPC(hipMemcpyAsync(m3mf, acc, nm3 * sizeof(double), hipMemcpyDeviceToHost, g.s));
   HIPC(hipStreamSynchronize(g.s));
   if (m3mr) {
      if (tile_m3mr) {
         // tile by tile, so that each pass reads one contiguous slot
         std::fill(m3mr, m3mr + nm3, 0.0);
         for (size_t k = 0; k < nt; ++k) {
            const double *c = tile_m3mr + k * nm3;
            for (size_t q = 0; q < nm3; ++q) m3mr[q] += c[q];
         }
      } else {
         HIPC(hipMemcpy(m3mr, acc + nm3, nm3 * sizeof(double), hipMemcpyDeviceToHost));
      }
   }
   if (counts) {
      counts[0] = m->evals_lv0;
      counts[1] = evals_post;
      counts[2] = brm;
   }
   return 0;
}

int bsahip_mesher_run(bsahip_mesher *m, const int32_t *levels, double *m3mf, double *m3mr,
                      int64_t *counts)
{
   return bsahip_mesher_run_tiles(m, levels, nullptr, m3mf, m3mr, counts);
}

int bsahip_mesher_run_batch(int32_t n, bsahip_mesher *const *ms, const int32_t *const *levels,
                            double *const *m3mf, double *const *m3mr, int64_t *counts)
{
   std::lock_guard<std::recursive_mutex> lk(g.mu);
   if (n < 0 || (n > 0 && !ms)) return fail(BSAHIP_EINVAL, "need n >= 0 and the Meshers");
   for (int i = 0; i < n; ++i) {
      if (!ms[i]) return fail(BSAHIP_EINVAL, "NULL Mesher in the batch");
      for (int j = 0; j < i; ++j)
         if (ms[j] == ms[i]) return fail(BSAHIP_EINVAL, "a Mesher appears twice in the batch");
      CHK(msh_check_levels(ms[i], levels ? levels[i] : nullptr));
   }
   std::vector<MshJob> jobs(n);
   std::vector<long> evals(n, 0);
   std::vector<double *> acc(n);
   for (int i = 0; i < n; ++i) {
      bsahip_mesher *m = ms[i];
      const size_t nm3 = (size_t)m->nc;
      CHK(m->B.get("m_acc", 2 * nm3, &acc[i]));
      CHK(fill(acc[i], 2 * nm3, 0.0));
      MshJob &J = jobs[i];
      J.M = m;
      J.m3f = acc[i];
      double *r = acc[i] + nm3;
      J.m3r_of = [r](size_t) { return r; };
      // each Mesher's last group of tiles waits for the shared launch below;
      // earlier groups (their staging buffer is reused) are integrated at once
      CHK(msh_post(
         m, levels ? levels[i] : nullptr,
         [&J](size_t k, const MshTile &, int l, const double *bf, long ldq, int nb1, int nb2) {
            J.refs.push_back(MshRef{k, l, nb1, nb2, bf, ldq});
            return 0;
         },
         evals[i], [&J](bool last) { return last ? 0 : msh_integrate_jobs({&J}); }));
   }
   std::vector<MshJob *> full, diag;
   for (MshJob &J : jobs) (J.M->diag ? diag : full).push_back(&J);
   if (!full.empty()) CHK(msh_integrate_jobs(full));
   if (!diag.empty()) CHK(msh_integrate_jobs(diag));
   for (int i = 0; i < n; ++i) {
      const size_t nm3 = (size_t)ms[i]->nc;
      if (m3mf && m3mf[i])
         HIPC(hipMemcpyAsync(m3mf[i], acc[i], nm3 * sizeof(double), hipMemcpyDeviceToHost, g.s));
      if (m3mr && m3mr[i])
         HIPC(hipMemcpyAsync(m3mr[i], acc[i] + nm3, nm3 * sizeof(double), hipMemcpyDeviceToHost, g.s));
   }
   HIPC(hipStreamSynchronize(g.s));
   if (counts)
      for (int i = 0; i < n; ++i) {
         counts[3 * i] = ms[i]->evals_lv0;
         counts[3 * i + 1] = evals[i];
         counts[3 * i + 2] = jobs[i].brm;
      }
   return 0;
}

}  // extern "C"

namespace {

// A-posteriori levels (bsahip_mesher_adapt).  With outputs (m3mf, m3mr or
// counts non-NULL) every refinement round runs the tiles it keeps at their
// current level instead of level 1, so that the last run is the run at the
// final levels and its moments are returned without another run.
int msh_adapt(bsahip_mesher *m, double tol, int32_t *levels, int64_t *evals, double *m3mf_out,
              double *m3mr_out, int64_t *counts_out)
{
   if (!m || !levels || !(tol > 0.0)) return fail(BSAHIP_EINVAL, "need a Mesher, levels and tol > 0");
   const bool want = m3mf_out || m3mr_out || counts_out;
   const size_t nt = m->tiles.size(), nm3 = (size_t)m->nc;
   std::vector<double> prev(nt * nm3), cur(nt * nm3), m3f(nm3), m3r(nm3);
   std::vector<int32_t> lev(nt, 1), lr(nt);
   std::vector<char> up(nt);
   int64_t cnt[3] = {0, 0, 0};
   // level 1 (the resident level-0 samples), then level 2 on every tile
   CHK(bsahip_mesher_run_tiles(m, lev.data(), prev.data(), m3f.data(), m3r.data(), nullptr));
   std::fill(lev.begin(), lev.end(), 2);
   CHK(bsahip_mesher_run_tiles(m, lev.data(), cur.data(), m3f.data(), m3r.data(), cnt));
   double M = 0.0;
   for (size_t q = 0; q < nm3; ++q) M = std::max(M, std::fabs(m3r[q]));
   const double budget = 0.5 * tol * M;
   // error estimate of a tile's contribution at level L from the change
   // D_L = c_L - c_L/2, as the tail of a geometric series of ratio rho:
   // err(L) ~ D_L rho / (1 - rho).  rho starts at 1/4 (bilinear interpolation
   // is second order, giving D_L / 3) and, once a tile has two changes, is the
   // observed ratio max|D_L| / max|D_L/2| clipped to [1/64, 1/2] (the change
   // falls much faster than 1/4 once a tile resolves its peak).  Kept signed
   // per entry: the error of m3mr is their sum over the tiles.
   std::vector<double> delta(nt * nm3), dnorm(nt), dlast(nt, 0.0);
   auto set_delta = [&](size_t t) {
      double dn = 0.0;
      for (size_t q = 0; q < nm3; ++q)
         dn = std::max(dn, std::fabs(cur[t * nm3 + q] - prev[t * nm3 + q]));
      double rho = 0.25;
      if (dlast[t] > 0.0) rho = std::min(0.5, std::max(1.0 / 64.0, dn / dlast[t]));
      const double f = rho / (1.0 - rho);
      for (size_t q = 0; q < nm3; ++q) delta[t * nm3 + q] = f * (cur[t * nm3 + q] - prev[t * nm3 + q]);
      dnorm[t] = f * dn;
      dlast[t] = dn;
   };
   for (size_t t = 0; t < nt; ++t) set_delta(t);
   std::swap(prev, cur);
   std::vector<double> E(nm3);
   auto emax = [&]() {
      double e = 0.0;
      for (size_t q = 0; q < nm3; ++q) e = std::max(e, std::fabs(E[q]));
      return e;
   };
   for (int round = 0; round < 2; ++round) {
      std::fill(E.begin(), E.end(), 0.0);
      for (size_t t = 0; t < nt; ++t)
         for (size_t q = 0; q < nm3; ++q) E[q] += delta[t * nm3 + q];
      if (emax() <= budget) break;
      // refine the tiles of largest estimate, in that order, until the
      // estimated error with theirs cut by one level (by about 1/4) fits
      std::vector<size_t> ord;
      for (size_t t = 0; t < nt; ++t)
         if (lev[t] < 8 && dnorm[t] > 0.0) ord.push_back(t);
      std::sort(ord.begin(), ord.end(), [&](size_t a, size_t b) { return dnorm[a] > dnorm[b]; });
      for (size_t t = 0; t < nt; ++t) lr[t] = want ? lev[t] : 1;
      std::fill(up.begin(), up.end(), 0);
      bool any = false;
      for (size_t t : ord) {
         if (emax() <= budget) break;
         lr[t] = 2 * lev[t];
         up[t] = 1;
         any = true;
         for (size_t q = 0; q < nm3; ++q) E[q] -= 0.75 * delta[t * nm3 + q];
      }
      if (!any) break;
      CHK(bsahip_mesher_run_tiles(m, lr.data(), cur.data(), m3f.data(), m3r.data(), cnt));
      for (size_t t = 0; t < nt; ++t)
         if (up[t]) {
            set_delta(t);
            std::copy(cur.begin() + t * nm3, cur.begin() + (t + 1) * nm3, prev.begin() + t * nm3);
            lev[t] = lr[t];
         }
   }
   int64_t n = m->evals_lv0;
   for (size_t t = 0; t < nt; ++t) {
      const MshTile &T = m->tiles[t];
      levels[t] = lev[t];
      n += (int64_t)(T.n1 * lev[t] + 1) * (T.n2 * lev[t] + 1) - (int64_t)(T.n1 + 1) * (T.n2 + 1);
   }
   if (evals) *evals = n;
   // with outputs, the last run was at the final levels: level 2 everywhere,
   // or a refinement round whose kept tiles ran at their own levels
   if (m3mf_out) std::copy(m3f.begin(), m3f.end(), m3mf_out);
   if (m3mr_out) std::copy(m3r.begin(), m3r.end(), m3mr_out);
   if (counts_out) std::copy(cnt, cnt + 3, counts_out);
   return 0;
}

}  // namespace

extern "C" {

int bsahip_mesher_adapt(bsahip_mesher *m, double tol, int32_t *levels, int64_t *evals)
{
   std::lock_guard<std::recursive_mutex> lk(g.mu);
   return msh_adapt(m, tol, levels, evals, nullptr, nullptr, nullptr);
}

int bsahip_mesher_adapt_run(bsahip_mesher *m, double tol, int32_t *levels, int64_t *evals,
                            double *m3mf, double *m3mr, int64_t *counts)
{
   std::lock_guard<std::recursive_mutex> lk(g.mu);
   return msh_adapt(m, tol, levels, evals, m3mf, m3mr, counts);
}

int bsahip_mesher_export(bsahip_mesher *m, const int32_t *levels, int32_t ncol, const double *coef,
                         int64_t cap, int32_t *dims, double *fi, double *fj, double *val,
                         int64_t *npts)
{
   std::lock_guard<std::recursive_mutex> lk(g.mu);
   if (!m || !npts) return fail(BSAHIP_EINVAL, "NULL Mesher or point count");
   CHK(msh_check_levels(m, levels));
   Prob &D = m->D;
   const int nm3 = m->nc;
   const size_t nt = m->tiles.size();
   // sizes: the BRM grid of every tile at its level
   std::vector<int64_t> off(nt + 1, 0);
   for (size_t i = 0; i < nt; ++i) {
      const MshTile &t = m->tiles[i];
      const MshBrmGrid G = msh_brm_grid(t, levels ? levels[i] : t.rule);
      if (dims) {
         dims[2 * i] = G.NA + 1;
         dims[2 * i + 1] = G.NB + 1;
      }
      off[i + 1] = off[i] + (int64_t)(G.NA + 1) * (G.NB + 1);
   }
   *npts = off[nt];
   if (!fi && !fj && !val) return 0;
   if (ncol < 0 || (ncol > 0 && !val) || (!coef && ncol != nm3))
      return fail(BSAHIP_EINVAL, "ncol < 0, NULL values, or NULL coef with ncol != entries");
   if (cap < off[nt]) return fail(BSAHIP_EINVAL, "cap below the number of export points");
   // unit columns take the single-entry path
   std::vector<int> sel(std::max(ncol, 1), -1);
   for (int c = 0; c < ncol; ++c) {
      if (!coef) {
         sel[c] = c;
         continue;
      }
      int one = -1, nz = 0;
      for (int q = 0; q < nm3; ++q)
         if (coef[(size_t)c * nm3 + q] != 0.0) {
            ++nz;
            if (coef[(size_t)c * nm3 + q] == 1.0) one = q;
         }
      if (nz == 1 && one >= 0) sel[c] = one;
   }
   int *dsel;
   double *dcoef = nullptr;
   CHK(upload(m->B, "m_xsel", sel, &dsel));
   if (coef && ncol > 0) {
      std::vector<double> cc(coef, coef + (size_t)ncol * nm3);
      CHK(upload(m->B, "m_xcoef", cc, &dcoef));
   }
   long evals = 0;
   CHK(msh_post(m, levels,
                [&](size_t k, const MshTile &t, int l, const double *bf, long ldq, int nb1, int nb2) {
                   MshBrmGrid G = msh_brm_grid(t, l);
                   CHK(msh_frf_tables(m, t, G));
                   const int64_t np = (int64_t)(G.NA + 1) * (G.NB + 1);
                   for (int b = 0; b <= G.NB; ++b)
                      for (int a = 0; a <= G.NA; ++a) {
                         const int64_t q = off[k] + a + (int64_t)b * (G.NA + 1);
                         if (fi) fi[q] = t.a1 + a * G.h1;
                         if (fj) fj[q] = t.a2 + b * G.h2;
                      }
                   if (ncol == 0) return 0;
                   double *dv;
                   CHK(m->B.get("m_xval", (size_t)np * ncol, &dv));
                   k_msh_export<<<(unsigned)((np + 255) / 256), 256, 0, g.s>>>(
                      bf, ldq, D.nm, m->nc, m->diag ? 1 : 0, nb1, nb2, G.s1, G.s2, G.H1, G.H2, G.H3, t.a1 + t.a2, G.h1,
                      G.h2, D.Mm, D.Km, D.Cm, ncol, dsel, dcoef, dv);
                   HIPC(hipGetLastError());
                   HIPC(hipMemcpyAsync(val + (size_t)off[k] * ncol, dv,
                                       (size_t)np * ncol * sizeof(double), hipMemcpyDeviceToHost,
                                       g.s));
                   HIPC(hipStreamSynchronize(g.s));
                   return 0;
                },
                evals));
   return 0;
}

int bsahip_mesher_save(const bsahip_mesher *m, const char *path)
{
   std::lock_guard<std::recursive_mutex> lk(g.mu);
   if (!m || !path) return fail(BSAHIP_EINVAL, "NULL Mesher or path");
   const size_t nt = m->tiles.size(), nm3 = (size_t)m->nc;
   std::unique_ptr<FILE, int (*)(FILE *)> f(std::fopen(path, "wb"), std::fclose);
   if (!f) return fail(BSAHIP_EINVAL, std::string("cannot open ") + path + " for writing");
   const int32_t head[4] = {1, m->D.nm, (int32_t)nt, m->diag ? 1 : 0};
   const int64_t sizes[2] = {m->nx, m->evals_lv0};
   std::vector<double> geom(10 * nt);
   std::vector<int32_t> rule(nt);
   msh_geom(m, geom.data(), rule.data());
   bool ok = std::fwrite(MSH_MAGIC, 1, 16, f.get()) == 16 &&
             std::fwrite(head, sizeof(head), 1, f.get()) == 1 &&
             std::fwrite(sizes, sizeof(sizes), 1, f.get()) == 1 &&
             std::fwrite(geom.data(), sizeof(double), geom.size(), f.get()) == geom.size() &&
             std::fwrite(rule.data(), sizeof(int32_t), nt, f.get()) == nt;
   const size_t tot = (size_t)m->nx * m->nx * nm3, chunk = (size_t)1 << 23;
   std::vector<double> h(std::min(tot, chunk));
   for (size_t o = 0; ok && o < tot; o += chunk) {
      const size_t n = std::min(chunk, tot - o);
      HIPC(hipMemcpy(h.data(), m->lv0 + o, n * sizeof(double), hipMemcpyDeviceToHost));
      ok = std::fwrite(h.data(), sizeof(double), n, f.get()) == n;
   }
   if (!ok || std::fflush(f.get()) != 0) return fail(BSAHIP_EINVAL, std::string("write error on ") + path);
   return 0;
}

int bsahip_mesher_load(const bsahip_problem *p, const bsahip_mesher_cfg *cfg, int32_t flags,
                       const char *path, bsahip_mesher **out)
{
   std::lock_guard<std::recursive_mutex> lk(g.mu);
   if (!out || !path) return fail(BSAHIP_EINVAL, "NULL Mesher argument or path");
   *out = nullptr;
   std::unique_ptr<bsahip_mesher> M;
   std::vector<double> x;
   CHK(msh_setup(p, cfg, flags, M, x));
   std::unique_ptr<FILE, int (*)(FILE *)> f(std::fopen(path, "rb"), std::fclose);
   if (!f) return fail(BSAHIP_EINVAL, std::string("cannot open ") + path);
   const size_t nt = M->tiles.size(), nm3 = (size_t)M->nc;
   char magic[16];
   int32_t head[4];
   int64_t sizes[2];
   if (std::fread(magic, 1, 16, f.get()) != 16 || std::memcmp(magic, MSH_MAGIC, 16) != 0 ||
       std::fread(head, sizeof(head), 1, f.get()) != 1 || head[0] != 1 ||
       std::fread(sizes, sizeof(sizes), 1, f.get()) != 1)
      return fail(BSAHIP_EINVAL, std::string(path) + " is not a Mesher checkpoint of this version");
   if (head[1] != M->D.nm || head[2] != (int32_t)nt || head[3] != (M->diag ? 1 : 0) ||
       sizes[0] != M->nx)
      return fail(BSAHIP_EINVAL,
                  "checkpoint modes, model or zone layout differ from this problem and configuration");
   std::vector<double> geom(10 * nt), want(10 * nt);
   std::vector<int32_t> rule(nt), rwant(nt);
   msh_geom(M.get(), want.data(), rwant.data());
   if (std::fread(geom.data(), sizeof(double), geom.size(), f.get()) != geom.size() ||
       std::fread(rule.data(), sizeof(int32_t), nt, f.get()) != nt)
      return fail(BSAHIP_EINVAL, std::string(path) + " is truncated");
   if (std::memcmp(geom.data(), want.data(), geom.size() * sizeof(double)) != 0 || rule != rwant)
      return fail(BSAHIP_EINVAL, "checkpoint zone layout differs from this configuration");
   const size_t tot = (size_t)M->nx * M->nx * nm3, chunk = (size_t)1 << 23;
   std::vector<double> h(std::min(tot, chunk));
   for (size_t o = 0; o < tot; o += chunk) {
      const size_t n = std::min(chunk, tot - o);
      if (std::fread(h.data(), sizeof(double), n, f.get()) != n)
         return fail(BSAHIP_EINVAL, std::string(path) + " is truncated");
      HIPC(hipMemcpy(M->lv0 + o, h.data(), n * sizeof(double), hipMemcpyHostToDevice));
   }
   if (std::fgetc(f.get()) != EOF) return fail(BSAHIP_EINVAL, std::string(path) + " has trailing data");
   M->evals_lv0 = (long)sizes[1];
   *out = M.release();
   return 0;
}

int bsahip_mesher_destroy(bsahip_mesher *m)
{
   std::lock_guard<std::recursive_mutex> lk(g.mu);
   delete m;
   return 0;
}

}  // extern "C"
