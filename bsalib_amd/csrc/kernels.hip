// Device kernels of the BFM hot path for gfx950 (CDNA4).
//
// Notation (the reference's, src/bsa/functions/functionsImpl.F90:1312-1587):
//   nodes i,j,k (NNL loaded), modes m,n,p (NM kept), components tc (NTC),
//   Psi_a / Psi_b = phi_times_A_ndegw_(:,:,tc) / (:,:,tc+3),
//   S_ij(f) = rho_ij^|f| sqrt(S_i(f) S_j(f)),   A(n,i;f) = sum_j Psi_a(n,j) S_ji(f).
// The reference's modal force bispectrum is
//   B(m,n,p)(f1,f2) = 2 sum_tc [ G[m](n,f1; p,f2) + G[n](m,f3; p,f2) + G[p](m,f3; n,f1) ],
//   G[x](y,f; z,g)  = sum_i Psi_b(x,i) A(y,i;f) A(z,i;g),          f3 = f1 + f2,
// i.e. its three node sums (functionsImpl.F90:1523-1528) are one tensor G read
// at three argument permutations.  G is symmetric in (y,f)<->(z,g) and depends
// on |f|, |g| only (PSD and coherence are even in f).  k_gsym_f64 computes each
// distinct G value once on FP64 MFMA and its epilogue adds it, weighted by the
// quadrature weight and the modal FRF product, to the up to 12 (m,n,p,f1,f2)
// places it occupies in the plane.
//
// K index of every contraction: k = tc*NNLp + i (NNLp = NNL padded, zeros).
// U table (A values): U[k*ldU + y*nhp + u], u = |f|/df on the half grid.
#include <hip/hip_runtime.h>
#include <stdint.h>

typedef double d4 __attribute__((ext_vector_type(4)));

#define KC 32          // K chunk staged in LDS by k_gsym_f64
#define TILE 64        // G tile edge (rows (y,u) x cols (z,v))

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------
__global__ void k_fill(double *__restrict__ p, size_t n, double v)
{
   size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
   const size_t st = (size_t)gridDim.x * blockDim.x;
   for (; i < n; i += st) p[i] = v;
}

// Nodal one-sided PSD, then its square root, on a list of |f| values.
// formula: 1 von Karman u (WindPSDImpl.F90:166-188), 2 von Karman v/w (:190-222),
// 3 Davenport-Greisch (:249-277), 5 Davenport-Uliege (:283-316), 0 zero (Kaimal,
// :231-243).  Pulsation convention divides by 4 pi (:143).
// prm[(tc*nnl + i)*4 + {0,1,2,3}] = {L, sigma, U_node, U_mean_ref(zone)}.
__global__ void k_psd_sqrt(const double *__restrict__ fabs_list, int nfl,
                           const double *__restrict__ prm, const int *__restrict__ formula,
                           int nnl, int ntc, int pulsation, double *__restrict__ sq)
{
   const int i = blockIdx.x * blockDim.x + threadIdx.x;
   const int q = blockIdx.y;
   const int tc = blockIdx.z;
   if (i >= nnl || q >= nfl) return;
   const double *pp = prm + ((size_t)tc * nnl + i) * 4;
   const double L = pp[0], sg = pp[1], U = pp[2], Um = pp[3];
   const double f = fabs_list[q];
   double S = 0.0;
   switch (formula[tc]) {
   case 1: {
      const double LU = L / U;
      const double x = f * LU;
      S = 4.0 * (LU * sg * sg) / pow(1.0 + 70.7 * x * x, 5.0 / 6.0);
   } break;
   case 2: {
      double d = 2.0 * f * (L / U);
      d = d * d;
      const double r2 = pow(1.0 + 70.7 * d, 11.0 / 6.0) * U;
      S = 4.0 * (L * (1.0 + 188.4 * d) / r2 * (sg * sg));
   } break;
   case 3: {
      const double c = 0.65 * 1200.0;
      const double x = c * f / Um;
      S = (c * sg * sg / Um) / pow(1.0 + x * x, 5.0 / 6.0);
   } break;
   case 5: {
      const double cLU = L / U;
      const double cF = f * cLU;
      S = ((2.0 / 3.0) * cF * cLU * (sg * sg)) / pow(1.0 + cF * cF, 4.0 / 3.0);
   } break;
   default: S = 0.0;
   }
   if (pulsation) S = S / (4.0 * M_PI);
   sq[((size_t)tc * nfl + q) * nnl + i] = sqrt(S);
}

// ln rho_ji for the loaded nodes from the packed coherence vector
// (util_getCorrVectIndex, src/utils/utility.F90:60-79): lnr[c][j][i].
// col[c] is the 0-based nod_corr column of component c; clamp > 0 replaces
// rho < clamp by clamp (BsaCL convention, bsacl.cl:243-245).
__global__ void k_lnrho(const double *__restrict__ corr, long ncorr_rows, const int *__restrict__ col,
                        const int *__restrict__ n_load, int nnl, long nn, double clamp,
                        double *__restrict__ lnr)
{
   const int i = blockIdx.x * blockDim.x + threadIdx.x;
   const int j = blockIdx.y;
   const int c = blockIdx.z;
   if (i >= nnl || j >= nnl) return;
   long a = n_load[j], b = n_load[i];
   if (a > b) {
      const long t = a;
      a = b;
      b = t;
   }
   const long id = (a - 1) * nn + b - (a * a - a) / 2;
   double r = corr[(size_t)col[c] * ncorr_rows + (id - 1)];
   if (clamp > 0.0 && r < clamp) r = clamp;
   lnr[((size_t)c * nnl + j) * nnl + i] = log(r);
}

// Modal FRF table H[m*nft + q] = 1/(K - w^2 M + i w C), w = 2 pi q df, q >= 0
// (functionsImpl.F90:1641-1651).  Negative frequencies use conj.
__global__ void k_frf_table(const double *__restrict__ Mm, const double *__restrict__ Km,
                            const double *__restrict__ Cm, int nm, int nft, double df,
                            double2 *__restrict__ H)
{
   const int q = blockIdx.x * blockDim.x + threadIdx.x;
   const int m = blockIdx.y;
   if (q >= nft || m >= nm) return;
   const double w = 2.0 * M_PI * ((double)q * df);
   const double r = -(w * w * Mm[m]) + Km[m];
   const double im = w * Cm[m];
   const double d = r * r + im * im;
   H[(size_t)m * nft + q] = make_double2(r / d, -im / d);
}

// ---------------------------------------------------------------------------
// projection A(y,i;f) = s_i(f) sum_j Psi_a(y,j) s_j(f) exp(|f| ln rho_ji) on
// FP64 MFMA: per frequency an (MT*16 x nnl) x (nnl x 64) product whose right
// operand, the coherence tile rho_ji^|f|, is generated in LDS chunk by chunk.
// grid (ceil(nnl/64), nfl, ntc * mode chunks), block 256 = 4 waves; a block
// takes the modes y0 .. y0 + 16 MT - 1 of its chunk (y0 = chunk * 16 MT); wave
// w owns the 16 node columns 16w.. of the block's 64 and all MT row blocks of
// 16 modes.
// v_mfma_f64_16x16x4f64: A lane l holds (row l%16, k l/16), B (k l/16, col l%16),
// D register t of lane l holds (row l/16 + 4t, col l%16).
// ---------------------------------------------------------------------------
#define PKC 8    // node (j) chunk staged in LDS

template <int MT>
__global__ __launch_bounds__(256) void k_project(
   const double *__restrict__ pa,      // [tc][nm][nnl]
   const double *__restrict__ lnr,     // [tc][nnl][nnl]  ln rho_ji (row j, col i)
   const double *__restrict__ sq,      // [tc][nfl][nnl]  sqrt(S)
   const double *__restrict__ fabs_list, int nfl, int nnl, int nm, int nnlp,
   int row_stride,                     // rows per mode in U (nhp)
   double *__restrict__ U, long ldU, int ntc)
{
   constexpr int NR = MT * 16;
   __shared__ double ps[PKC][NR];
   __shared__ double ee[PKC][64];
   const int tc = blockIdx.z % ntc;
   const int y0 = (blockIdx.z / ntc) * NR;
   const int q = blockIdx.y;
   const int i0 = blockIdx.x * 64;
   const int tid = threadIdx.x, lane = tid & 63;
   const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
   const double af = fabs_list[q];
   const double *sqq = sq + ((size_t)tc * nfl + q) * nnl;
   const double *lnt = lnr + (size_t)tc * nnl * nnl;
   const double *pat = pa + (size_t)tc * nm * nnl;
   d4 acc[MT];
#pragma unroll
   for (int r = 0; r < MT; ++r) acc[r] = (d4){0.0, 0.0, 0.0, 0.0};

   for (int j0 = 0; j0 < nnl; j0 += PKC) {
      for (int t = tid; t < PKC * NR; t += 256) {
         const int jj = t / NR, y = t % NR;
         const int j = j0 + jj;
         ps[jj][y] = (j < nnl && y0 + y < nm) ? pat[(size_t)(y0 + y) * nnl + j] * sqq[j] : 0.0;
      }
      for (int t = tid; t < PKC * 64; t += 256) {
         const int jj = t >> 6, ii = t & 63;
         const int j = j0 + jj, i = i0 + ii;
         double e = 0.0;
         if (j < nnl && i < nnl) {
            const double l = lnt[(size_t)j * nnl + i];
            e = (af == 0.0) ? 1.0 : exp(af * l);     // rho^0 == 1 even for rho == 0
         }
         ee[jj][ii] = e;
      }
      __syncthreads();
#pragma unroll
      for (int kk = 0; kk < PKC; kk += 4) {
         const int kr = kk + (lane >> 4);
         const double b = ee[kr][16 * w + (lane & 15)];
#pragma unroll
         for (int r = 0; r < MT; ++r)
            acc[r] = __builtin_amdgcn_mfma_f64_16x16x4f64(ps[kr][16 * r + (lane & 15)], b, acc[r],
                                                          0, 0, 0);
      }
      __syncthreads();
   }
   const int i = i0 + 16 * w + (lane & 15);
   if (i < nnl) {
      const double si = sqq[i];
#pragma unroll
      for (int r = 0; r < MT; ++r)
#pragma unroll
         for (int t = 0; t < 4; ++t) {
            const int y = y0 + 16 * r + (lane >> 4) + 4 * t;
            if (y < nm) U[((size_t)tc * nnlp + i) * ldU + (size_t)y * row_stride + q] = acc[r][t] * si;
         }
   }
}

// ---------------------------------------------------------------------------
// quadrature weight and FRF helpers (FP64)
// ---------------------------------------------------------------------------
// rule 0: intgSpectraVect_ (BsaClassicImpl.F90:604-631): interior 1, edge 1/2,
// corner 1 - 3/2;  rule 1: plain Riemann (:331,341);  rule 2: tensor trapezoid.
__device__ __forceinline__ double qweight(int F1, int F2, int nh1, int rule, double d2)
{
   const int a1 = abs(F1), a2 = abs(F2);
   if (a1 > nh1 || a2 > nh1) return 0.0;
   const bool e1 = (a1 == nh1), e2 = (a2 == nh1);
   if (rule == 1) return d2;
   if (rule == 0) return (e1 && e2) ? (1.0 - 1.5) * d2 : ((e1 || e2) ? 0.5 * d2 : d2);
   return d2 * (e1 ? 0.5 : 1.0) * (e2 ? 0.5 : 1.0);
}

__device__ __forceinline__ double2 hget(const double2 *__restrict__ H, int nft, int m, int F)
{
   double2 h = H[(size_t)m * nft + abs(F)];
   if (F < 0) h.y = -h.y;
   return h;
}

// Re[a b conj(c)]  (functionsImpl.F90:1706-1710 / bsacl.cl:892-897)
__device__ __forceinline__ double re3(double2 a, double2 b, double2 c)
{
   const double pr = a.x * b.x - a.y * b.y;
   const double pi = a.x * b.y + a.y * b.x;
   return pr * c.x + pi * c.y;
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
   for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
   return v;
}

// ---------------------------------------------------------------------------
// G contraction on FP64 MFMA + plane epilogue
// grid: (n_tile_pairs * nm) workgroups, 256 threads (4 waves, 2x2 of 32x32)
// ---------------------------------------------------------------------------
struct GsymArgs {
   const double *U;       // [K][ldU]
   long ldU;
   const double *pb;      // [nm][K]  Psi_b per K index
   int K;                 // multiple of KC
   int nm, nhp, nh;       // modes, rows per mode, half-grid points of the plane
   int nu;                // table rows holding A (nh, or 2nh-1 when S(f1+f2)
                          // is evaluated beyond the plane)
   int ntile;             // tiles per side
   long npairs;           // ntile*(ntile+1)/2
   const double2 *H;      // [nm][nft]
   int nft;               // >= 2*nu - 1
   double d2;             // (2 pi df)^2 or df^2
   int rule;
   int epi_plain;         // 1: every 16x16 block through the generic epilogue
   double *m3f, *m3r;     // [nm^3] accumulated (atomic)
   double *slots;         // NULL, or fixed-order mode: each flush of a wave writes
                          // its 12 m3 contributions to slots + ((item*4 + wave)
                          // * mflush + flush) * 12 instead of the atomics
   int mflush;            // flushes per wave at most (fixed-order mode)
};

// Quadrature weight / d^2 from edge flags of |F1| (p) and |F2| (q) on the
// half grid [0, n]: outside the plane 0; corner c_corner; edge c_edge; else 1.
struct QRule {
   double c_edge, c_corner;
};

__device__ __forceinline__ QRule qrule(int rule)
{
   if (rule == 0) return QRule{0.5, 1.0 - 1.5};   // intgSpectraVect_ (BsaClassicImpl.F90:604-631)
   if (rule == 1) return QRule{1.0, 1.0};         // Riemann (BsaClassicImpl.F90:331,341)
   return QRule{0.5, 0.25};                       // tensor trapezoid (_shared_poly2d.fi:827-832)
}

__device__ __forceinline__ double qw(bool ep, bool op, bool eq, bool oq, const QRule &r)
{
   return (op || oq) ? 0.0 : ((ep && eq) ? r.c_corner : ((ep || eq) ? r.c_edge : 1.0));
}

// Plane epilogue of one 16x16 block of G[x] values (rows r0.., cols q0..),
// mode rows y and cols z fixed per block.  Each lane holds four values of
// column q0 + lane%16, rows given by the MFMA accumulator layout: interleaved
// (f64: (l/16) + 4t) or grouped (f32 outputs: 4(l/16) + t).  Adds 2*G times
// the quadrature weight / d^2 and the FRF product to nine per-lane sums:
//   s0 force (x,y,z)=(x,z,y)  s1 resp (x,y,z)  s2 force (y,x,z)=(y,z,x)
//   s3 resp (y,x,z)  s4 resp (y,z,x)  s5 resp (x,z,y)
//   s6 force (z,x,y)=(z,y,x)  s7 resp (z,x,y)  s8 resp (z,y,x)
// (transposed entries s5..s8 only when !DIAG; a compile-time flag keeps the
// four rows of a lane in one basic block, so their loads can overlap).
// Re[a b conj c] is unchanged when all three factors are conjugated, so of the
// eight FRF products these sums weight only four differ:
//   (y,z,x) at (v,u-v):  Re[H_y(v) H_z(u-v) conj H_x(u)]       = F1o
//   (y,z,x) at (-v,u+v): Re[conj H_y(v) H_z(u+v) conj H_x(u)]  = F1s
//   (z,y,x) at (u,v-u):  Re[H_z(u) H_y(v-u) conj H_x(v)]       = F2s
//   (z,y,x) at (-u,u+v): Re[conj H_z(u) H_y(u+v) conj H_x(v)]  = F2o
template <bool DIAG>
__device__ __forceinline__ void epi_accum(const GsymArgs &a, int x, int y, int z, int ur0,
                                          int vq0, int lane, const double (&gval)[4],
                                          bool interleaved, double (&s)[9])
{
   const int n = a.nh - 1;
   const QRule qr = qrule(a.rule);
   const double2 *Hx = a.H + (size_t)x * a.nft;
   const double2 *Hy = a.H + (size_t)y * a.nft;
   const double2 *Hz = a.H + (size_t)z * a.nft;
   // branch-free over the lanes: out-of-range rows/columns read row/column 0
   // and contribute with weight 0 (divergent exits would make the compiler
   // copy the nine accumulators at every join)
   const int vr = vq0 + (lane & 15);
   const bool vok = vr >= 0 && vr < a.nu;          // a block may start in the previous mode
   const int v = vok ? vr : 0;
   const double2 hxv = Hx[v], hyv = Hy[v];
   const double2 cyv = make_double2(hyv.x, -hyv.y);
   const double2 cxv = make_double2(hxv.x, -hxv.y);
   const bool ev = (v == n), ov = (v > n);
#pragma unroll
   for (int t = 0; t < 4; ++t) {
      // rows past the mode's last row (the next mode's rows, or padding) add
      // nothing: stop once the lowest row of this register is one of them (uniform)
      if (ur0 + (interleaved ? 4 * t : t) >= a.nu) break;
      const int ur = ur0 + (interleaved ? ((lane >> 4) + 4 * t) : (4 * (lane >> 4) + t));
      const bool ok = vok && ur >= 0 && ur < a.nu;
      const int u = ok ? ur : 0;
      const double g = ok ? 2.0 * gval[t] : 0.0;   // B = 2 * sum of the three G terms
      const double ms = (u == 0 && v == 0) ? 1.0 : 2.0;
      const double mo = (u > 0 && v > 0) ? 2.0 : 0.0;
      const int sp = u + v;
      const int dm = u - v;
      const int ad = dm < 0 ? -dm : dm;
      const double2 hxu = Hx[u], hzu = Hz[u], hyp = Hy[sp], hzp = Hz[sp];
      double2 hym = Hy[ad], hzm = Hz[ad];
      if (dm < 0) {
         hym.y = -hym.y;
         hzm.y = -hzm.y;
      }
      const bool eu = (u == n), ou = (u > n), es = (sp == n), os = (sp > n);
      const bool ed = (ad == n), od = (ad > n);
      const double Wa = qw(eu, ou, ev, ov, qr);
      const double Wb = qw(ed, od, ev, ov, qr);
      const double Wd = qw(es, os, ev, ov, qr);
      // FRF products Re[H_m(f1) H_n(f2) conj H_p(f1+f2)]; Re[conj w] = Re[w]
      // makes the other eight of the twelve contributions reuse these four:
      const double F1s = re3(hxu, hyv, hzp);      // (x,y,z) at (u,v);    = (z,x,y) at (v+u,-u)
      const double F1o = re3(hxu, cyv, hzm);      // (x,y,z) at (u,-v)
      const double F2s = re3(hym, hxv, hzu);      // (y,x,z) at (u-v,v);  = (x,z,y) at (v,-u)
      const double F2o = re3(hyp, cxv, hzu);      // (y,x,z) at (u+v,-v); = (x,z,y) at (v,u)
      const double ga = g * Wa, gbs = ms * g * Wb, gdo = mo * g * Wd;
      s[0] += ga * (ms + mo);
      s[1] += ga * (ms * F1s + mo * F1o);
      s[2] += gbs + gdo;
      s[3] += gbs * F2s + gdo * F2o;
      s[4] += gbs * F1o + gdo * F1s;
      if constexpr (!DIAG) {
         const double We = qw(ed, od, eu, ou, qr);
         const double Wf = qw(es, os, eu, ou, qr);
         const double ges = ms * g * We, gfo = mo * g * Wf;
         s[5] += ga * (ms * F2o + mo * F2s);
         s[6] += ges + gfo;
         s[7] += ges * F1o + gfo * F1s;
         s[8] += ges * F2s + gfo * F2o;
      }
   }
}

// The same nine sums for a 16x16 block whose every point has the weight of
// the plane's interior in all five quadrature weights (CLS 1: 1 <= u, v and
// u + v <= n - 1) or lies beyond the plane's f1+f2 edge with the other four
// weights interior (CLS 2: u, v <= n - 1 and u + v >= n + 1, so Wd = Wf = 0).
// CLS 3: the band between them (u, v, |u-v| interior, u + v on both sides of
// n), where Wd = Wf = w(u+v): 1 below n, the edge weight at n, 0 beyond.
// With Wa = Wb = We = 1, ms = mo = 2 and h = 2 g = 4 G the contributions of
// epi_accum reduce to three (CLS 1) or five (CLS 2, 3) distinct sums, kept per
// block and spread over the nine at its end; no lane is out of range.
template <int CLS, bool DIAG, bool WIN>
__device__ __forceinline__ void epi_accum_inner(const GsymArgs &a, int x, int y, int z, int ur0,
                                                int vq0, int lane, const double (&gval)[4],
                                                bool interleaved, double (&s)[9],
                                                const double2 *win, int wub, int wvb, int wdb)
{
   const double2 *Hx = a.H + (size_t)x * a.nft;
   const double2 *Hy = a.H + (size_t)y * a.nft;
   const double2 *Hz = a.H + (size_t)z * a.nft;
   const int v = vq0 + (lane & 15);
   const int wsb = wub + wvb;
   double2 hxv, hyv;
   if constexpr (WIN) {
      hxv = win[640 + v - wvb];
      hyv = win[704 + v - wvb];
   } else {
      hxv = Hx[v];
      hyv = Hy[v];
   }
   const double2 cyv = make_double2(hyv.x, -hyv.y);
   const double2 cxv = make_double2(hxv.x, -hxv.y);
   // lh = sum h, lA = sum h (F1s + F1o), lB = sum h F2s (CLS 2), h (F2s + F2o)
   // (CLS 1) or h (F2s + w F2o) (CLS 3), lC = sum h F1o (CLS 2) or h (F1o + w F1s)
   // (CLS 3), lD = sum h (F2s + F2o) (CLS 2, 3), lW = sum h w (CLS 3)
   const int n = a.nh - 1;
   const double ce = qrule(a.rule).c_edge;
   double lh = 0.0, lA = 0.0, lB = 0.0, lC = 0.0, lD = 0.0, lW = 0.0;
#pragma unroll
   for (int t = 0; t < 4; ++t) {
      const int u = ur0 + (interleaved ? ((lane >> 4) + 4 * t) : (4 * (lane >> 4) + t));
      const double h = 4.0 * gval[t];
      const int sp = u + v;
      const int dm = u - v;
      const int ad = dm < 0 ? -dm : dm;
      double2 hxu, hzu, hzp, hym, hzm;
      if constexpr (WIN) {                           // the wave's LDS window
         hxu = win[512 + u - wub];
         hzu = win[576 + u - wub];
         hzp = win[sp - wsb];
         hym = win[256 + dm - wdb];
         hzm = win[384 + dm - wdb];
      } else {
         hxu = Hx[u];
         hzu = Hz[u];
         hzp = Hz[sp];
         hym = Hy[ad];
         hzm = Hz[ad];
         if (dm < 0) {
            hym.y = -hym.y;
            hzm.y = -hzm.y;
         }
      }
      auto hys = [&]() {                                              // H_y(u+v)
         if constexpr (WIN) return win[128 + sp - wsb];
         else return Hy[sp];
      };
      const double F1s = re3(hxu, hyv, hzp);
      const double F1o = re3(hxu, cyv, hzm);
      const double F2s = re3(hym, hxv, hzu);
      lh += h;
      lA += h * (F1s + F1o);
      if constexpr (CLS == 1) {
         const double F2o = re3(hys(), cxv, hzu);
         lB += h * (F2s + F2o);
      } else if constexpr (CLS == 2) {
         lB += h * F2s;
         lC += h * F1o;
         if constexpr (!DIAG) {
            const double F2o = re3(hys(), cxv, hzu);
            lD += h * (F2s + F2o);
         }
      } else {
         const double w = sp < n ? 1.0 : (sp == n ? ce : 0.0);
         const double F2o = re3(hys(), cxv, hzu);
         const double hw = h * w;
         lW += hw;
         lB += h * F2s + hw * F2o;
         lC += h * F1o + hw * F1s;
         if constexpr (!DIAG) lD += h * (F2s + F2o);
      }
   }
   if constexpr (CLS == 1) {
      // s0 = s2 = s6 = 2h, s1 = s4 = s7 = h(F1s+F1o), s3 = s5 = s8 = h(F2s+F2o)
      s[0] += 2.0 * lh;
      s[1] += lA;
      s[2] += 2.0 * lh;
      s[3] += lB;
      s[4] += lA;
      if constexpr (!DIAG) {
         s[5] += lB;
         s[6] += 2.0 * lh;
         s[7] += lA;
         s[8] += lB;
      }
   } else if constexpr (CLS == 3) {
      s[0] += 2.0 * lh;
      s[1] += lA;
      s[2] += lh + lW;
      s[3] += lB;
      s[4] += lC;
      if constexpr (!DIAG) {
         s[5] += lD;
         s[6] += lh + lW;
         s[7] += lC;
         s[8] += lB;
      }
   } else {
      // beyond f1+f2 = fmax: Wd = Wf = 0
      s[0] += 2.0 * lh;
      s[1] += lA;
      s[2] += lh;
      s[3] += lB;
      s[4] += lC;
      if constexpr (!DIAG) {
         s[5] += lD;
         s[6] += lh;
         s[7] += lC;
         s[8] += lB;
      }
   }
}

// Class of a 16x16 block with half-grid origin (u0, v0): 1, 2 or 3 as above,
// 0 for every other block (edges of the plane, rows beyond it).
__device__ __forceinline__ int epi_class(const GsymArgs &a, int u0, int v0)
{
   const int n = a.nh - 1;
   if (a.epi_plain || u0 < 1 || v0 < 1 || u0 + 15 > n - 1 || v0 + 15 > n - 1) return 0;
   if (u0 + v0 + 30 <= n - 1) return 1;
   if (u0 + v0 >= n + 1) return 2;
   return 3;
}

template <bool WIN>
__device__ __forceinline__ void epi_accum_any(const GsymArgs &a, int x, int y, int z, bool diag,
                                              int ur0, int vq0, int lane, const double (&g)[4],
                                              bool interleaved, double (&s)[9],
                                              const double2 *win, int wub, int wvb, int wdb)
{
   const int cls = epi_class(a, ur0, vq0);           // uniform per block
   if (diag) {
      if (cls == 1)
         epi_accum_inner<1, true, WIN>(a, x, y, z, ur0, vq0, lane, g, interleaved, s, win, wub, wvb, wdb);
      else if (cls == 2)
         epi_accum_inner<2, true, WIN>(a, x, y, z, ur0, vq0, lane, g, interleaved, s, win, wub, wvb, wdb);
      else if (cls == 3)
         epi_accum_inner<3, true, WIN>(a, x, y, z, ur0, vq0, lane, g, interleaved, s, win, wub, wvb, wdb);
      else
         epi_accum<true>(a, x, y, z, ur0, vq0, lane, g, interleaved, s);
   } else {
      if (cls == 1)
         epi_accum_inner<1, false, WIN>(a, x, y, z, ur0, vq0, lane, g, interleaved, s, win, wub, wvb, wdb);
      else if (cls == 2)
         epi_accum_inner<2, false, WIN>(a, x, y, z, ur0, vq0, lane, g, interleaved, s, win, wub, wvb, wdb);
      else if (cls == 3)
         epi_accum_inner<3, false, WIN>(a, x, y, z, ur0, vq0, lane, g, interleaved, s, win, wub, wvb, wdb);
      else
         epi_accum<false>(a, x, y, z, ur0, vq0, lane, g, interleaved, s);
   }
}

// Wave-reduce the nine sums and add them (times d^2) to the m3 entries, or
// in the fixed-order mode store them in the wave's slot `f` (the order of the
// twelve values is the one of the atomics below; bsahip_api.hip builds the
// matching slot lists, det_slot_lists).
__device__ __forceinline__ void epi_flush(const GsymArgs &a, int x, int y, int z, bool diag,
                                          int lane, double (&s)[9], long sbase, int f)
{
   const int nm = a.nm;
   const size_t nm2 = (size_t)nm * nm;
#pragma unroll
   for (int t = 0; t < 9; ++t) s[t] = wave_sum(s[t]) * a.d2;
   if (a.slots) {
      if (lane == 0) {
         double *d = a.slots + sbase + (long)f * 12;
         d[0] = s[0];
         d[1] = s[1];
         d[2] = s[2];
         d[3] = s[3];
         d[4] = s[2];
         d[5] = s[4];
         if (!diag) {
            d[6] = s[0];
            d[7] = s[5];
            d[8] = s[6];
            d[9] = s[7];
            d[10] = s[6];
            d[11] = s[8];
         }
      }
#pragma unroll
      for (int t = 0; t < 9; ++t) s[t] = 0.0;
      return;
   }
   if (lane == 0) {
      const size_t i_xyz = x + (size_t)nm * y + nm2 * z;
      const size_t i_yxz = y + (size_t)nm * x + nm2 * z;
      const size_t i_yzx = y + (size_t)nm * z + nm2 * x;
      atomicAdd(a.m3f + i_xyz, s[0]);
      atomicAdd(a.m3r + i_xyz, s[1]);
      atomicAdd(a.m3f + i_yxz, s[2]);
      atomicAdd(a.m3r + i_yxz, s[3]);
      atomicAdd(a.m3f + i_yzx, s[2]);
      atomicAdd(a.m3r + i_yzx, s[4]);
      if (!diag) {
         const size_t i_xzy = x + (size_t)nm * z + nm2 * y;
         const size_t i_zxy = z + (size_t)nm * x + nm2 * y;
         const size_t i_zyx = z + (size_t)nm * y + nm2 * x;
         atomicAdd(a.m3f + i_xzy, s[0]);
         atomicAdd(a.m3r + i_xzy, s[5]);
         atomicAdd(a.m3f + i_zxy, s[6]);
         atomicAdd(a.m3r + i_zxy, s[7]);
         atomicAdd(a.m3f + i_zyx, s[6]);
         atomicAdd(a.m3r + i_zyx, s[8]);
      }
   }
#pragma unroll
   for (int t = 0; t < 9; ++t) s[t] = 0.0;
}

// Epilogue of a wave's BR x BC blocks of 16x16 (rows rw0.., cols cw0..): one
// accumulation / reduction when all blocks share (y,z), else one per (y,z) pair.
// WIN: win holds 768 double2 of LDS owned by this wave (free after the K loop).
// When all blocks share (y,z), the FRF values the interior classes read per
// point are staged there first, each indexed by its offset from the wave's
// origin (ub, vb) within the modes:
//   [0, 128)   H_z(u+v)      at u+v - (ub+vb)
//   [128, 256) H_y(u+v)      at u+v - (ub+vb)
//   [256, 384) H_y(u-v)      at u-v - (ub-vb-16BC+1), conjugated for u < v
//   [384, 512) H_z(u-v)      likewise
//   [512, 576) H_x(u), [576, 640) H_z(u)   at u - ub
//   [640, 704) H_x(v), [704, 768) H_y(v)   at v - vb
// (u+v and u-v span 16(BR+BC)-1 <= 127 values over the wave's blocks).
template <int BR, int BC, bool WIN, class GetVal>
__device__ __forceinline__ void epi_wave(const GsymArgs &a, int x, bool diag, int rw0, int cw0,
                                         int lane, bool interleaved, long sbase, GetVal getv,
                                         double2 *win)
{
   static_assert(16 * (BR + BC) - 1 <= 128 && 16 * BR <= 64 && 16 * BC <= 64,
                 "FRF windows of 128 and 64 entries");
   const int nm = a.nm;
   int f = 0;                                         // flushes of this wave so far
   const int y0 = rw0 / a.nhp, y1 = (rw0 + 16 * BR - 1) / a.nhp;
   const int z0 = cw0 / a.nhp, z1 = (cw0 + 16 * BC - 1) / a.nhp;
   double s[9];
#pragma unroll
   for (int t = 0; t < 9; ++t) s[t] = 0.0;
   if (y0 == y1 && z0 == z1) {
      if (y0 >= nm || z0 >= nm) return;
      const int ub = rw0 - y0 * a.nhp, vb = cw0 - z0 * a.nhp;
      const int wsb = ub + vb, wdb = ub - vb - (16 * BC - 1);
      if constexpr (WIN) {
         const double2 *Hy = a.H + (size_t)y0 * a.nft;
         const double2 *Hz = a.H + (size_t)z0 * a.nft;
         for (int e = lane; e < 128; e += 64) {
            const int sp = min(wsb + e, a.nft - 1);   // entries past the wave's range unused
            const int dm = wdb + e;
            const int ad = min(dm < 0 ? -dm : dm, a.nft - 1);
            double2 hym = Hy[ad], hzm = Hz[ad];
            if (dm < 0) {
               hym.y = -hym.y;
               hzm.y = -hzm.y;
            }
            win[e] = Hz[sp];
            win[128 + e] = Hy[sp];
            win[256 + e] = hym;
            win[384 + e] = hzm;
         }
         if (lane < 16 * BR) {
            const double2 *Hx = a.H + (size_t)x * a.nft;
            win[512 + lane] = Hx[ub + lane];
            win[576 + lane] = Hz[ub + lane];
         }
         if (lane < 16 * BC) {
            const double2 *Hx = a.H + (size_t)x * a.nft;
            win[640 + lane] = Hx[vb + lane];
            win[704 + lane] = Hy[vb + lane];
         }
         // the wave's own LDS writes before its reads (other waves never touch it)
         __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
         __builtin_amdgcn_wave_barrier();
         __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      }
#pragma unroll 1
      for (int b = 0; b < BR * BC; ++b) {
         const int i = b / BC, j = b % BC;
         double g[4];
         getv(b, g);
         epi_accum_any<WIN>(a, x, y0, z0, diag, ub + 16 * i, vb + 16 * j, lane, g, interleaved, s,
                            win, ub, vb, wdb);
      }
      epi_flush(a, x, y0, z0, diag, lane, s, sbase, f++);
   } else {
      // the wave's rows (cols) span modes y0..y1 (z0..z1): one reduction per
      // (y, z) pair over the blocks that meet both.  The rows of consecutive
      // modes are packed, so a 16-block can hold the end of one mode and the
      // start of the next: it is visited for each, the rows (cols) of the
      // other mode masked by the generic epilogue (u < 0 or u >= nu there)
#pragma unroll 1
      for (int y = y0; y <= y1; ++y)
#pragma unroll 1
         for (int z = z0; z <= z1; ++z) {
            if (y >= nm || z >= nm) continue;          // uniform per wave
#pragma unroll 1
            for (int b = 0; b < BR * BC; ++b) {
               const int i = b / BC, j = b % BC;
               const int r0 = rw0 + 16 * i, q0 = cw0 + 16 * j;
               if (r0 / a.nhp > y || (r0 + 15) / a.nhp < y || q0 / a.nhp > z ||
                   (q0 + 15) / a.nhp < z)
                  continue;
               double g[4];
               getv(b, g);
               epi_accum_any<false>(a, x, y, z, diag, r0 - y * a.nhp, q0 - z * a.nhp, lane, g,
                                    interleaved, s, nullptr, 0, 0, 0);
            }
            epi_flush(a, x, y, z, diag, lane, s, sbase, f++);
         }
   }
}

__device__ __forceinline__ void pair_from_index(long t, int &R, int &Q)
{
   long q = (long)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
   while ((q + 1) * (q + 2) / 2 <= t) ++q;
   while (q * (q + 1) / 2 > t) --q;
   Q = (int)q;
   R = (int)(t - q * (q + 1) / 2);
}

// XCD-aware order: workgroups dispatched to one XCD take consecutive logical
// ids, so the nm modes x that share one (R,Q) tile pair run on the same L2.
__device__ __forceinline__ long xcd_lid()
{
   const long nwg = (long)gridDim.x;
   const long wid = blockIdx.x;
   long lid = wid;
   const long per = nwg / 8;
   if (wid < per * 8) lid = (wid % 8) * per + wid / 8;
   return lid;
}

// One work item of the FP64 G contraction: mode x, tile pair tp.
__device__ __forceinline__ void gsym_f64_item(const GsymArgs &a, int x, long tp)
{
   __shared__ double As[KC][TILE];
   __shared__ double Bs[KC][TILE];
   int R, Q;
   pair_from_index(tp, R, Q);
   const bool diag = (R == Q);

   const int tid = threadIdx.x;
   const int lane = tid & 63;
   const int w = __builtin_amdgcn_readfirstlane(tid >> 6);   // wave-uniform
   const int wr = w >> 1, wc = w & 1;

   // staging: each thread moves 8 consecutive doubles of one k row per tile
   const int sk = tid >> 3;            // 0..31
   const int so = (tid & 7) * 8;       // 0..56
   const double *gA = a.U + (size_t)R * TILE + so;
   const double *gB = a.U + (size_t)Q * TILE + so;
   const double *pbx = a.pb + (size_t)x * a.K;

   d4 acc[2][2];
#pragma unroll
   for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = (d4){0.0, 0.0, 0.0, 0.0};

   double ra[8], rb[8];
   double sc;
   {
      const double *pA = gA + (size_t)sk * a.ldU;
      const double *pB = gB + (size_t)sk * a.ldU;
#pragma unroll
      for (int t = 0; t < 8; ++t) { ra[t] = pA[t]; rb[t] = pB[t]; }
      sc = pbx[sk];
   }
   __builtin_amdgcn_s_setprio(1);   // K loop at raised wave priority, as k_gsym_bf16x3
   for (int k0 = 0; k0 < a.K; k0 += KC) {
      __syncthreads();
#pragma unroll
      for (int t = 0; t < 8; ++t) { As[sk][so + t] = ra[t] * sc; Bs[sk][so + t] = rb[t]; }
      __syncthreads();
      if (k0 + KC < a.K) {
         const double *pA = gA + (size_t)(k0 + KC + sk) * a.ldU;
         const double *pB = gB + (size_t)(k0 + KC + sk) * a.ldU;
#pragma unroll
         for (int t = 0; t < 8; ++t) { ra[t] = pA[t]; rb[t] = pB[t]; }
         sc = pbx[k0 + KC + sk];
      }
#pragma unroll
      for (int kk = 0; kk < KC; kk += 4) {
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

   // ---- epilogue: for v_mfma_f64_16x16x4f64, register t of lane l holds
   // D[(l/16) + 4t][l%16] of its 16x16 block (rows interleaved by 4).
   const long sbase = ((tp * a.nm + x) * 4 + w) * (long)a.mflush * 12;
   __builtin_amdgcn_s_setprio(0);
   epi_wave<2, 2, false>(a, x, diag, R * TILE + wr * 32, Q * TILE + wc * 32, lane, true, sbase,
                  [&](int b, double (&g)[4]) {
                     d4 v;
                     switch (b) {       // uniform branch: constant register indices
                     case 0: v = acc[0][0]; break;
                     case 1: v = acc[0][1]; break;
                     case 2: v = acc[1][0]; break;
                     default: v = acc[1][1]; break;
                     }
#pragma unroll
                     for (int t = 0; t < 4; ++t) g[t] = v[t];
                  },
                  nullptr);
}

// Four workgroups per CU (at most 128 VGPRs: a few spill in the epilogue):
// more waves hide the barriers and loads of the single-buffered K chunks,
// 337 -> 277 ms at the deck against two workgroups per CU.
__global__ __launch_bounds__(256, 4) void k_gsym_f64(GsymArgs a)
{
   const long lid = xcd_lid();
   gsym_f64_item(a, (int)(lid % a.nm), lid / a.nm);
}

// ---------------------------------------------------------------------------
// G contraction in split-bf16 x3 on v_mfma_f32_16x16x32_bf16 (FP32 accumulate)
//   a*b ~= a_hi*b_hi + a_hi*b_lo + a_lo*b_hi,  a_hi = bf16(a), a_lo = bf16(a - a_hi)
// Operands come from the FP32 copy of the A table; the row operand is scaled
// by Psi_b(x,k) in FP32 before the split.  Same plane epilogue (FP64) as
// k_gsym_f64.  Workgroup tile 128x128, 4 waves of 64x64 (4x4 blocks of
// 16x16), K chunk 32 double-buffered in LDS.  LDS rows (64 bytes) hold four
// 16-byte k-groups stored at slot g ^ ((row >> 1) & 3): the 8 rows 0-7 (8-15)
// of a fragment read fill the 8 distinct 16-byte slots of a 128-byte line
// (SQ_LDS_BANK_CONFLICT reads zero for this kernel).
// ---------------------------------------------------------------------------
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf2v __attribute__((ext_vector_type(2)));
typedef float f2v __attribute__((ext_vector_type(2)));
typedef float f4v __attribute__((ext_vector_type(4)));

#define T2 128
#define KC2 32

__device__ __forceinline__ uint32_t bf16_rne_bits(float f)
{
   const uint32_t b = __float_as_uint(f);
   return (b + 0x7FFFu + ((b >> 16) & 1u)) >> 16;
}

__device__ __forceinline__ void split_bf16(float v, uint32_t &hi, uint32_t &lo)
{
   hi = bf16_rne_bits(v);
   const float r = v - __uint_as_float(hi << 16);     // exact in FP32
   lo = bf16_rne_bits(r);
}

__device__ __forceinline__ int swz(int row, int g) { return row * 4 + (g ^ ((row >> 1) & 3)); }

struct GsymArgs16 {
   const float *UR;       // [rows][ldk]  A table, FP32, K contiguous (row operand source)
   const uint4 *UCh;      // [rows][ldk/8] bf16 hi of the A table (column operand)
   const uint4 *UCl;      // [rows][ldk/8] bf16 lo
   const float *pb32;     // [nm][ldk]  Psi_b per K index, FP32
   long ldk;              // K stride (multiple of 32)
};

// One work item of the split-bf16x3 G contraction: mode x, tile pair tp.
__device__ __forceinline__ void gsym_bf16x3_item(const GsymArgs &a, const GsymArgs16 &b, int x,
                                                 long tp)
{
   __shared__ uint4 Ah[2][T2 * 4], Al[2][T2 * 4], Bh[2][T2 * 4], Bl[2][T2 * 4];
   int R, Q;
   pair_from_index(tp, R, Q);
   const bool diag = (R == Q);

   const int tid = threadIdx.x;
   const int lane = tid & 63;
   const int w = __builtin_amdgcn_readfirstlane(tid >> 6);   // wave-uniform
   const int wr = w >> 1, wc = w & 1;

   // staging: items tid and tid+256 -> (row = item/4, k-group = item%4 of 8)
   const int srow0 = tid >> 2, srow1 = srow0 + 64;
   const int sgrp = tid & 3;
   const long ldk = b.ldk;
   const float *gA0 = b.UR + ((size_t)R * T2 + srow0) * ldk + sgrp * 8;
   const float *gA1 = b.UR + ((size_t)R * T2 + srow1) * ldk + sgrp * 8;
   const uint4 *gBh0 = b.UCh + (((size_t)Q * T2 + srow0) * ldk + sgrp * 8) / 8;
   const uint4 *gBh1 = b.UCh + (((size_t)Q * T2 + srow1) * ldk + sgrp * 8) / 8;
   const uint4 *gBl0 = b.UCl + (((size_t)Q * T2 + srow0) * ldk + sgrp * 8) / 8;
   const uint4 *gBl1 = b.UCl + (((size_t)Q * T2 + srow1) * ldk + sgrp * 8) / 8;
   const float *pbx = b.pb32 + (size_t)x * ldk + sgrp * 8;

   f4v acc[4][4];
#pragma unroll
   for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = (f4v){0.f, 0.f, 0.f, 0.f};

   float4 ra0, ra1, ra2, ra3, sc0, sc1;
   uint4 rbh0, rbh1, rbl0, rbl1;
   auto gload = [&](int k0) {
      ra0 = *(const float4 *)(gA0 + k0);
      ra1 = *(const float4 *)(gA0 + k0 + 4);
      ra2 = *(const float4 *)(gA1 + k0);
      ra3 = *(const float4 *)(gA1 + k0 + 4);
      sc0 = *(const float4 *)(pbx + k0);
      sc1 = *(const float4 *)(pbx + k0 + 4);
      rbh0 = gBh0[k0 / 8];
      rbh1 = gBh1[k0 / 8];
      rbl0 = gBl0[k0 / 8];
      rbl1 = gBl1[k0 / 8];
   };
   auto split2 = [](float v0, float v1, uint32_t &h, uint32_t &l) {
      const f2v vv = (f2v){v0, v1};
      const bf2v hv = __builtin_convertvector(vv, bf2v);              // v_cvt_pk_bf16_f32 (RNE)
      const f2v r = vv - __builtin_convertvector(hv, f2v);             // exact in FP32
      h = __builtin_bit_cast(uint32_t, hv);
      l = __builtin_bit_cast(uint32_t, __builtin_convertvector(r, bf2v));
   };
   auto split8 = [&](float4 p0, float4 p1, uint4 &hi, uint4 &lo) {
      split2(p0.x * sc0.x, p0.y * sc0.y, hi.x, lo.x);
      split2(p0.z * sc0.z, p0.w * sc0.w, hi.y, lo.y);
      split2(p1.x * sc1.x, p1.y * sc1.y, hi.z, lo.z);
      split2(p1.z * sc1.z, p1.w * sc1.w, hi.w, lo.w);
   };
   auto lstore = [&](int buf) {
      uint4 hi, lo;
      split8(ra0, ra1, hi, lo);
      Ah[buf][swz(srow0, sgrp)] = hi;
      Al[buf][swz(srow0, sgrp)] = lo;
      split8(ra2, ra3, hi, lo);
      Ah[buf][swz(srow1, sgrp)] = hi;
      Al[buf][swz(srow1, sgrp)] = lo;
      Bh[buf][swz(srow0, sgrp)] = rbh0;
      Bh[buf][swz(srow1, sgrp)] = rbh1;
      Bl[buf][swz(srow0, sgrp)] = rbl0;
      Bl[buf][swz(srow1, sgrp)] = rbl1;
   };

   const int nchunk = a.K / KC2;
   // the K loop at raised wave priority, the epilogue at the default: where the
   // two workgroups of a CU are in different phases, the MFMA-issuing waves are
   // served first and the epilogue's FP64 VALU work fills the gaps (measured:
   // 1-1.4 % shorter kernel; the reverse order gains nothing)
   __builtin_amdgcn_s_setprio(1);
   gload(0);
   __syncthreads();      // a previous item's epilogue (persistent batch) may still read LDS
   lstore(0);
   __syncthreads();
   for (int c = 0; c < nchunk; ++c) {
      const int buf = c & 1;
      if (c + 1 < nchunk) gload((c + 1) * KC2);
      const int g = lane >> 4;
      bf8 ah[4], al[4], bh[4], bl[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
         const int row = wr * 64 + i * 16 + (lane & 15);
         const int col = wc * 64 + i * 16 + (lane & 15);
         ah[i] = __builtin_bit_cast(bf8, Ah[buf][swz(row, g)]);
         al[i] = __builtin_bit_cast(bf8, Al[buf][swz(row, g)]);
         bh[i] = __builtin_bit_cast(bf8, Bh[buf][swz(col, g)]);
         bl[i] = __builtin_bit_cast(bf8, Bl[buf][swz(col, g)]);
      }
      // small products first; product-major order keeps consecutive MFMAs independent
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
         for (int j = 0; j < 4; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
         for (int j = 0; j < 4; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
         for (int j = 0; j < 4; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
      if (c + 1 < nchunk) lstore(buf ^ 1);
      __syncthreads();
   }

   __builtin_amdgcn_s_setprio(0);
   // epilogue: f32-output 16x16 MFMAs hold D[4(l/16)+t][l%16]
   const long sbase = ((tp * a.nm + x) * 4 + w) * (long)a.mflush * 12;
   // each wave's FRF window: 12 KB of one operand plane (16 KB), free after the
   // last chunk's barrier
   uint4 *plane = w == 0 ? &Ah[0][0] : (w == 1 ? &Al[0][0] : (w == 2 ? &Bh[0][0] : &Bl[0][0]));
   double2 *win = reinterpret_cast<double2 *>(plane);
   epi_wave<4, 4, true>(a, x, diag, R * T2 + wr * 64, Q * T2 + wc * 64, lane, false, sbase,
                  [&](int b, double (&g)[4]) {
                     f4v v;
                     switch (b) {       // uniform branch: constant register indices
#define CASE_(k) case k: v = acc[k / 4][k % 4]; break;
                     CASE_(0) CASE_(1) CASE_(2) CASE_(3) CASE_(4) CASE_(5) CASE_(6) CASE_(7)
                     CASE_(8) CASE_(9) CASE_(10) CASE_(11) CASE_(12) CASE_(13) CASE_(14)
#undef CASE_
                     default: v = acc[3][3]; break;
                     }
#pragma unroll
                     for (int t = 0; t < 4; ++t) g[t] = (double)v[t];
                  },
                  win);
}

__global__ __launch_bounds__(256, 2) void k_gsym_bf16x3(GsymArgs a, GsymArgs16 b)
{
   const long lid = xcd_lid();
   gsym_bf16x3_item(a, b, (int)(lid % a.nm), lid / a.nm);
}

// ---------------------------------------------------------------------------
// G contraction in FP32 on v_mfma_f32_16x16x4f32 (FP32 operands, FP32
// accumulate); same plane epilogue (FP64) as k_gsym_f64.  Workgroup tile
// 128x128, 4 waves of 64x64 (4x4 blocks of 16x16), K chunk 32 double-buffered
// in LDS (two FP32 planes, 64 KB).  Within a chunk, MFMA step s (0..7) takes
// k = 8 (l/16) + s for lane l in both operands (any assignment of the chunk's
// 32 k to (step, lane group) gives the same sum), so a lane reads its 8 k of
// one row as two 16-byte LDS loads.  LDS rows (128 bytes) hold eight 16-byte
// k-groups stored at slot q ^ (row & 7).  Measured at NN=1000, NM=30: half the
// bank-conflict cycles of slot q ^ ((row >> 1) & 7) and 2.4 % faster; slot
// q ^ ((row ^ (row >> 3)) & 7) has none but ran 1.7 % slower than this one.
// ---------------------------------------------------------------------------
#define KC4 32

__device__ __forceinline__ int swz8(int row, int q) { return row * 8 + (q ^ (row & 7)); }

__device__ __forceinline__ void gsym_f32_item(const GsymArgs &a, const GsymArgs16 &b, int x, long tp)
{
   __shared__ float4 As[2][T2 * 8], Bs[2][T2 * 8];
   int R, Q;
   pair_from_index(tp, R, Q);
   const bool diag = (R == Q);

   const int tid = threadIdx.x;
   const int lane = tid & 63;
   const int w = __builtin_amdgcn_readfirstlane(tid >> 6);   // wave-uniform
   const int wr = w >> 1, wc = w & 1;

   // staging: thread -> (row = tid/2, k-groups 4h..4h+3 of 4 floats), h = tid%2
   const int srow = tid >> 1, sh = tid & 1;
   const long ldk = b.ldk;
   const float *gA = b.UR + ((size_t)R * T2 + srow) * ldk + sh * 16;
   const float *gB = b.UR + ((size_t)Q * T2 + srow) * ldk + sh * 16;
   const float *pbx = b.pb32 + (size_t)x * ldk + sh * 16;

   f4v acc[4][4];
#pragma unroll
   for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = (f4v){0.f, 0.f, 0.f, 0.f};

   float4 ra[4], rb[4], sc[4];
   auto gload = [&](int k0) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
         ra[j] = *(const float4 *)(gA + k0 + 4 * j);
         rb[j] = *(const float4 *)(gB + k0 + 4 * j);
         sc[j] = *(const float4 *)(pbx + k0 + 4 * j);
      }
   };
   auto lstore = [&](int buf) {
#pragma unroll
      for (int j = 0; j < 4; ++j) {
         const float4 v = ra[j], s = sc[j];
         As[buf][swz8(srow, 4 * sh + j)] = make_float4(v.x * s.x, v.y * s.y, v.z * s.z, v.w * s.w);
         Bs[buf][swz8(srow, 4 * sh + j)] = rb[j];
      }
   };

   const int nchunk = a.K / KC4;
   __builtin_amdgcn_s_setprio(1);   // K loop at raised wave priority, as k_gsym_bf16x3
   gload(0);
   __syncthreads();      // a previous item's epilogue (persistent batch) may still read LDS
   lstore(0);
   __syncthreads();
   const int g = lane >> 4;            // lane group: k = 8g .. 8g+7 of the chunk
   for (int c = 0; c < nchunk; ++c) {
      const int buf = c & 1;
      if (c + 1 < nchunk) gload((c + 1) * KC4);
#pragma unroll
      for (int hsel = 0; hsel < 2; ++hsel) {       // steps 4 hsel .. 4 hsel + 3
         float4 fa[4], fb[4];
#pragma unroll
         for (int i = 0; i < 4; ++i) {
            const int row = wr * 64 + i * 16 + (lane & 15);
            const int col = wc * 64 + i * 16 + (lane & 15);
            fa[i] = As[buf][swz8(row, 2 * g + hsel)];
            fb[i] = Bs[buf][swz8(col, 2 * g + hsel)];
         }
#pragma unroll
         for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int i = 0; i < 4; ++i)
#pragma unroll
               for (int j = 0; j < 4; ++j)
                  acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(
                     ((const float *)&fa[i])[s], ((const float *)&fb[j])[s], acc[i][j], 0, 0, 0);
      }
      if (c + 1 < nchunk) lstore(buf ^ 1);
      __syncthreads();
   }

   // epilogue: f32-output 16x16 MFMAs hold D[4(l/16)+t][l%16]
   const long sbase = ((tp * a.nm + x) * 4 + w) * (long)a.mflush * 12;
   __builtin_amdgcn_s_setprio(0);
   // each wave's FRF window: 12 KB of the operand planes (2 x 32 KB), free after
   // the last chunk's barrier
   double2 *win = reinterpret_cast<double2 *>(w < 2 ? &As[0][0] : &Bs[0][0]) + (w & 1) * 768;
   epi_wave<4, 4, true>(a, x, diag, R * T2 + wr * 64, Q * T2 + wc * 64, lane, false, sbase,
                  [&](int bb, double (&gv)[4]) {
                     f4v v;
                     switch (bb) {      // uniform branch: constant register indices
#define CASE_(k) case k: v = acc[k / 4][k % 4]; break;
                     CASE_(0) CASE_(1) CASE_(2) CASE_(3) CASE_(4) CASE_(5) CASE_(6) CASE_(7)
                     CASE_(8) CASE_(9) CASE_(10) CASE_(11) CASE_(12) CASE_(13) CASE_(14)
#undef CASE_
                     default: v = acc[3][3]; break;
                     }
#pragma unroll
                     for (int t = 0; t < 4; ++t) gv[t] = (double)v[t];
                  },
                  win);
}

__global__ __launch_bounds__(256, 2) void k_gsym_f32(GsymArgs a, GsymArgs16 b)
{
   const long lid = xcd_lid();
   gsym_f32_item(a, b, (int)(lid % a.nm), lid / a.nm);
}

// Lane-layout self test of v_mfma_f32_16x16x4f32 (one wave, 512 floats):
// out[t*64 + l] = D[4(l/16)+t][l%16] of A[i][k] = i + 16 k, B = identity
// columns k == j % 4 (so D[i][j] = i + 16 (j % 4)); out[256 + t*64 + l] of
// A[i][k] = 1, B[k][j] = k + 4 j (D[i][j] = 6 + 16 j).
__global__ void k_mfma_f32_layout(float *__restrict__ out)
{
   const int l = threadIdx.x;
   const int i = l & 15, k = l >> 4;
   const f4v z = (f4v){0.f, 0.f, 0.f, 0.f};
   const f4v d1 = __builtin_amdgcn_mfma_f32_16x16x4f32((float)(i + 16 * k),
                                                       (k == (i & 3)) ? 1.f : 0.f, z, 0, 0, 0);
   const f4v d2 = __builtin_amdgcn_mfma_f32_16x16x4f32(1.f, (float)(k + 4 * i), z, 0, 0, 0);
#pragma unroll
   for (int t = 0; t < 4; ++t) {
      out[t * 64 + l] = d1[t];
      out[256 + t * 64 + l] = d2[t];
   }
}

// A table [K][ldU] (FP64) -> row operand [rows][ldk] FP32 and column operand
// [rows][ldk] split into bf16 hi/lo (Ch, Cl NULL: the FP32 copy only).  32x32
// tiles transposed through LDS.
__global__ void k_split_table(const double *__restrict__ U, long ldU, int K, int rows, long ldk,
                              float *__restrict__ UR, uint16_t *__restrict__ Ch,
                              uint16_t *__restrict__ Cl)
{
   __shared__ float tile[32][33];
   const int r0 = blockIdx.x * 32, k0 = blockIdx.y * 32;
   const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;     // 256 threads: 32 x 8
   for (int j = ty; j < 32; j += 8) {
      const int k = k0 + j, r = r0 + tx;
      tile[j][tx] = (k < K && r < rows) ? (float)U[(size_t)k * ldU + r] : 0.f;
   }
   __syncthreads();
   for (int j = ty; j < 32; j += 8) {
      const int r = r0 + j, k = k0 + tx;
      if (r < rows && k < ldk) {
         const float v = tile[tx][j];
         const size_t o = (size_t)r * ldk + k;
         UR[o] = v;
         if (Ch) {                 // bf16 planes (split-bf16x3 only)
            const __bf16 h = (__bf16)v;
            const __bf16 l = (__bf16)(v - (float)h);
            Ch[o] = __builtin_bit_cast(uint16_t, h);
            Cl[o] = __builtin_bit_cast(uint16_t, l);
         }
      }
   }
}

// FP64 -> FP32 copy of the A table
__global__ void k_to_f32(const double *__restrict__ src, float *__restrict__ dst, size_t n)
{
   size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
   const size_t st = (size_t)gridDim.x * blockDim.x;
   for (; i < n; i += st) dst[i] = (float)src[i];
}

// Lane-layout self test of v_mfma_f32_16x16x32_bf16 (one wave, 768 floats):
// three products with known answers, see bsahip_selftest_mfma_layout.
__global__ void k_mfma_bf16_layout(float *__restrict__ out)
{
   const int l = threadIdx.x;
   const int row = l & 15, g = l >> 4;
   uint32_t a1[8], b1[8], a2[8], b2[8], b3[8];
#pragma unroll
   for (int j = 0; j < 8; ++j) {
      const int k = 8 * g + j;                    // k index this element holds (assumed)
      a1[j] = bf16_rne_bits(k == 0 ? (float)row : 0.f);
      b1[j] = bf16_rne_bits(k == 0 ? 1.f : 0.f);
      a2[j] = bf16_rne_bits((float)k);
      b2[j] = bf16_rne_bits(k == row ? 1.f : 0.f);          // B[k][col=row-of-lane]
      b3[j] = bf16_rne_bits(k == row + 16 ? 1.f : 0.f);
   }
   auto pack = [](const uint32_t (&v)[8]) {
      uint4 q = make_uint4(v[0] | (v[1] << 16), v[2] | (v[3] << 16), v[4] | (v[5] << 16), v[6] | (v[7] << 16));
      return __builtin_bit_cast(bf8, q);
   };
   const f4v z = (f4v){0.f, 0.f, 0.f, 0.f};
   const f4v d1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(pack(a1), pack(b1), z, 0, 0, 0);
   const f4v d2 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(pack(a2), pack(b2), z, 0, 0, 0);
   const f4v d3 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(pack(a2), pack(b3), z, 0, 0, 0);
#pragma unroll
   for (int t = 0; t < 4; ++t) {
      out[t * 64 + l] = d1[t];
      out[256 + t * 64 + l] = d2[t];
      out[512 + t * 64 + l] = d3[t];
   }
}

// ---------------------------------------------------------------------------
// modal PSD and second moments (functionsImpl.F90:1546-1558, 1686-1689;
// BsaClassicImpl.F90:197-210, 567-589).
// psd(n,p;u) = sum_k U[k][n*nhp+u] Psi_a(p,k); the grid holds +-u.
// ---------------------------------------------------------------------------
// Modal PSD moments: PSD(n,p;u) = sum_k A(n,k;u) Psi_a(p,k) on the half grid,
// weighted with the vect rule (+u and -u), times Re[H_p conj(H_n)] and w^2.
// grid (nm*nm, ceil(nh/64)), block 256 = 64 frequencies x 4 slices of K;
// partial sums are added atomically (the outputs start at zero), or with
// m2part stored per block for a fixed-order sum (k_sum_fixed).
__global__ __launch_bounds__(256) void k_m2(const double *__restrict__ U, long ldU,
                                            const double *__restrict__ pa, int K, int nm,
                                            int nhp, int nh, const double2 *__restrict__ H,
                                            int nft, double d, double om_step,
                                            double *__restrict__ m2f, double *__restrict__ m2r,
                                            double *__restrict__ m2o, double *__restrict__ m2part)
{
   __shared__ double part[4][64];
   __shared__ double red[3][64];
   const int n = blockIdx.x % nm, p = blockIdx.x / nm;
   const int tu = threadIdx.x & 63, tk = threadIdx.x >> 6;
   const int u = blockIdx.y * 64 + tu;
   double s = 0.0;
   if (u < nh) {
      const double *col = U + (size_t)n * nhp + u;
      const double *pp = pa + (size_t)p * K;
      for (int k = tk; k < K; k += 4) s = fma(col[(size_t)k * ldU], pp[k], s);
   }
   part[tk][tu] = s;
   __syncthreads();
   if (tk == 0) {
      double sf = 0.0, sr = 0.0, so = 0.0;
      if (u < nh) {
         const double v = part[0][tu] + part[1][tu] + part[2][tu] + part[3][tu];
         double w = (u == nh - 1) ? 0.5 * d : d;
         if (u > 0) w *= 2.0;                                 // +u and -u
         const double2 hp = H[(size_t)p * nft + u], hn = H[(size_t)n * nft + u];
         const double fr = hp.x * hn.x + hp.y * hn.y;          // Re[H_p conj(H_n)]
         const double om = om_step * u;
         sf = w * v;
         sr = w * v * fr;
         so = w * v * fr * om * om;
      }
      red[0][tu] = wave_sum(sf);
      red[1][tu] = wave_sum(sr);
      red[2][tu] = wave_sum(so);
      if (tu == 0) {
         const int id = n + nm * p;
         if (m2part) {                 // fixed order: per-block partials [by][3][nm^2]
            const size_t nm2 = (size_t)nm * nm, b0 = (size_t)blockIdx.y * 3 * nm2;
            m2part[b0 + id] = red[0][0];
            m2part[b0 + nm2 + id] = red[1][0];
            m2part[b0 + 2 * nm2 + id] = red[2][0];
         } else {
            if (m2f) atomicAdd(m2f + id, red[0][0]);
            if (m2r) atomicAdd(m2r + id, red[1][0]);
            if (m2o) atomicAdd(m2o + id, red[2][0]);
         }
      }
   }
}

// Fixed-order compensated sums.  k_sum_lists: out[e] = sum over the slot list
// idx[off[e] .. off[e+1]) of vals (ascending slot order), Kahan-compensated;
// k_sum_fixed: out[i] = sum_b part[b * stride + i] for b = 0 .. nb-1, likewise.
__device__ __forceinline__ void kahan_add(double &s, double &c, double v)
{
   const double y = v - c;
   const double t = s + y;
   c = (t - s) - y;
   s = t;
}

__global__ void k_sum_lists(const double *__restrict__ vals, const long *__restrict__ off,
                            const long *__restrict__ idx, long n, double *__restrict__ out)
{
   const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
   if (e >= n) return;
   double s = 0.0, c = 0.0;
   for (long q = off[e]; q < off[e + 1]; ++q) kahan_add(s, c, vals[idx[q]]);
   out[e] = s;
}

__global__ void k_sum_fixed(const double *__restrict__ part, int nb, long stride, long n,
                            double *__restrict__ out)
{
   const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
   if (i >= n) return;
   double s = 0.0, c = 0.0;
   for (int b = 0; b < nb; ++b) kahan_add(s, c, part[(size_t)b * stride + i]);
   out[i] = s;
}

// Modal PSD on grid indices (for the spectra output): psd(q, n+p*nm).
__global__ void k_psd_grid(const double *__restrict__ U, long ldU, const double *__restrict__ pa,
                           int K, int nm, int nhp, int nf, int c, const double2 *__restrict__ H,
                           int nft, double *__restrict__ psd_f, double *__restrict__ psd_r)
{
   const int q = blockIdx.x * blockDim.x + threadIdx.x;
   const int np = blockIdx.y;
   if (q >= nf) return;
   const int n = np % nm, p = np / nm;
   const int u = abs(q - c);
   double s = 0.0;
   for (int k = 0; k < K; ++k) s = fma(U[(size_t)k * ldU + (size_t)n * nhp + u], pa[(size_t)p * K + k], s);
   const double2 hp = H[(size_t)p * nft + u], hn = H[(size_t)n * nft + u];
   if (psd_f) psd_f[(size_t)np * nf + q] = s;
   if (psd_r) psd_r[(size_t)np * nf + q] = s * (hp.x * hn.x + hp.y * hn.y);
}

// ---------------------------------------------------------------------------
// Modal force bispectrum of triple (m, n, p) at one point from the projection
// table U: rows a1 -> |f1|, a2 -> |f2|, a3 -> |f1+f2| (a3 < 0: S(f1+f2) = 0).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double bfm_point(const double *__restrict__ U, long ldU, int rows,
                                            const double *__restrict__ pb, int K, int m, int n,
                                            int p, int a1, int a2, int a3)
{
   double s = 0.0;
   for (int k = 0; k < K; ++k) {
      const double *row = U + (size_t)k * ldU;
      const double An1 = row[(size_t)n * rows + a1], Ap2 = row[(size_t)p * rows + a2];
      double v = pb[(size_t)m * K + k] * An1 * Ap2;
      if (a3 >= 0) {
         const double Am3 = row[(size_t)m * rows + a3];
         v += pb[(size_t)n * K + k] * Am3 * Ap2 + pb[(size_t)p * K + k] * Am3 * An1;
      }
      s += v;
   }
   return 2.0 * s;
}

// One G value G[x](y,u; z,v) = sum_k Psi_b(x,k) A(y,k;u) A(z,k;v) with the
// operand rounding of the mixed-precision G kernels: the operand of the lower
// table row (the kernels' row operand) is fl32(A) fl32(Psi_b) in FP32, the
// other fl32(A); PREC 1 (FP32) accumulates their FP32 product, PREC 2
// (split-bf16x3) the three bf16 products lo*hi + hi*lo + hi*hi, all in FP32
// and in K order (the MFMA path sums in another order; its error per term is
// the same).
template <int PREC>
__device__ __forceinline__ float g_value_mixed(const double *__restrict__ U, long ldU, int rows,
                                               const double *__restrict__ pb, int K, int x, int y,
                                               int u, int z, int v)
{
   if ((long)z * rows + v < (long)y * rows + u) {
      const int ty = y, tu = u;
      y = z;
      u = v;
      z = ty;
      v = tu;
   }
   float acc = 0.f;
   for (int k = 0; k < K; ++k) {
      const double *row = U + (size_t)k * ldU;
      const float a = (float)row[(size_t)y * rows + u] * (float)pb[(size_t)x * K + k];
      const float b = (float)row[(size_t)z * rows + v];
      if constexpr (PREC == 1) {
         acc = fmaf(a, b, acc);
      } else {
         const float ah = (float)(__bf16)a, bh = (float)(__bf16)b;
         const float al = (float)(__bf16)(a - ah), bl = (float)(__bf16)(b - bh);
         acc = fmaf(al, bh, acc);
         acc = fmaf(ah, bl, acc);
         acc = fmaf(ah, bh, acc);
      }
   }
   return acc;
}

// bfm_point with the G values of the mixed-precision kernels (g_value_mixed).
template <int PREC>
__device__ __forceinline__ double bfm_point_mixed(const double *__restrict__ U, long ldU, int rows,
                                                  const double *__restrict__ pb, int K, int m,
                                                  int n, int p, int a1, int a2, int a3)
{
   double s = (double)g_value_mixed<PREC>(U, ldU, rows, pb, K, m, n, a1, p, a2);
   if (a3 >= 0) {
      s += (double)g_value_mixed<PREC>(U, ldU, rows, pb, K, n, m, a3, p, a2);
      s += (double)g_value_mixed<PREC>(U, ldU, rows, pb, K, p, m, a3, n, a1);
   }
   return 2.0 * s;
}

// Response factor Re[H_m(f1) H_n(f2) conj H_p(f1+f2)] from the modal mass,
// stiffness and damping; H at f1+f2 uses w1 + w2 as the reference does
// (functionsImpl.F90:1659-1663).
__device__ __forceinline__ double brm_factor(double f1, double f2, int m, int n, int p,
                                             const double *__restrict__ Mm,
                                             const double *__restrict__ Km,
                                             const double *__restrict__ Cm)
{
   auto H = [&](int mo, double w) {
      const double r = -(w * w * Mm[mo]) + Km[mo];
      const double im = w * Cm[mo];
      const double d = r * r + im * im;
      return make_double2(r / d, -im / d);
   };
   const double w1 = 2.0 * M_PI * f1, w2 = 2.0 * M_PI * f2;
   return re3(H(m, w1), H(n, w2), H(p, w1 + w2));
}

// ---------------------------------------------------------------------------
// Per-point bispectrum: B(m,n,p) at points given by row indices into U
// (i1 -> f1, i2 -> f2, i3 -> f1+f2, i3 < 0 means S(f1+f2) = 0) and the
// signed frequencies for the FRF.  out index = pt*spt + mnp*smn.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_bfm_points(
   const double *__restrict__ U, long ldU, int rows, const double *__restrict__ pb, int K,
   int nm, int npts, const int *__restrict__ i1, const int *__restrict__ i2,
   const int *__restrict__ i3, const double *__restrict__ f1s, const double *__restrict__ f2s,
   const double *__restrict__ Mm, const double *__restrict__ Km, const double *__restrict__ Cm,
   double *__restrict__ bfm, double *__restrict__ brm, long spt, long smn)
{
   const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
   const long nm3 = (long)nm * nm * nm;
   if (gid >= nm3 * npts) return;
   const long pt = gid / nm3;
   const int mnp = (int)(gid % nm3);
   const int m = mnp % nm, n = (mnp / nm) % nm, p = mnp / (nm * nm);
   const double s = bfm_point(U, ldU, rows, pb, K, m, n, p, i1[pt], i2[pt], i3[pt]);
   if (bfm) bfm[pt * spt + mnp * smn] = s;
   if (brm) brm[pt * spt + mnp * smn] = s * brm_factor(f1s[pt], f2s[pt], m, n, p, Mm, Km, Cm);
}

// ---------------------------------------------------------------------------
// Bispectra of selected mode triples on the Classic grid f = (q - c) df,
// q = 0 .. nf-1: out[q1 + nf*q2 + nf*nf*s] for triple sel[s] (linear index
// m + n*nm + p*nm^2, m fastest), the reference's bisp(i, j, k) layout.
// Rows of U: |F| for F within the grid; with exact, |f1+f2| up to 2c as well,
// otherwise S(f1+f2) = 0 beyond the grid.  grid (ceil(nf*nf/256), nsel)
// ---------------------------------------------------------------------------
// PREC 0: FP64 (bfm_point); 1, 2: the G values of the FP32 / split-bf16x3
// kernels (bfm_point_mixed).
template <int PREC>
__global__ __launch_bounds__(256) void k_bisp_grid_sel(
   const double *__restrict__ U, long ldU, int rows, const double *__restrict__ pb, int K,
   int nm, int nf, int exact, double df, const int *__restrict__ sel,
   const double *__restrict__ Mm, const double *__restrict__ Km, const double *__restrict__ Cm,
   double *__restrict__ bfm, double *__restrict__ brm)
{
   const long npts = (long)nf * nf;
   const long pt = (long)blockIdx.x * blockDim.x + threadIdx.x;
   if (pt >= npts) return;
   const int c = (nf - 1) / 2;
   const int mnp = sel[blockIdx.y];
   const int m = mnp % nm, n = (mnp / nm) % nm, p = mnp / (nm * nm);
   const int F1 = (int)(pt % nf) - c, F2 = (int)(pt / nf) - c, F3 = F1 + F2;
   const int a3 = (exact || abs(F3) <= c) ? abs(F3) : -1;
   const double s = PREC == 0 ? bfm_point(U, ldU, rows, pb, K, m, n, p, abs(F1), abs(F2), a3)
                              : bfm_point_mixed<PREC>(U, ldU, rows, pb, K, m, n, p, abs(F1),
                                                      abs(F2), a3);
   const size_t o = (size_t)blockIdx.y * npts + pt;
   if (bfm) bfm[o] = s;
   if (brm) brm[o] = s * brm_factor(F1 * df, F2 * df, m, n, p, Mm, Km, Cm);
}

// Lane-layout self test of v_mfma_f64_16x16x4f64 (one wave, 512 doubles out):
// out[t*64 + l] = row index seen by register t of lane l, out[256 + t*64 + l]
// = its column index (A = row index in column 0, B = column index in row 0).
__global__ void k_mfma_f64_layout(double *__restrict__ out)
{
   const int l = threadIdx.x;
   const bool k0 = (l >> 4) == 0;
   d4 z = (d4){0.0, 0.0, 0.0, 0.0};
   d4 rows = __builtin_amdgcn_mfma_f64_16x16x4f64(k0 ? (double)(l & 15) : 0.0, k0 ? 1.0 : 0.0, z, 0, 0, 0);
   d4 cols = __builtin_amdgcn_mfma_f64_16x16x4f64(k0 ? 1.0 : 0.0, k0 ? (double)(l & 15) : 0.0, z, 0, 0, 0);
#pragma unroll
   for (int t = 0; t < 4; ++t) {
      out[t * 64 + l] = rows[t];
      out[256 + t * 64 + l] = cols[t];
   }
}

// ---------------------------------------------------------------------------
// Sum over a point list of w_pt * B(m,n,p)(pt) [* FRF]: one thread per mode
// triple and chunk of points, one atomic per thread.  Used for point sets that
// are not a symmetric uniform grid (BsaCL computation frequencies).
// grid (ceil(nm3/256), nchunk).
// ---------------------------------------------------------------------------
// FRF weighting of given modal spectra (getRM_full_vect_cls_,
// functionsImpl.F90:1598-1738, and getRM_full_scalar_msh_, :1031-1115), in
// place, H_m(f) = 1 / (K_m - w^2 M_m + i w C_mm), w = 2 pi f:
//   kind 0: psd(n1, NM^2), column n + p NM: times Re[H_p(f1_i) conj H_n(f1_i)]
//   kind 1: bisp(n1, n2, NM^3), entry m + n NM + p NM^2 at (f1_i, f2_j):
//           times Re[H_m(f1_i) H_n(f2_j) conj H_p(f1_i + f2_j)]
//   kind 2: bisp(NM^3, n1) at the points (f1_q, f2_q): the same factor.
// grid (ceil(total/256)), block 256
// ---------------------------------------------------------------------------
// U[k*ldU + y*nu + u] = 0 for u >= u0 (n = K * nm * (nu - u0) values)
__global__ void k_zero_cols(double *__restrict__ U, long ldU, int nu, long u0, int nm, long n)
{
   const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
   if (t >= n) return;
   const long w = nu - u0;
   const long u = u0 + t % w, y = (t / w) % nm, k = t / (w * nm);
   U[k * ldU + y * nu + u] = 0.0;
}

__device__ __forceinline__ double2 frf_at(double f, const double *Mm, const double *Km,
                                          const double *Cm, int m)
{
   const double w = 2.0 * M_PI * f;
   const double r = Km[m] - w * w * Mm[m], im = w * Cm[m], d = r * r + im * im;
   return make_double2(r / d, -im / d);
}

// FRF weighting in place (bsahip_frf_apply).  kind 0: (n1, E2) spectra at f1,
// column c = n + p NM times Re[H_p conj H_n] (diag: column m times |H_m|^2);
// kind 1: (n1, n2, E3) at (f1_i, f2_j); kind 2: (E3, n1) at (f1_q, f2_q); entry
// e = m + n NM + p NM^2 times Re[H_m(a) H_n(b) conj H_p(a+b)] (diag: e = m, all
// three modes m).  E2 = NM^2, E3 = NM^3, or NM both for the diagonal model.
__global__ void k_frf_apply(int kind, int diag, long n1, const double *__restrict__ f1, long n2,
                            const double *__restrict__ f2, int nm, const double *__restrict__ Mm,
                            const double *__restrict__ Km, const double *__restrict__ Cm,
                            long total, double *__restrict__ data)
{
   const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
   if (t >= total) return;
   if (kind == 0) {
      const long i = t % n1, c = t / n1;
      const int n = diag ? (int)c : (int)(c % nm), p = diag ? (int)c : (int)(c / nm);
      const double2 hn = frf_at(f1[i], Mm, Km, Cm, n), hp = frf_at(f1[i], Mm, Km, Cm, p);
      data[t] *= hp.x * hn.x + hp.y * hn.y;
      return;
   }
   const long ne = diag ? (long)nm : (long)nm * nm * nm;
   double a, b;
   long e;
   if (kind == 1) {
      const long i = t % n1, j = (t / n1) % n2;
      e = t / (n1 * n2);
      a = f1[i];
      b = f2[j];
   } else {
      const long q = t / ne;
      e = t % ne;
      a = f1[q];
      b = f2[q];
   }
   int m, n, p;
   if (diag) {
      m = n = p = (int)e;
   } else {
      m = (int)(e % nm);
      n = (int)((e / nm) % nm);
      p = (int)(e / ((long)nm * nm));
   }
   data[t] *= re3(frf_at(a, Mm, Km, Cm, m), frf_at(b, Mm, Km, Cm, n), frf_at(a + b, Mm, Km, Cm, p));
}

// Diagonal-model modal PSD at points: psd[q nm + m] = sum_k pa(m,k)^2 S_k(f_q),
// S from the sqrt-PSD table sq (ntc, nu, nnl) at row iu[q] (getFM_diag_tnlm_*'s
// psd, functionsImpl.F90:1747-1970).
__global__ void k_diag_psd_points(const double *__restrict__ sq, int nu, int nnl, int kv,
                                  const double *__restrict__ pa, int K, int nm, int npts,
                                  const int *__restrict__ iu, double *__restrict__ psd)
{
   const long gid = (long)blockIdx.x * blockDim.x + threadIdx.x;
   if (gid >= (long)npts * nm) return;
   const int q = (int)(gid / nm), m = (int)(gid % nm);
   const int u = iu[q];
   double s = 0.0;
   for (int k = 0; k < kv; ++k) {                      // rows k >= kv are padding
      const int c = k / nnl, n = k % nnl;
      const double r = sq[((size_t)c * nu + u) * nnl + n];
      const double a = pa[(size_t)m * K + k];
      s += a * a * r * r;
   }
   psd[gid] = s;
}

// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_bfm_points_sum(
   const double *__restrict__ U, long ldU, int rows, const double *__restrict__ pb, int K,
   int nm, int npts, const int *__restrict__ i1, const int *__restrict__ i2,
   const int *__restrict__ i3, const double *__restrict__ f1s, const double *__restrict__ f2s,
   const double *__restrict__ wpt, const double *__restrict__ Mm, const double *__restrict__ Km,
   const double *__restrict__ Cm, int with_frf, double *__restrict__ out)
{
   const int nm3 = nm * nm * nm;
   const int mnp = blockIdx.x * blockDim.x + threadIdx.x;
   if (mnp >= nm3) return;
   const int m = mnp % nm, n = (mnp / nm) % nm, p = mnp / (nm * nm);
   const int per = (npts + gridDim.y - 1) / gridDim.y;
   const int p0 = blockIdx.y * per, p1 = min(npts, p0 + per);
   double acc = 0.0;
   for (int pt = p0; pt < p1; ++pt) {
      const int a1 = i1[pt], a2 = i2[pt], a3 = i3[pt];
      double s = 0.0;
      for (int k = 0; k < K; ++k) {
         const double *row = U + (size_t)k * ldU;
         const double An1 = row[(size_t)n * rows + a1], Ap2 = row[(size_t)p * rows + a2];
         double v = pb[(size_t)m * K + k] * An1 * Ap2;
         if (a3 >= 0) {
            const double Am3 = row[(size_t)m * rows + a3];
            v += pb[(size_t)n * K + k] * Am3 * Ap2 + pb[(size_t)p * K + k] * Am3 * An1;
         }
         s += v;
      }
      s *= 2.0 * wpt[pt];
      if (with_frf) {
         const double w1 = 2.0 * M_PI * f1s[pt], w2 = 2.0 * M_PI * f2s[pt], w3 = w1 + w2;
         double r, im, d;
         r = -(w1 * w1 * Mm[m]) + Km[m]; im = w1 * Cm[m]; d = r * r + im * im;
         const double2 h1 = make_double2(r / d, -im / d);
         r = -(w2 * w2 * Mm[n]) + Km[n]; im = w2 * Cm[n]; d = r * r + im * im;
         const double2 h2 = make_double2(r / d, -im / d);
         r = -(w3 * w3 * Mm[p]) + Km[p]; im = w3 * Cm[p]; d = r * r + im * im;
         const double2 h3 = make_double2(r / d, -im / d);
         s *= re3(h1, h2, h3);
      }
      acc += s;
   }
   atomicAdd(out + mnp, acc);
}
