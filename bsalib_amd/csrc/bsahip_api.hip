// C ABI of libbsahip.so (declarations and layouts: include/bsahip.h).
//
// Host side of the MI355X BFM path.  A plan holds one structure resident on
// the device (Psi per K index, ln rho, PSD parameters, modal data) and every
// buffer its Classic run needs; bsahip_plan_run is the per-step hot path
// (nodal PSD -> projection A(f) -> [bf16 split] -> G contraction with the
// plane epilogue -> modal PSD moments -> copy of the moments).  There is no
// CPU fallback: without a usable GPU every entry point returns BSAHIP_EDEVICE.
#include "kernels.hip"
#include "gemm.hip"
#include "pod.hip"
#include "diag.hip"
#include "../../include/bsahip.h"
#include "bsahip_internal.h"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
   g_err = msg;
   return code;
}

#define HIPC(x)                                                                              \
   do {                                                                                      \
      hipError_t e_ = (x);                                                                   \
      if (e_ != hipSuccess)                                                                  \
         return fail(BSAHIP_EDEVICE, std::string(#x) + ": " + hipGetErrorString(e_));       \
   } while (0)

#define CHK(x)                                                                               \
   do {                                                                                      \
      int r_ = (x);                                                                          \
      if (r_ != 0) return r_;                                                                \
   } while (0)

// Named, grow-only device buffers.  With an arena (set_arena), get() carves
// the buffers out of that fixed device region instead and never allocates.
struct Bufs {
   struct Buf {
      void *p = nullptr;
      size_t n = 0;
   };
   std::map<std::string, Buf> m;
   char *arena = nullptr;
   size_t arena_cap = 0, arena_used = 0;
   ~Bufs() { release(); }
   void release()
   {
      if (!arena)
         for (auto &kv : m)
            if (kv.second.p) (void)hipFree(kv.second.p);
      m.clear();
      arena_used = 0;
   }
   void set_arena(void *base, size_t cap)
   {
      release();
      arena = static_cast<char *>(base);
      arena_cap = cap;
   }
   template <class T>
   int get(const char *name, size_t count, T **out)
   {
      Buf &b = m[name];
      const size_t bytes = std::max<size_t>(count * sizeof(T), 256);
      if (arena) {
         if (b.n < bytes) {
            const size_t need = (bytes + 255) / 256 * 256;
            if (arena_used + need > arena_cap)
               return fail(BSAHIP_EINVAL,
                           std::string("the problem needs more than the ") +
                              std::to_string(arena_cap >> 20) +
                              " MiB of static device scratch of this entry point (at buffer " + name +
                              "); use the allocating entry point");
            b.p = arena + arena_used;
            b.n = need;
            arena_used += need;
         }
         *out = static_cast<T *>(b.p);
         return 0;
      }
      if (b.n < bytes) {
         // work queued earlier may still read the old buffer
         if (b.p) {
            (void)hipDeviceSynchronize();
            (void)hipFree(b.p);
         }
         b.p = nullptr;
         b.n = 0;
         hipError_t e = hipMalloc(&b.p, bytes);
         if (e != hipSuccess)
            return fail(BSAHIP_ENOMEM, std::string("hipMalloc(") + name + ", " +
                                           std::to_string(bytes) + " B): " + hipGetErrorString(e));
         b.n = bytes;
      }
      *out = static_cast<T *>(b.p);
      return 0;
   }
};

// Static device scratch of bsahip_classic_run_noalloc: every buffer of its
// run (problem tables, A table, split planes, FRF table, moment accumulators)
// is carved out of this array, so that entry point never allocates.  Each
// device holds its own copy, at the address hipGetSymbolAddress returns with
// that device current (kept in Ctx::scratch_dev); the name itself is a host
// stand-in and is never passed to a kernel or a copy.
#define BSAHIP_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
#define BSAHIP_MAX_DEVICES 64
static __device__ __align__(256) unsigned char scratch[BSAHIP_NOALLOC_SCRATCH_BYTES];

struct Ctx {
   bool ready = false;
   int dev = 0;
   void *scratch_dev[BSAHIP_MAX_DEVICES] = {};   // device address of `scratch` per device
   hipStream_t s = nullptr;
   hipStream_t s2 = nullptr;             // side stream: modal PSD moments beside the G kernel
   hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr, e3 = nullptr;
   Bufs *bufs = nullptr;                 // scratch of the one-shot entry points
   double last_ms = 0.0;
   int last_launches = 0;
   rocblas_handle blas = nullptr;        // created on first use (truncated POD)
   std::recursive_mutex mu;
};

Ctx g;

template <class T>
int upload(Bufs &B, const char *name, const std::vector<T> &h, T **out)
{
   CHK(B.get(name, h.size(), out));
   if (!h.empty())
      HIPC(hipMemcpyAsync(*out, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, g.s));
   return 0;
}

int fill(double *p, size_t n, double v)
{
   if (n == 0) return 0;
   const int blocks = (int)std::min<size_t>((n + 255) / 256, 65536);
   k_fill<<<blocks, 256, 0, g.s>>>(p, n, v);
   HIPC(hipGetLastError());
   return 0;
}

int ensure_init()
{
   if (g.ready) return 0;
   return bsahip_init(0) == 0 ? 0 : BSAHIP_EDEVICE;
}

inline int round_up(int a, int b) { return (a + b - 1) / b * b; }

int validate(const bsahip_problem *p)
{
   if (!p) return fail(BSAHIP_EINVAL, "problem pointer is NULL");
   if (p->nn <= 0 || p->nnl <= 0 || p->nm <= 0 || p->ntc <= 0 || p->ntc > 3 || p->nwz <= 0)
      return fail(BSAHIP_EINVAL, "invalid sizes (nn, nnl, nm, ntc in 1..3, nwz must be > 0)");
   if (p->nnl > p->nn) return fail(BSAHIP_EINVAL, "nnl > nn");
   if (p->nm > 256) return fail(BSAHIP_EUNSUPPORTED, "more than 256 modes");
   if (p->ndegw < 6) return fail(BSAHIP_EINVAL, "ndegw must be >= 6 (3 linear + 3 quadratic)");
   if ((long)p->ncorr_rows != ((long)p->nn * p->nn + p->nn) / 2)
      return fail(BSAHIP_EINVAL, "ncorr_rows must equal (nn*nn+nn)/2");
   if (!p->n_load || !p->tc || !p->psi || !p->nod_corr || !p->u_node || !p->wz_node ||
       !p->turb_scales || !p->turb_std || !p->Mm || !p->Km || !p->Cm)
      return fail(BSAHIP_EINVAL, "a required problem array is NULL");
   if (p->psd_id == 3 && !p->u_mean_ref) return fail(BSAHIP_EINVAL, "psd_id 3 needs u_mean_ref");
   if (!(p->psd_id == 1 || p->psd_id == 2 || p->psd_id == 3 || p->psd_id == 5))
      return fail(BSAHIP_EINVAL, "psd_id must be 1, 2, 3 or 5");
   for (int a = 0; a < p->nnl; ++a)
      if (p->n_load[a] < 1 || p->n_load[a] > p->nn) return fail(BSAHIP_EINVAL, "n_load out of range");
   for (int c = 0; c < p->ntc; ++c)
      if (p->tc[c] < 1 || p->tc[c] > 3) return fail(BSAHIP_EINVAL, "tc must be in 1..3");
   for (int a = 0; a < p->nn; ++a)
      if (p->wz_node[a] < 1 || p->wz_node[a] > p->nwz) return fail(BSAHIP_EINVAL, "wz_node out of range");
   return 0;
}

// Device-resident description of one structure.
struct Prob {
   // K rows: component c, loaded node i at k = c*nnlp + i (nnlp = nnl), the
   // kv = ntc*nnl valid rows padded with zero rows to a multiple of the K chunk
   int nn = 0, nnl = 0, nnlp = 0, nm = 0, ntc = 0, K = 0, kv = 0, pulsation = 1;
   double *pa_t = nullptr;   // [tc][m][nnl]
   double *pa_k = nullptr;   // [m][K]
   double *pb_k = nullptr;   // [m][K]
   float *pb32 = nullptr;    // [m][K]
   double *lnr = nullptr;    // [tc][nnl(j)][nnl(i)]
   double *prm = nullptr;    // [tc][nnl][4]
   int *formula = nullptr;   // [tc]
   double *Mm = nullptr, *Km = nullptr, *Cm = nullptr;
   // truncated POD of S(f) in the point tables (0/0: none, direct projection)
   int pod_nmodes = 0;
   double pod_lim = 0.0;
};

// psd_mode 0: reference Classic/Mesher semantics (evalPSD per component);
// psd_mode 1: BsaCL semantics (u formula for every component, zone 1, rho clamped).
int load_problem(const bsahip_problem *p, int psd_mode, Bufs &B, Prob &D)
{
   CHK(validate(p));
   D.nn = p->nn;
   D.nnl = p->nnl;
   D.nm = p->nm;
   D.ntc = p->ntc;
   D.nnlp = p->nnl;
   D.kv = D.ntc * D.nnl;
   D.K = round_up(D.kv, KC);
   D.pulsation = p->pulsation;
   const int nm = D.nm, nnl = D.nnl, K = D.K;
   std::vector<double> pa_t((size_t)D.ntc * nm * nnl, 0.0), pa_k((size_t)nm * K, 0.0),
      pb_k((size_t)nm * K, 0.0);
   for (int c = 0; c < D.ntc; ++c) {
      const int t = p->tc[c];
      for (int i = 0; i < nnl; ++i)
         for (int m = 0; m < nm; ++m) {
            const double a = p->psi[m + (size_t)nm * (i + (size_t)nnl * (t - 1))];
            const double b = p->psi[m + (size_t)nm * (i + (size_t)nnl * (t + 2))];
            pa_t[((size_t)c * nm + m) * nnl + i] = a;
            pa_k[(size_t)m * K + (size_t)c * D.nnlp + i] = a;
            pb_k[(size_t)m * K + (size_t)c * D.nnlp + i] = b;
         }
   }
   std::vector<float> pb32(pb_k.begin(), pb_k.end());
   // coherence columns actually read (nod_corr_(:, tc), or column 1 for all)
   std::vector<int> col(D.ntc);
   for (int c = 0; c < D.ntc; ++c)
      col[c] = (p->corr_col_mode == 1 || psd_mode == 1) ? 0 : p->tc[c] - 1;
   std::vector<int> ucol = col;
   std::sort(ucol.begin(), ucol.end());
   ucol.erase(std::unique(ucol.begin(), ucol.end()), ucol.end());
   std::vector<int> colmap(D.ntc);
   for (int c = 0; c < D.ntc; ++c)
      colmap[c] = (int)(std::lower_bound(ucol.begin(), ucol.end(), col[c]) - ucol.begin());
   std::vector<double> corr((size_t)ucol.size() * p->ncorr_rows);
   for (size_t k = 0; k < ucol.size(); ++k)
      std::memcpy(corr.data() + k * p->ncorr_rows, p->nod_corr + (size_t)ucol[k] * p->ncorr_rows,
                  sizeof(double) * p->ncorr_rows);
   std::vector<int> nl(p->n_load, p->n_load + nnl);
   std::vector<double> prm((size_t)D.ntc * nnl * 4, 0.0);
   std::vector<int> formula(D.ntc, 0);
   for (int c = 0; c < D.ntc; ++c) {
      const int t = p->tc[c];
      int f = 0;
      if (psd_mode == 1) f = (p->psd_id == 5) ? 5 : (p->psd_id == 1 ? 1 : 0);
      else if (p->psd_id == 1) f = (t == 1) ? 1 : 2;
      else f = (p->psd_id == 2) ? 0 : p->psd_id;
      formula[c] = f;
      for (int i = 0; i < nnl; ++i) {
         const int node = p->n_load[i] - 1;
         const int wz = (psd_mode == 1) ? 0 : p->wz_node[node] - 1;
         // L(component, direction 1, zone): von Karman u reads component 1
         // (WindPSDImpl.F90:167), the other formulas component tc.
         const int lc = (psd_mode == 0 && f == 1) ? 1 : t;
         double *q = prm.data() + ((size_t)c * nnl + i) * 4;
         q[0] = p->turb_scales[(lc - 1) + 9 * (size_t)wz];
         q[1] = p->turb_std[(t - 1) + 3 * (size_t)wz];
         q[2] = p->u_node[node];
         q[3] = p->u_mean_ref ? p->u_mean_ref[wz] : 1.0;
      }
   }
   std::vector<double> Mm(p->Mm, p->Mm + nm), Km(p->Km, p->Km + nm), Cm(p->Cm, p->Cm + nm);
   double *dcorr;
   int *dcol, *dnl;
   CHK(upload(B, "pa_t", pa_t, &D.pa_t));
   CHK(upload(B, "pa_k", pa_k, &D.pa_k));
   CHK(upload(B, "pb_k", pb_k, &D.pb_k));
   CHK(upload(B, "pb32", pb32, &D.pb32));
   CHK(upload(B, "corr", corr, &dcorr));
   CHK(upload(B, "colmap", colmap, &dcol));
   CHK(upload(B, "n_load", nl, &dnl));
   CHK(upload(B, "prm", prm, &D.prm));
   CHK(upload(B, "formula", formula, &D.formula));
   CHK(upload(B, "Mm", Mm, &D.Mm));
   CHK(upload(B, "Km", Km, &D.Km));
   CHK(upload(B, "Cm", Cm, &D.Cm));
   CHK(B.get("lnr", (size_t)D.ntc * nnl * nnl, &D.lnr));
   dim3 gl((nnl + 255) / 256, nnl, D.ntc);
   k_lnrho<<<gl, 256, 0, g.s>>>(dcorr, p->ncorr_rows, dcol, dnl, nnl, p->nn,
                                psd_mode == 1 ? 2.2250738585072014e-308 : 0.0, D.lnr);
   HIPC(hipGetLastError());
   return 0;
}

// Truncated POD for the point paths (pod_nmodes / pod_trunc_lim of the problem).
int set_pod(const bsahip_problem *p, Prob &D)
{
   D.pod_nmodes = 0;
   D.pod_lim = 0.0;
   if (p->pod_nmodes < 0 || !std::isfinite(p->pod_trunc_lim) || p->pod_trunc_lim < 0.0)
      return fail(BSAHIP_EINVAL, "pod_nmodes must be >= 0 and pod_trunc_lim finite and >= 0");
   if (p->pod_nmodes == 0 && p->pod_trunc_lim == 0.0) return 0;
   if (p->corr_col_mode != 1)
      return fail(BSAHIP_EINVAL, "a truncated POD reads coherence column 1 (corr_col_mode must be 1)");
   D.pod_nmodes = p->pod_nmodes;
   D.pod_lim = p->pod_trunc_lim;
   return 0;
}

// A table on a list of |f| values (device array fl): U[K][ldU], rows y*rows_per_mode + q.
#define BLASC(x)                                                                             \
   do {                                                                                      \
      rocblas_status st_ = (x);                                                              \
      if (st_ != rocblas_status_success)                                                     \
         return fail(BSAHIP_EDEVICE, std::string(#x) + ": " + rocblas_status_to_string(st_)); \
   } while (0)

// rocBLAS / rocSOLVER handle on the library stream, created on first use.
int blas_handle()
{
   if (!g.blas) {
      BLASC(rocblas_create_handle(&g.blas));
      BLASC(rocblas_set_stream(g.blas, g.s));
   }
   return 0;
}

// C_b = alpha op(A_b) op(B_b) + beta C_b on the library stream (k_dgemm_sb).
int dgemm_sb(int transa, int transb, int M, int N, int K, double alpha, const double *A, long lda,
             long sA, const double *B, long ldb, long sB, double beta, double *C, long ldc,
             long sC, int batch)
{
   if (M <= 0 || N <= 0 || batch <= 0) return 0;
   if (batch > 65535) return fail(BSAHIP_EUNSUPPORTED, "GEMM batch above 65535");
   dim3 gr((unsigned)((M + 63) / 64), (unsigned)((N + 63) / 64), (unsigned)batch);
   k_dgemm_sb<<<gr, 256, 0, g.s>>>(transa, transb, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C,
                                    ldc, sC);
   HIPC(hipGetLastError());
   return 0;
}

// A_K(y, i; f) = (Psi_a V) diag(w) V^T for the truncated POD of S(f) (see
// pod.hip), written into the table layout of k_project.  sq: sqrt(S_i) at fl.
int pod_project(Bufs &B, const Prob &D, const double *sq, const double *fl, int nfl,
                int rows_per_mode, double *U, long ldU)
{
   CHK(blas_handle());
   const int n = D.nnl, nm = D.nm;
   const size_t nn2 = (size_t)n * n, mn = (size_t)nm * n;
   const long total = (long)D.ntc * nfl;
   // at most 2^28 doubles (2 GiB) of matrices and 65535 matrices per chunk
   const long cap = std::max<long>(1, std::min<long>({total, 65535L, (1L << 28) / (long)nn2}));
   double *S, *lam, *E, *w, *T, *A;
   int *info, *kept;
   CHK(B.get("pod_S", (size_t)cap * nn2, &S));
   CHK(B.get("pod_lam", (size_t)cap * n, &lam));
   CHK(B.get("pod_E", (size_t)cap * n, &E));
   CHK(B.get("pod_w", (size_t)cap * n, &w));
   CHK(B.get("pod_T", (size_t)cap * mn, &T));
   CHK(B.get("pod_A", (size_t)cap * mn, &A));
   CHK(B.get("pod_info", (size_t)cap, &info));
   CHK(B.get("pod_kept", (size_t)cap, &kept));
   std::vector<int> hk((size_t)cap);
   for (long b0 = 0; b0 < total; b0 += cap) {
      const int nb = (int)std::min<long>(cap, total - b0);
      k_pod_smat<<<dim3((unsigned)((nn2 + 255) / 256), nb), 256, 0, g.s>>>(D.lnr, sq, fl, nfl, n, b0, S);
      HIPC(hipGetLastError());
      BLASC(rocsolver_dsyevd_strided_batched(g.blas, rocblas_evect_original, rocblas_fill_lower, n, S,
                                             n, (rocblas_stride)nn2, lam, n, E, n, info, nb));
      k_pod_weights<<<nb, 64, 0, g.s>>>(lam, info, n, D.pod_nmodes, D.pod_lim, w, kept);
      HIPC(hipGetLastError());
      // T = Psi_a(c) V: Psi_a(c) is pa_t[c] read as the column-major n x nm
      // matrix P, so T = P^T V; one strided batch per component in the chunk
      for (long bb = b0; bb < b0 + nb;) {
         const int c = (int)(bb / nfl);
         const long be = std::min<long>(b0 + nb, (long)(c + 1) * nfl);
         CHK(dgemm_sb(1, 0, nm, n, n, 1.0, D.pa_t + (size_t)c * mn, n, 0, S + (size_t)(bb - b0) * nn2,
                      n, (long)nn2, 0.0, T + (size_t)(bb - b0) * mn, nm, (long)mn, (int)(be - bb)));
         bb = be;
      }
      k_pod_scale<<<dim3((unsigned)((mn + 255) / 256), nb), 256, 0, g.s>>>(T, w, nm, n);
      HIPC(hipGetLastError());
      CHK(dgemm_sb(0, 1, nm, n, n, 1.0, T, nm, (long)mn, S, n, (long)nn2, 0.0, A, nm, (long)mn, nb));
      k_pod_scatter<<<dim3((unsigned)((mn + 255) / 256), nb), 256, 0, g.s>>>(A, nm, n, nfl, b0, D.nnlp,
                                                                         rows_per_mode, U, ldU);
      HIPC(hipGetLastError());
      HIPC(hipMemcpyAsync(hk.data(), kept, (size_t)nb * sizeof(int), hipMemcpyDeviceToHost, g.s));
      HIPC(hipStreamSynchronize(g.s));
      for (int b = 0; b < nb; ++b)
         if (hk[b] < 0) return fail(BSAHIP_EDEVICE, "POD eigendecomposition of S(f) did not converge");
   }
   return 0;
}

int build_table(Bufs &B, const Prob &D, const double *fl, int nfl, int rows_per_mode, double *U,
                long ldU)
{
   if (nfl > 65535) return fail(BSAHIP_EUNSUPPORTED, "more than 65535 distinct frequencies in one table");
   double *sq;
   CHK(B.get("sq", (size_t)D.ntc * nfl * D.nnl, &sq));
   CHK(fill(U, (size_t)D.K * ldU, 0.0));
   dim3 gp((D.nnl + 63) / 64, nfl, D.ntc);
   k_psd_sqrt<<<gp, 64, 0, g.s>>>(fl, nfl, D.prm, D.formula, D.nnl, D.ntc, D.pulsation, sq);
   HIPC(hipGetLastError());
   if (D.pod_nmodes > 0 || D.pod_lim > 0.0)
      return pod_project(B, D, sq, fl, nfl, rows_per_mode, U, ldU);
   // modes in chunks of at most 64 (grid z = component + NTC * chunk)
   const int mt = (D.nm + 15) / 16;
   const int nchunk = (D.nm + 63) / 64;
   dim3 gpm(gp.x, gp.y, (unsigned)(D.ntc * nchunk));
#define LAUNCH_PROJ(MT)                                                                       \
   k_project<MT><<<gpm, 256, 0, g.s>>>(D.pa_t, D.lnr, sq, fl, nfl, D.nnl, D.nm, D.nnlp,       \
                                       rows_per_mode, U, ldU, D.ntc)
   if (mt <= 1) LAUNCH_PROJ(1);
   else if (mt <= 2) LAUNCH_PROJ(2);
   else if (mt <= 3) LAUNCH_PROJ(3);
   else LAUNCH_PROJ(4);
#undef LAUNCH_PROJ
   HIPC(hipGetLastError());
   return 0;
}

int launch_frf(const Prob &D, int nft, double df, double2 *H)
{
   dim3 gh((nft + 255) / 256, D.nm);
   k_frf_table<<<gh, 256, 0, g.s>>>(D.Mm, D.Km, D.Cm, D.nm, nft, df, H);
   HIPC(hipGetLastError());
   return 0;
}

int check_flags(int32_t flags, bool grid, bool allow_mixed)
{
   const int prec = flags & BSAHIP_PREC_MASK;
   const int quad = (flags & BSAHIP_QUAD_MASK) >> 4;
   if (prec != BSAHIP_PREC_FP64 &&
       !(allow_mixed && (prec == BSAHIP_PREC_BF16X3 || prec == BSAHIP_PREC_FP32)))
      return fail(BSAHIP_EUNSUPPORTED, "precision not available in this entry point");
   if (grid && quad > 2) return fail(BSAHIP_EINVAL, "unknown quadrature rule");
   return 0;
}

// Distinct |f| over f1, f2, f1+f2 of a point list and the index maps into it.
void point_tables(int npts, const double *fi, const double *fj, std::vector<double> &uq,
                  std::vector<int> &i1, std::vector<int> &i2, std::vector<int> &i3)
{
   std::vector<double> all;
   all.reserve(3 * (size_t)npts);
   for (int q = 0; q < npts; ++q) {
      all.push_back(std::fabs(fi[q]));
      all.push_back(std::fabs(fj[q]));
      all.push_back(std::fabs(fi[q] + fj[q]));
   }
   uq = all;
   std::sort(uq.begin(), uq.end());
   uq.erase(std::unique(uq.begin(), uq.end()), uq.end());
   auto idx = [&](double v) { return (int)(std::lower_bound(uq.begin(), uq.end(), v) - uq.begin()); };
   i1.resize(npts);
   i2.resize(npts);
   i3.resize(npts);
   for (int q = 0; q < npts; ++q) {
      i1[q] = idx(all[3 * q]);
      i2[q] = idx(all[3 * q + 1]);
      i3[q] = idx(all[3 * q + 2]);
   }
}

}  // namespace

// A Classic run of one structure on one grid, resident on the device.
struct bsahip_plan {
   Bufs B;
   Prob D;
   int32_t flags = 0, prec = 0;
   int nh = 0, nf = 0, nu = 0, nhp = 0, rows = 0, ntile = 0, nft = 0;
   long ldU = 0;
   double df = 0.0, d = 0.0;
   double *fl = nullptr, *U = nullptr, *m3 = nullptr, *m2 = nullptr;
   double2 *H = nullptr;
   float *UR = nullptr;
   uint16_t *Ch = nullptr, *Cl = nullptr;
   double last_ms = 0.0;
   // fixed-order summation (BSAHIP_DETERMINISTIC): per-flush slots of the G
   // epilogue, the slot list of every m3 entry, per-block m2 partials
   double *slots = nullptr, *m2part = nullptr;
   long *det_off = nullptr, *det_idx = nullptr;
   int mflush = 0;
};

namespace {

// Host pair index of a tile pair as the kernels' pair_from_index: t ->
// (R, Q), R <= Q, t = Q (Q + 1) / 2 + R.
void host_pair(long t, int &R, int &Q)
{
   long q = (long)((std::sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
   while ((q + 1) * (q + 2) / 2 <= t) ++q;
   while (q * (q + 1) / 2 > t) --q;
   Q = (int)q;
   R = (int)(t - q * (q + 1) / 2);
}

// Slot lists of the fixed-order mode: the flushes of every wave of every G
// item in the order epi_wave makes them (kernels.hip), each with the twelve m3
// entries of epi_flush (m3f entries 0 .. NM^3-1, m3r entries NM^3 .. 2 NM^3-1).
// off (2 NM^3 + 1) and idx list, per entry, its slots in ascending order;
// mflush is the largest flush count of a wave.
void det_slot_lists(const bsahip_plan *P, std::vector<long> &off, std::vector<long> &idx,
                    int &mflush)
{
   const int nm = P->D.nm, nhp = P->nhp;
   const bool f64 = (P->prec == BSAHIP_PREC_FP64);
   const int T = f64 ? TILE : T2, WT = f64 ? 32 : 64, BR = WT / 16, BC = BR;
   const long nm3 = (long)nm * nm * nm, nent = 2 * nm3;
   const long npairs = (long)P->ntile * (P->ntile + 1) / 2, nitems = npairs * nm;
   auto walk = [&](auto emit) {
      for (long lid = 0; lid < nitems; ++lid) {
         const int x = (int)(lid % nm);
         int R, Q;
         host_pair(lid / nm, R, Q);
         const bool diag = (R == Q);
         for (int w = 0; w < 4; ++w) {
            const int rw0 = R * T + (w >> 1) * WT, cw0 = Q * T + (w & 1) * WT;
            const int y0 = rw0 / nhp, y1 = (rw0 + 16 * BR - 1) / nhp;
            const int z0 = cw0 / nhp, z1 = (cw0 + 16 * BC - 1) / nhp;
            int f = 0;
            if (y0 == y1 && z0 == z1) {
               if (y0 < nm && z0 < nm) emit(lid, w, f++, x, y0, z0, diag);
            } else {
               for (int y = y0; y <= y1; ++y)
                  for (int z = z0; z <= z1; ++z)
                     if (y < nm && z < nm) emit(lid, w, f++, x, y, z, diag);
            }
         }
      }
   };
   auto entries = [&](int x, int y, int z, bool diag, long (&e)[12]) {
      const long xyz = x + (long)nm * y + (long)nm * nm * z, yxz = y + (long)nm * x + (long)nm * nm * z,
                 yzx = y + (long)nm * z + (long)nm * nm * x, xzy = x + (long)nm * z + (long)nm * nm * y,
                 zxy = z + (long)nm * x + (long)nm * nm * y, zyx = z + (long)nm * y + (long)nm * nm * x;
      const long v[12] = {xyz, nm3 + xyz, yxz, nm3 + yxz, yzx, nm3 + yzx,
                          xzy, nm3 + xzy, zxy, nm3 + zxy, zyx, nm3 + zyx};
      for (int t = 0; t < 12; ++t) e[t] = (diag && t >= 6) ? -1 : v[t];
   };
   std::vector<long> cnt(nent + 1, 0);
   mflush = 1;
   walk([&](long, int, int f, int x, int y, int z, bool diag) {
      mflush = std::max(mflush, f + 1);
      long e[12];
      entries(x, y, z, diag, e);
      for (int t = 0; t < 12; ++t)
         if (e[t] >= 0) ++cnt[e[t] + 1];
   });
   off.assign(nent + 1, 0);
   for (long e = 0; e < nent; ++e) off[e + 1] = off[e] + cnt[e + 1];
   idx.assign(off[nent], 0);
   std::vector<long> pos(off.begin(), off.end() - 1);
   walk([&](long lid, int w, int f, int x, int y, int z, bool diag) {
      long e[12];
      entries(x, y, z, diag, e);
      const long base = ((lid * 4 + w) * (long)mflush + f) * 12;
      for (int t = 0; t < 12; ++t)
         if (e[t] >= 0) idx[pos[e[t]]++] = base + t;
   });
}

int plan_setup(bsahip_plan *P, const bsahip_problem *p, int32_t nfh, double df, int32_t flags,
               int psd_mode)
{
   CHK(check_flags(flags, true, true));
   if (nfh < 2 || !(df > 0.0)) return fail(BSAHIP_EINVAL, "need nfh >= 2 and df > 0");
   if (flags & BSAHIP_ONLY_DIAG)
      return fail(BSAHIP_EUNSUPPORTED, "the ONLY_DIAG model runs through bsahip_classic_run");
   CHK(load_problem(p, psd_mode, P->B, P->D));
   P->flags = flags;
   P->prec = flags & BSAHIP_PREC_MASK;
   P->df = df;
   P->nh = nfh;
   P->nf = 2 * nfh - 1;
   P->nu = (flags & BSAHIP_SUM_EXACT) ? P->nf : P->nh;          // rows of A per mode
   // rows of consecutive modes packed (no padding to 16 per mode): a 16-row
   // block may span two modes, and the epilogue masks the other mode's rows
   P->nhp = P->nu;
   P->rows = P->D.nm * P->nhp;
   const int tile = (P->prec == BSAHIP_PREC_FP64) ? TILE : T2;
   P->ntile = (P->rows + tile - 1) / tile;
   P->ldU = (long)round_up(P->rows, T2);
   P->nft = 2 * P->nu - 1;
   P->d = p->pulsation ? df * 2.0 * M_PI : df;
   const long npairs = (long)P->ntile * (P->ntile + 1) / 2;
   // one workgroup of 256 threads per (tile pair, mode): the launch's global
   // size must fit 32 bits
   if (npairs * P->D.nm * 256L > 0xffffffffL)
      return fail(BSAHIP_EUNSUPPORTED, "grid too large for one launch: (tile pairs x modes) x 256 "
                                       "threads exceeds 2^32; use a coarser grid or fewer modes");
   std::vector<double> fl(P->nu);
   for (int u = 0; u < P->nu; ++u) fl[u] = (double)u * df;
   CHK(upload(P->B, "fl", fl, &P->fl));
   CHK(P->B.get("U", (size_t)P->D.K * P->ldU, &P->U));
   CHK(P->B.get("H", (size_t)P->D.nm * P->nft, &P->H));
   const size_t nm3 = (size_t)P->D.nm * P->D.nm * P->D.nm, nm2 = (size_t)P->D.nm * P->D.nm;
   CHK(P->B.get("m3", 2 * nm3, &P->m3));
   CHK(P->B.get("m2", 3 * nm2, &P->m2));
   if ((flags & BSAHIP_DETERMINISTIC) && !(flags & BSAHIP_NO_BISP)) {
      std::vector<long> off, idx;
      det_slot_lists(P, off, idx, P->mflush);
      CHK(upload(P->B, "det_off", off, &P->det_off));
      CHK(upload(P->B, "det_idx", idx, &P->det_idx));
      CHK(P->B.get("det_slots", (size_t)npairs * P->D.nm * 4 * P->mflush * 12, &P->slots));
   }
   if ((flags & BSAHIP_DETERMINISTIC) && !(flags & BSAHIP_NO_PSD))
      CHK(P->B.get("m2part", (size_t)((P->nh + 63) / 64) * 3 * nm2, &P->m2part));
   if (P->prec != BSAHIP_PREC_FP64) {
      const size_t n = (size_t)P->ldU * P->D.K;
      CHK(P->B.get("UR32", n, &P->UR));
      if (P->prec == BSAHIP_PREC_BF16X3) {
         CHK(P->B.get("UCh", n, &P->Ch));
         CHK(P->B.get("UCl", n, &P->Cl));
      }
   }
   HIPC(hipStreamSynchronize(g.s));
   return 0;
}

// Phases of a plan run (all stream-ordered on g.s, no host synchronisation):
// tables (PSD, projection, FRF, split planes), G contraction, modal PSD moments.
int plan_prepare(bsahip_plan *P)
{
   Prob &D = P->D;
   const size_t nm3 = (size_t)D.nm * D.nm * D.nm;
   CHK(build_table(P->B, D, P->fl, P->nu, P->nhp, P->U, P->ldU));
   CHK(launch_frf(D, P->nft, P->df, P->H));
   if (!(P->flags & BSAHIP_NO_BISP)) {
      CHK(fill(P->m3, 2 * nm3, 0.0));
      if (P->prec != BSAHIP_PREC_FP64) {
         dim3 gs((unsigned)(P->ldU / 32), (unsigned)((D.K + 31) / 32));
         k_split_table<<<gs, 256, 0, g.s>>>(P->U, P->ldU, D.K, (int)P->ldU, D.K, P->UR, P->Ch, P->Cl);
         HIPC(hipGetLastError());
      }
   }
   return 0;
}

void plan_args(const bsahip_plan *P, GsymArgs &a, GsymArgs16 &b16)
{
   const Prob &D = P->D;
   const size_t nm3 = (size_t)D.nm * D.nm * D.nm;
   a.U = P->U;
   a.ldU = P->ldU;
   a.pb = D.pb_k;
   a.K = D.K;
   a.nm = D.nm;
   a.nhp = P->nhp;
   a.nh = P->nh;
   a.nu = P->nu;
   a.ntile = P->ntile;
   a.npairs = (long)P->ntile * (P->ntile + 1) / 2;
   a.H = P->H;
   a.nft = P->nft;
   a.d2 = P->d * P->d;
   a.rule = (P->flags & BSAHIP_QUAD_MASK) >> 4;
   a.epi_plain = (P->flags & BSAHIP_EPI_GENERIC) ? 1 : 0;
   a.m3f = P->m3;
   a.m3r = P->m3 + nm3;
   a.slots = P->slots;
   a.mflush = P->mflush;
   b16.UR = P->UR;
   b16.UCh = reinterpret_cast<const uint4 *>(P->Ch);
   b16.UCl = reinterpret_cast<const uint4 *>(P->Cl);
   b16.pb32 = D.pb32;
   b16.ldk = D.K;
}

long plan_items(const bsahip_plan *P)
{
   return (long)P->ntile * (P->ntile + 1) / 2 * P->D.nm;
}

int plan_gsym(bsahip_plan *P)
{
   GsymArgs a;
   GsymArgs16 b16;
   plan_args(P, a, b16);
   const long nwg = plan_items(P);
   if (P->prec == BSAHIP_PREC_BF16X3)
      k_gsym_bf16x3<<<dim3((unsigned)nwg), 256, 0, g.s>>>(a, b16);
   else if (P->prec == BSAHIP_PREC_FP32)
      k_gsym_f32<<<dim3((unsigned)nwg), 256, 0, g.s>>>(a, b16);
   else
      k_gsym_f64<<<dim3((unsigned)nwg), 256, 0, g.s>>>(a);
   HIPC(hipGetLastError());
   if (P->slots) {
      // fixed order: every m3 entry sums its slots in ascending order
      const long n = 2 * (long)P->D.nm * P->D.nm * P->D.nm;
      k_sum_lists<<<(unsigned)((n + 255) / 256), 256, 0, g.s>>>(P->slots, P->det_off, P->det_idx, n,
                                                                P->m3);
      HIPC(hipGetLastError());
   }
   return 0;
}

int plan_m2(bsahip_plan *P)
{
   Prob &D = P->D;
   const size_t nm2 = (size_t)D.nm * D.nm;
   CHK(fill(P->m2, 3 * nm2, 0.0));
   const int nby = (P->nh + 63) / 64;
   dim3 gm((unsigned)(D.nm * D.nm), (unsigned)nby);
   k_m2<<<gm, 256, 0, g.s>>>(P->U, P->ldU, D.pa_k, D.K, D.nm, P->nhp, P->nh, P->H, P->nft, P->d,
                             2.0 * M_PI * P->df, P->m2, P->m2 + nm2, P->m2 + 2 * nm2, P->m2part);
   HIPC(hipGetLastError());
   if (P->m2part) {
      const long n = 3 * (long)nm2;
      k_sum_fixed<<<(unsigned)((n + 255) / 256), 256, 0, g.s>>>(P->m2part, nby, n, n, P->m2);
      HIPC(hipGetLastError());
   }
   return 0;
}

int plan_copyout(bsahip_plan *P, double *m2mf, double *m2mr, double *m2o2mr, double *m3mf,
                 double *m3mr)
{
   const Prob &D = P->D;
   const size_t nm3 = (size_t)D.nm * D.nm * D.nm, nm2 = (size_t)D.nm * D.nm;
   if (!(P->flags & BSAHIP_NO_BISP)) {
      if (m3mf) HIPC(hipMemcpyAsync(m3mf, P->m3, nm3 * sizeof(double), hipMemcpyDeviceToHost, g.s));
      if (m3mr) HIPC(hipMemcpyAsync(m3mr, P->m3 + nm3, nm3 * sizeof(double), hipMemcpyDeviceToHost, g.s));
   }
   if (!(P->flags & BSAHIP_NO_PSD)) {
      if (m2mf) HIPC(hipMemcpyAsync(m2mf, P->m2, nm2 * sizeof(double), hipMemcpyDeviceToHost, g.s));
      if (m2mr) HIPC(hipMemcpyAsync(m2mr, P->m2 + nm2, nm2 * sizeof(double), hipMemcpyDeviceToHost, g.s));
      if (m2o2mr) HIPC(hipMemcpyAsync(m2o2mr, P->m2 + 2 * nm2, nm2 * sizeof(double), hipMemcpyDeviceToHost, g.s));
   }
   return 0;
}

int plan_exec(bsahip_plan *P, double *m2mf, double *m2mr, double *m2o2mr, double *m3mf, double *m3mr)
{
   const bool want_psd = !(P->flags & BSAHIP_NO_PSD);
   const bool want_bisp = !(P->flags & BSAHIP_NO_BISP);
   CHK(plan_prepare(P));
   P->last_ms = 0.0;
   // the modal PSD moments only read the tables: they run on the side stream
   // while the G kernel runs, and the copy-out waits for both
   const bool side = want_psd && want_bisp;
   if (side) {
      HIPC(hipEventRecord(g.e2, g.s));
      HIPC(hipStreamWaitEvent(g.s2, g.e2, 0));
      hipStream_t main_s = g.s;
      g.s = g.s2;
      const int rc = plan_m2(P);
      g.s = main_s;
      CHK(rc);
      HIPC(hipEventRecord(g.e3, g.s2));
   }
   if (want_bisp) {
      HIPC(hipEventRecord(g.e0, g.s));
      CHK(plan_gsym(P));
      HIPC(hipEventRecord(g.e1, g.s));
   }
   if (want_psd && !side) CHK(plan_m2(P));
   if (side) HIPC(hipStreamWaitEvent(g.s, g.e3, 0));
   CHK(plan_copyout(P, m2mf, m2mr, m2o2mr, m3mf, m3mr));
   HIPC(hipStreamSynchronize(g.s));
   g.last_ms = 0.0;
   g.last_launches = 0;
   if (want_bisp) {
      float ms = 0.f;
      HIPC(hipEventElapsedTime(&ms, g.e0, g.e1));
      P->last_ms = ms;
      g.last_ms = ms;
      g.last_launches = 1;
   }
   return 0;
}

Bufs &oneshot_bufs()
{
   if (!g.bufs) g.bufs = new Bufs();
   return *g.bufs;
}

}  // namespace

extern "C" {

int bsahip_version(void) { return 110; }

const char *bsahip_last_error(void) { return g_err.c_str(); }

int bsahip_init(int device)
{
   std::lock_guard<std::recursive_mutex> lk(g.mu);
   if (g.ready) return 0;
   int n = 0;
   hipError_t e = hipGetDeviceCount(&n);
   if (e != hipSuccess || n <= 0)
      return fail(BSAHIP_EDEVICE, std::string("no HIP device available: ") + hipGetErrorString(e));
   if (device < 0 || device >= n) return fail(BSAHIP_EINVAL, "device index out of range");
   if (device >= BSAHIP_MAX_DEVICES) return fail(BSAHIP_EUNSUPPORTED, "device index above 63");
   HIPC(hipSetDevice(device));
   if (!g.scratch_dev[device]) {
      void *sp = nullptr;
      HIPC(hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)));
      if (!sp) return fail(BSAHIP_EDEVICE, "hipGetSymbolAddress(scratch) returned NULL");
      g.scratch_dev[device] = sp;
   }
   HIPC(hipStreamCreateWithFlags(&g.s, hipStreamNonBlocking));
   HIPC(hipStreamCreateWithFlags(&g.s2, hipStreamNonBlocking));
   HIPC(hipEventCreate(&g.e0));
   HIPC(hipEventCreate(&g.e1));
   HIPC(hipEventCreateWithFlags(&g.e2, hipEventDisableTiming));
   HIPC(hipEventCreateWithFlags(&g.e3, hipEventDisableTiming));
   g.dev = device;
   g.ready = true;
   return 0;
}

int bsahip_finalise(void)
{
   std::lock_guard<std::recursive_mutex> lk(g.mu);
   if (!g.ready) return 0;
   (void)hipStreamSynchronize(g.s);
   delete g.bufs;
   g.bufs = nullptr;
   if (g.blas) (void)rocblas_destroy_handle(g.blas);
   g.blas = nullptr;
   (void)hipStreamSynchronize(g.s2);
   (void)hipEventDestroy(g.e0);
   (void)hipEventDestroy(g.e1);
   (void)hipEventDestroy(g.e2);
   (void)hipEventDestroy(g.e3);
   (void)hipStreamDestroy(g.s2);
   (void)hipStreamDestroy(g.s);
   g.ready = false;
   return 0;
}

int bsahip_last_kernel_time(double *ms, int32_t *launches)
{
   if (ms) *ms = g.last_ms;
   if (launches) *launches = g.last_launches;
   return 0;
}

int bsahip_plan_create(const bsahip_problem *p, int32_t nfh, double df, int32_t flags,
                       bsahip_plan **out)
{
   std::lock_guard<std::recursive_mutex> lk(g.mu);
   if (!out) return fail(BSAHIP_EINVAL, "NULL plan output");
   *out = nullptr;
   CHK(ensure_init());
   std::unique_ptr<bsahip_plan> P(new bsahip_plan());
   CHK(plan_setup(P.get(), p, nfh, df, flags, 0));
   *out = P.release();
   return 0;
}

int bsahip_plan_run(bsahip_plan *plan, double *m2mf, double *m2mr, double *m2o2mr,
                    double *m3mf, double *m3mr)
{
   std::lock_guard<std::recursive_mutex> lk(g.mu);
   if (!plan) return fail(BSAHIP_EINVAL, "NULL plan");
   CHK(ensure_init());
   return plan_exec(plan, m2mf, m2mr, m2o2mr, m3mf, m3mr);
}

int bsahip_plan_destroy(bsahip_plan *plan)
{
   std::lock_guard<std::recursive_mutex> lk(g.mu);
   if (plan) {
      if (g.ready) (void)hipStreamSynchronize(g.s);
      delete plan;
   }
   return 0;
}

int bsahip_classic_run(const bsahip_problem *p, int32_t nfh, double df, int32_t flags,
                       double *m2mf, double *m2mr, double *m2o2mr, double *m3mf, double *m3mr)
{
   return bsahip__classic_moments(p, nfh, df, flags, 0, m2mf, m2mr, m2o2mr, m3mf, m3mr);
}

int bsahip_classic_run_noalloc(const bsahip_problem *p, int32_t nfh, double df, int32_t flags,
                               double *m2mf, double *m2mr, double *m2o2mr, double *m3mf,
                               double *m3mr)
{
   std::lock_guard<std::recursive_mutex> lk(g.mu);
   if (!g.ready) return fail(BSAHIP_ESTATE, "call bsahip_init first (it sets up the static scratch)");
   if (flags & BSAHIP_ONLY_DIAG)
      return fail(BSAHIP_EUNSUPPORTED, "the ONLY_DIAG model runs through bsahip_classic_run");
   void *base = g.scratch_dev[g.dev];
   if (!base) return fail(BSAHIP_ESTATE, "no static scratch on the current device");
   // the plan's buffers come from the static array; the plan is never
   // destroyed with buffers of its own (Bufs in arena mode frees nothing)
   bsahip_plan P;
   P.B.set_arena(base, BSAHIP_NOALLOC_SCRATCH_BYTES);
   CHK(plan_setup(&P, p, nfh, df, flags, 0));
   return plan_exec(&P, m2mf, m2mr, m2o2mr, m3mf, m3mr);
}

}  // extern "C"

namespace {

// Tables of the ONLY_DIAG model on the half grid u df, u < nh: X (K, nh) the
// nodal PSDs, T (nm, nh, nh) the per-mode GEMM T_m = S W_m S^T (when with_T),
// H (nm, 2 nh - 1) the FRFs.
int diag_tables(Bufs &B, const Prob &D, int nh, double df, bool with_T, double **X, double **T,
                double2 **H)
{
   const int nft = 2 * nh - 1, nm = D.nm, K = D.K;
   std::vector<double> fl(nh);
   for (int u = 0; u < nh; ++u) fl[u] = (double)u * df;
   double *dfl, *sq, *Y;
   CHK(upload(B, "fl", fl, &dfl));
   CHK(B.get("sq", (size_t)D.ntc * nh * D.nnl, &sq));
   CHK(B.get("dg_X", (size_t)K * nh, X));
   CHK(B.get("dg_Y", (size_t)nm * K * nh, &Y));
   CHK(B.get("dg_H", (size_t)nm * nft, H));
   dim3 gp((D.nnl + 63) / 64, nh, D.ntc);
   k_psd_sqrt<<<gp, 64, 0, g.s>>>(dfl, nh, D.prm, D.formula, D.nnl, D.ntc, D.pulsation, sq);
   HIPC(hipGetLastError());
   k_diag_tables<<<dim3((unsigned)((nh + 63) / 64), (unsigned)K), 64, 0, g.s>>>(
      sq, nh, D.nnl, D.kv, D.pa_k, D.pb_k, K, nm, *X, Y);
   HIPC(hipGetLastError());
   CHK(launch_frf(D, nft, df, *H));
   *T = nullptr;
   if (with_T) {
      CHK(B.get("dg_T", (size_t)nm * nh * nh, T));
      CHK(dgemm_sb(0, 1, nh, nh, K, 1.0, *X, nh, 0, Y, nh, (long)K * nh, 0.0, *T, nh, (long)nh * nh,
                   nm));
   }
   return 0;
}

// The ONLY_DIAG Classic model (diag.hip): outputs have NM entries each.
int classic_diag(const bsahip_problem *p, int32_t nfh, double df, int32_t flags, double *m2mf,
                 double *m2mr, double *m2o2mr, double *m3mf, double *m3mr)
{
   if ((flags & BSAHIP_PREC_MASK) != BSAHIP_PREC_FP64)
      return fail(BSAHIP_EUNSUPPORTED, "the ONLY_DIAG model runs in FP64");
   if (nfh < 2 || !(df > 0.0)) return fail(BSAHIP_EINVAL, "need nfh >= 2 and df > 0");
   Bufs &B = oneshot_bufs();
   Prob D;
   CHK(load_problem(p, 0, B, D));
   const int nh = nfh, nf = 2 * nh - 1, nft = 2 * nh - 1, nm = D.nm, K = D.K;
   const int rule = (flags & BSAHIP_QUAD_MASK) >> 4;
   const double d = p->pulsation ? df * 2.0 * M_PI : df;
   double *X, *T, *mom;
   double2 *H;
   CHK(diag_tables(B, D, nh, df, !(flags & BSAHIP_NO_BISP), &X, &T, &H));
   CHK(B.get("dg_mom", 5 * (size_t)nm, &mom));
   CHK(fill(mom, 5 * (size_t)nm, 0.0));
   double *m2f = mom, *m2r = mom + nm, *m2o = mom + 2 * nm, *m3f = mom + 3 * nm, *m3r = mom + 4 * nm;
   if (!(flags & BSAHIP_NO_BISP)) {
      k_diag_m3<<<dim3((unsigned)nf, (unsigned)nm), 256, 0, g.s>>>(T, nh, H, nft, rule, d * d, m3f, m3r);
      HIPC(hipGetLastError());
   }
   if (!(flags & BSAHIP_NO_PSD)) {
      k_diag_m2<<<dim3((unsigned)((nh + 63) / 64), (unsigned)nm), 64, 0, g.s>>>(
         X, nh, D.pa_k, K, H, nft, d, 2.0 * M_PI * df, m2f, m2r, m2o);
      HIPC(hipGetLastError());
   }
   std::vector<double> h(5 * (size_t)nm);
   HIPC(hipMemcpyAsync(h.data(), mom, h.size() * sizeof(double), hipMemcpyDeviceToHost, g.s));
   HIPC(hipStreamSynchronize(g.s));
   auto put = [&](double *dst, int slot, bool on) {
      if (dst && on) std::memcpy(dst, h.data() + (size_t)slot * nm, nm * sizeof(double));
   };
   const bool psd = !(flags & BSAHIP_NO_PSD), bisp = !(flags & BSAHIP_NO_BISP);
   put(m2mf, 0, psd);
   put(m2mr, 1, psd);
   put(m2o2mr, 2, psd);
   put(m3mf, 3, bisp);
   put(m3mr, 4, bisp);
   return 0;
}

}  // namespace

extern "C" {

int bsahip__classic_moments(const bsahip_problem *p, int32_t nfh, double df, int32_t flags,
                            int32_t psd_mode, double *m2mf, double *m2mr, double *m2o2mr,
                            double *m3mf, double *m3mr)
{
   std::lock_guard<std::recursive_mutex> lk(g.mu);
   CHK(ensure_init());
   if (flags & BSAHIP_ONLY_DIAG) {
      if (psd_mode != 0) return fail(BSAHIP_EUNSUPPORTED, "the ONLY_DIAG model has no BsaCL form");
      if (flags & BSAHIP_DETERMINISTIC)
         return fail(BSAHIP_EUNSUPPORTED, "the ONLY_DIAG model has no fixed-order mode");
      return classic_diag(p, nfh, df, flags, m2mf, m2mr, m2o2mr, m3mf, m3mr);
   }
   bsahip_plan P;
   CHK(plan_setup(&P, p, nfh, df, flags, psd_mode));
   return plan_exec(&P, m2mf, m2mr, m2o2mr, m3mf, m3mr);
}

int bsahip__points_sum(const bsahip_problem *p, int32_t npts, const double *fi, const double *fj,
                       const double *w, int32_t psd_mode, int32_t with_frf, double *m3)
{
   std::lock_guard<std::recursive_mutex> lk(g.mu);
   CHK(ensure_init());
   if (npts < 0 || (npts > 0 && (!fi || !fj || !w || !m3)))
      return fail(BSAHIP_EINVAL, "npts < 0 or NULL point/weight/output arrays");
   Bufs &B = oneshot_bufs();
   Prob D;
   CHK(load_problem(p, psd_mode, B, D));
   const size_t nm3 = (size_t)D.nm * D.nm * D.nm;
   if (npts == 0) {
      std::fill(m3, m3 + nm3, 0.0);
      return 0;
   }
   std::vector<double> uq;
   std::vector<int> i1, i2, i3;
   point_tables(npts, fi, fj, uq, i1, i2, i3);
   const int nu = (int)uq.size();
   const long ldU = (long)D.nm * nu;
   double *fl, *U;
   CHK(upload(B, "fl", uq, &fl));
   CHK(B.get("U", (size_t)D.K * ldU, &U));
   CHK(build_table(B, D, fl, nu, nu, U, ldU));
   std::vector<double> f1(fi, fi + npts), f2(fj, fj + npts), ww(w, w + npts);
   int *di1, *di2, *di3;
   double *df1, *df2, *dw, *out;
   CHK(upload(B, "i1", i1, &di1));
   CHK(upload(B, "i2", i2, &di2));
   CHK(upload(B, "i3", i3, &di3));
   CHK(upload(B, "f1", f1, &df1));
   CHK(upload(B, "f2", f2, &df2));
   CHK(upload(B, "wpt", ww, &dw));
   CHK(B.get("m3sum", nm3, &out));
   CHK(fill(out, nm3, 0.0));
   const int nchunk = (int)std::min<long>(1024, std::max<long>(1, npts / 16));
   dim3 gs((unsigned)((nm3 + 255) / 256), nchunk);
   k_bfm_points_sum<<<gs, 256, 0, g.s>>>(U, ldU, nu, D.pb_k, D.K, D.nm, npts, di1, di2, di3,
                                         df1, df2, dw, D.Mm, D.Km, D.Cm, with_frf, out);
   HIPC(hipGetLastError());
   HIPC(hipMemcpyAsync(m3, out, nm3 * sizeof(double), hipMemcpyDeviceToHost, g.s));
   HIPC(hipStreamSynchronize(g.s));
   return 0;
}

int bsahip_classic_spectra(const bsahip_problem *p, int32_t nfh, double df, int32_t flags,
                           double *psd_f, double *psd_r, double *bfm, double *brm)
{
   std::lock_guard<std::recursive_mutex> lk(g.mu);
   CHK(ensure_init());
   CHK(check_flags(flags, true, false));
   if (flags & BSAHIP_ONLY_DIAG)
      return fail(BSAHIP_EUNSUPPORTED, "bsahip_classic_spectra is the full model; the diagonal model's "
                                       "spectra: bsahip_diag_psd_points, bsahip_classic_bispectra_sel");
   if (nfh < 2 || !(df > 0.0)) return fail(BSAHIP_EINVAL, "need nfh >= 2 and df > 0");
   Bufs &B = oneshot_bufs();
   Prob D;
   CHK(load_problem(p, 0, B, D));
   const int nh = nfh, nf = 2 * nh - 1, c = nh - 1;
   const bool exact = (flags & BSAHIP_SUM_EXACT) != 0;
   const int nu = exact ? nf : nh;
   const long ldU = (long)D.nm * nu;
   std::vector<double> fl(nu);
   for (int u = 0; u < nu; ++u) fl[u] = (double)u * df;
   double *dfl, *U;
   double2 *H;
   CHK(upload(B, "fl", fl, &dfl));
   CHK(B.get("U", (size_t)D.K * ldU, &U));
   CHK(build_table(B, D, dfl, nu, nu, U, ldU));
   CHK(B.get("H", (size_t)D.nm * nf, &H));
   CHK(launch_frf(D, nf, df, H));
   const size_t nm2 = (size_t)D.nm * D.nm, nm3 = nm2 * D.nm;
   if (psd_f || psd_r) {
      double *pf, *pr;
      CHK(B.get("psd_f", nm2 * nf, &pf));
      CHK(B.get("psd_r", nm2 * nf, &pr));
      dim3 gq((nf + 255) / 256, (unsigned)nm2);
      k_psd_grid<<<gq, 256, 0, g.s>>>(U, ldU, D.pa_k, D.K, D.nm, nu, nf, c, H, nf, pf, pr);
      HIPC(hipGetLastError());
      if (psd_f) HIPC(hipMemcpyAsync(psd_f, pf, nm2 * nf * sizeof(double), hipMemcpyDeviceToHost, g.s));
      if (psd_r) HIPC(hipMemcpyAsync(psd_r, pr, nm2 * nf * sizeof(double), hipMemcpyDeviceToHost, g.s));
   }
   if (bfm || brm) {
      const long npts = (long)nf * nf;
      std::vector<int> i1(npts), i2(npts), i3(npts);
      std::vector<double> f1(npts), f2(npts);
      for (int q2 = 0; q2 < nf; ++q2)
         for (int q1 = 0; q1 < nf; ++q1) {
            const long pt = q1 + (long)nf * q2;
            const int F1 = q1 - c, F2 = q2 - c, F3 = F1 + F2;
            i1[pt] = std::abs(F1);
            i2[pt] = std::abs(F2);
            i3[pt] = (exact || std::abs(F3) <= c) ? std::abs(F3) : -1;
            f1[pt] = F1 * df;
            f2[pt] = F2 * df;
         }
      int *di1, *di2, *di3;
      double *df1, *df2, *ob, *orr;
      CHK(upload(B, "i1", i1, &di1));
      CHK(upload(B, "i2", i2, &di2));
      CHK(upload(B, "i3", i3, &di3));
      CHK(upload(B, "f1", f1, &df1));
      CHK(upload(B, "f2", f2, &df2));
      CHK(B.get("bfm_out", nm3 * npts, &ob));
      CHK(B.get("brm_out", nm3 * npts, &orr));
      const long tot = (long)nm3 * npts;
      if (tot > 0xffffff00L)
         return fail(BSAHIP_EUNSUPPORTED, "NF^2 * NM^3 exceeds 2^32 bispectrum values in one call");
      k_bfm_points<<<(unsigned)((tot + 255) / 256), 256, 0, g.s>>>(
         U, ldU, nu, D.pb_k, D.K, D.nm, (int)npts, di1, di2, di3, df1, df2, D.Mm, D.Km, D.Cm,
         ob, orr, 1, npts);
      HIPC(hipGetLastError());
      if (bfm) HIPC(hipMemcpyAsync(bfm, ob, nm3 * npts * sizeof(double), hipMemcpyDeviceToHost, g.s));
      if (brm) HIPC(hipMemcpyAsync(brm, orr, nm3 * npts * sizeof(double), hipMemcpyDeviceToHost, g.s));
   }
   HIPC(hipStreamSynchronize(g.s));
   return 0;
}

int bsahip_classic_bispectra_sel(const bsahip_problem *p, int32_t nfh, double df, int32_t flags,
                                 int32_t nsel, const int32_t *sel, double *bfm, double *brm)
{
   std::lock_guard<std::recursive_mutex> lk(g.mu);
   CHK(ensure_init());
   CHK(check_flags(flags, true, true));
   const int prec = flags & BSAHIP_PREC_MASK;
   if (nfh < 2 || !(df > 0.0)) return fail(BSAHIP_EINVAL, "need nfh >= 2 and df > 0");
   if (prec != BSAHIP_PREC_FP64 && (flags & BSAHIP_ONLY_DIAG))
      return fail(BSAHIP_EUNSUPPORTED, "the ONLY_DIAG model runs in FP64");
   if (nsel < 0 || nsel > 65535 || (nsel > 0 && !sel))
      return fail(BSAHIP_EINVAL, "nsel outside 0..65535 or NULL selection");
   const bool diag = (flags & BSAHIP_ONLY_DIAG) != 0, exact = (flags & BSAHIP_SUM_EXACT) != 0;
   if (diag && exact)
      return fail(BSAHIP_EUNSUPPORTED, "the ONLY_DIAG model pads S(f1+f2) with zeros (no SUM_EXACT)");
   if (nsel == 0 || (!bfm && !brm)) return 0;
   if (nfh > 46341) return fail(BSAHIP_EINVAL, "nfh too large for one selected bispectrum");
   Bufs &B = oneshot_bufs();
   Prob D;
   CHK(load_problem(p, 0, B, D));
   const long ncol = diag ? D.nm : (long)D.nm * D.nm * D.nm;
   std::vector<int> s0(nsel);
   for (int s = 0; s < nsel; ++s) {
      if (sel[s] < 1 || sel[s] > ncol)
         return fail(BSAHIP_EINVAL, "selection index out of 1..NM^3 (1..NM with ONLY_DIAG)");
      s0[s] = sel[s] - 1;
   }
   const int nh = nfh, nf = 2 * nh - 1;
   const size_t npts = (size_t)nf * nf, nout = npts * (size_t)nsel;
   int *dsel;
   double *ob = nullptr, *orr = nullptr;
   CHK(upload(B, "sel", s0, &dsel));
   if (bfm) CHK(B.get("bisp_sel_f", nout, &ob));
   if (brm) CHK(B.get("bisp_sel_r", nout, &orr));
   const dim3 gs((unsigned)((npts + 255) / 256), (unsigned)nsel);
   if (diag) {
      double *X, *T;
      double2 *H;
      CHK(diag_tables(B, D, nh, df, true, &X, &T, &H));
      k_diag_bisp_sel<<<gs, 256, 0, g.s>>>(T, nh, H, 2 * nh - 1, dsel, ob, orr);
   } else {
      const int nu = exact ? nf : nh;
      const long ldU = (long)D.nm * nu;
      std::vector<double> fl(nu);
      for (int u = 0; u < nu; ++u) fl[u] = (double)u * df;
      double *dfl, *U;
      CHK(upload(B, "fl", fl, &dfl));
      CHK(B.get("U", (size_t)D.K * ldU, &U));
      CHK(build_table(B, D, dfl, nu, nu, U, ldU));
#define LAUNCH_SEL(PR)                                                                        \
   k_bisp_grid_sel<PR><<<gs, 256, 0, g.s>>>(U, ldU, nu, D.pb_k, D.K, D.nm, nf, exact ? 1 : 0, df, \
                                            dsel, D.Mm, D.Km, D.Cm, ob, orr)
      if (prec == BSAHIP_PREC_FP32) LAUNCH_SEL(1);
      else if (prec == BSAHIP_PREC_BF16X3) LAUNCH_SEL(2);
      else LAUNCH_SEL(0);
#undef LAUNCH_SEL
   }
   HIPC(hipGetLastError());
   if (bfm) HIPC(hipMemcpyAsync(bfm, ob, nout * sizeof(double), hipMemcpyDeviceToHost, g.s));
   if (brm) HIPC(hipMemcpyAsync(brm, orr, nout * sizeof(double), hipMemcpyDeviceToHost, g.s));
   HIPC(hipStreamSynchronize(g.s));
   return 0;
}

}  // extern "C"

namespace {
int bfm_points_impl(const bsahip_problem *p, int32_t npts, const double *fi, const double *fj,
                    double *bfm, double *brm, double fmax, bool diag);
int diag_sq_table(Bufs &B, const Prob &D, const std::vector<double> &uq, double fmax, double **sq);
}  // namespace

extern "C" {

int bsahip_bfm_points(const bsahip_problem *p, int32_t npts, const double *fi, const double *fj,
                      int32_t flags, double *bfm, double *brm)
{
   std::lock_guard<std::recursive_mutex> lk(g.mu);
   CHK(ensure_init());
   CHK(check_flags(flags, false, false));
   if (flags & ~(BSAHIP_PREC_MASK | BSAHIP_ONLY_DIAG))
      return fail(BSAHIP_EINVAL, "bsahip_bfm_points takes BSAHIP_ONLY_DIAG and FP64 only");
   return bfm_points_impl(p, npts, fi, fj, bfm, brm, -1.0, (flags & BSAHIP_ONLY_DIAG) != 0);
}

int bsahip_bfm_points_cls(const bsahip_problem *p, int32_t nfh, double df, int32_t flags,
                          int32_t npts, const double *fi, const double *fj, double *bfm,
                          double *brm)
{
   std::lock_guard<std::recursive_mutex> lk(g.mu);
   CHK(ensure_init());
   if (nfh < 2 || !(df > 0.0)) return fail(BSAHIP_EINVAL, "need nfh >= 2 and df > 0");
   if (flags & ~BSAHIP_ONLY_DIAG)
      return fail(BSAHIP_EINVAL, "bsahip_bfm_points_cls takes BSAHIP_ONLY_DIAG only");
   return bfm_points_impl(p, npts, fi, fj, bfm, brm, (nfh - 1) * df,
                          (flags & BSAHIP_ONLY_DIAG) != 0);
}

int bsahip_diag_psd_points(const bsahip_problem *p, int32_t npts, const double *f, double *psd)
{
   std::lock_guard<std::recursive_mutex> lk(g.mu);
   CHK(ensure_init());
   if (npts < 0 || (npts > 0 && (!f || !psd))) return fail(BSAHIP_EINVAL, "npts < 0 or NULL arrays");
   if (npts == 0) return 0;
   Bufs &B = oneshot_bufs();
   Prob D;
   CHK(load_problem(p, 0, B, D));
   std::vector<double> af(npts);
   for (int q = 0; q < npts; ++q) af[q] = std::fabs(f[q]);
   std::vector<double> uq(af);
   std::sort(uq.begin(), uq.end());
   uq.erase(std::unique(uq.begin(), uq.end()), uq.end());
   std::vector<int> iu(npts);
   for (int q = 0; q < npts; ++q)
      iu[q] = (int)(std::lower_bound(uq.begin(), uq.end(), af[q]) - uq.begin());
   double *sq, *out;
   int *diu;
   CHK(diag_sq_table(B, D, uq, -1.0, &sq));
   CHK(upload(B, "dp_iu", iu, &diu));
   const long tot = (long)npts * D.nm;
   CHK(B.get("dp_out", (size_t)tot, &out));
   k_diag_psd_points<<<(unsigned)((tot + 255) / 256), 256, 0, g.s>>>(
      sq, (int)uq.size(), D.nnl, D.kv, D.pa_k, D.K, D.nm, npts, diu, out);
   HIPC(hipGetLastError());
   HIPC(hipMemcpyAsync(psd, out, (size_t)tot * sizeof(double), hipMemcpyDeviceToHost, g.s));
   HIPC(hipStreamSynchronize(g.s));
   return 0;
}

}  // extern "C"

namespace {

// sqrt-PSD table (ntc, nu, nnl) of the nodes at the distinct |f| uq; fmax > 0:
// rows with |f| beyond fmax zero (the Classic grid's truncation)
int diag_sq_table(Bufs &B, const Prob &D, const std::vector<double> &uq, double fmax, double **sq)
{
   const int nu = (int)uq.size();
   if (nu > 65535) return fail(BSAHIP_EUNSUPPORTED, "more than 65535 distinct frequencies in one table");
   double *fl;
   CHK(upload(B, "dg_fl", uq, &fl));
   CHK(B.get("dg_sq", (size_t)D.ntc * nu * D.nnl, sq));
   dim3 gp((D.nnl + 63) / 64, nu, D.ntc);
   k_psd_sqrt<<<gp, 64, 0, g.s>>>(fl, nu, D.prm, D.formula, D.nnl, D.ntc, D.pulsation, *sq);
   HIPC(hipGetLastError());
   if (fmax > 0.0) {
      const long u0 = std::upper_bound(uq.begin(), uq.end(), fmax * (1.0 + 1e-12)) - uq.begin();
      for (int c = 0; c < D.ntc && u0 < nu; ++c)
         CHK(fill(*sq + ((size_t)c * nu + u0) * D.nnl, (size_t)(nu - u0) * D.nnl, 0.0));
   }
   return 0;
}

// Point BFM/BRM; fmax > 0: the Classic grid's truncation, S(|f|) = 0 beyond
// fmax (it reaches only f1 + f2 of grid points); diag: the ONLY_DIAG model,
// NM entries per point (k_diag_points, the Mesher's diagonal evaluation)
int bfm_points_impl(const bsahip_problem *p, int32_t npts, const double *fi, const double *fj,
                    double *bfm, double *brm, double fmax, bool diag)
{
   if (npts < 0 || (npts > 0 && (!fi || !fj || !bfm)))
      return fail(BSAHIP_EINVAL, "npts < 0 or NULL point/output arrays");
   if (npts == 0) return 0;
   Bufs &B = oneshot_bufs();
   Prob D;
   CHK(load_problem(p, 0, B, D));
   std::vector<double> uq;
   std::vector<int> i1, i2, i3;
   point_tables(npts, fi, fj, uq, i1, i2, i3);
   if (diag) {
      const long tot = (long)npts * D.nm;
      if (tot > 0xffffff00L) return fail(BSAHIP_EUNSUPPORTED, "npts * NM exceeds 2^32 values");
      double *sq, *ob, *df1, *df2;
      int *d1, *d2, *d3;
      long *ds;
      std::vector<long> slot(npts);
      for (int q = 0; q < npts; ++q) slot[q] = q;
      CHK(diag_sq_table(B, D, uq, fmax, &sq));
      CHK(upload(B, "i1", i1, &d1));
      CHK(upload(B, "i2", i2, &d2));
      CHK(upload(B, "i3", i3, &d3));
      CHK(upload(B, "dg_slot", slot, &ds));
      CHK(B.get("bfm_out", (size_t)tot, &ob));
      k_diag_points<<<(unsigned)((tot + 255) / 256), 256, 0, g.s>>>(
         sq, (int)uq.size(), D.nnl, D.kv, D.pa_k, D.pb_k, D.K, D.nm, npts, d1, d2, d3, ds, ob);
      HIPC(hipGetLastError());
      HIPC(hipMemcpyAsync(bfm, ob, (size_t)tot * sizeof(double), hipMemcpyDeviceToHost, g.s));
      if (brm) {
         std::vector<double> f1(fi, fi + npts), f2(fj, fj + npts);
         double *orr;
         CHK(upload(B, "f1", f1, &df1));
         CHK(upload(B, "f2", f2, &df2));
         CHK(B.get("brm_out", (size_t)tot, &orr));
         HIPC(hipMemcpyAsync(orr, ob, (size_t)tot * sizeof(double), hipMemcpyDeviceToDevice, g.s));
         k_frf_apply<<<(unsigned)((tot + 255) / 256), 256, 0, g.s>>>(
            BSAHIP_FRF_POINTS, 1, npts, df1, npts, df2, D.nm, D.Mm, D.Km, D.Cm, tot, orr);
         HIPC(hipGetLastError());
         HIPC(hipMemcpyAsync(brm, orr, (size_t)tot * sizeof(double), hipMemcpyDeviceToHost, g.s));
      }
      HIPC(hipStreamSynchronize(g.s));
      return 0;
   }
   CHK(set_pod(p, D));
   const int nu = (int)uq.size();
   const long ldU = (long)D.nm * nu;
   double *dfl, *U;
   CHK(upload(B, "fl", uq, &dfl));
   CHK(B.get("U", (size_t)D.K * ldU, &U));
   CHK(build_table(B, D, dfl, nu, nu, U, ldU));
   if (fmax > 0.0) {
      // uq is ascending: columns u0.. lie beyond the grid
      const long u0 = std::upper_bound(uq.begin(), uq.end(), fmax * (1.0 + 1e-12)) - uq.begin();
      if (u0 < nu) {
         const long n = (long)D.K * D.nm * (nu - u0);
         k_zero_cols<<<(unsigned)((n + 255) / 256), 256, 0, g.s>>>(U, ldU, nu, u0, D.nm, n);
         HIPC(hipGetLastError());
      }
   }
   std::vector<double> f1(fi, fi + npts), f2(fj, fj + npts);
   int *di1, *di2, *di3;
   double *df1, *df2, *ob, *orr = nullptr;
   CHK(upload(B, "i1", i1, &di1));
   CHK(upload(B, "i2", i2, &di2));
   CHK(upload(B, "i3", i3, &di3));
   CHK(upload(B, "f1", f1, &df1));
   CHK(upload(B, "f2", f2, &df2));
   const size_t nm3 = (size_t)D.nm * D.nm * D.nm;
   CHK(B.get("bfm_out", nm3 * npts, &ob));
   if (brm) CHK(B.get("brm_out", nm3 * npts, &orr));
   const long tot = (long)nm3 * npts;
   if (tot > 0xffffff00L)
      return fail(BSAHIP_EUNSUPPORTED, "npts * NM^3 exceeds 2^32 bispectrum values in one call");
   k_bfm_points<<<(unsigned)((tot + 255) / 256), 256, 0, g.s>>>(
      U, ldU, nu, D.pb_k, D.K, D.nm, npts, di1, di2, di3, df1, df2, D.Mm, D.Km, D.Cm, ob, orr,
      (long)nm3, 1);
   HIPC(hipGetLastError());
   HIPC(hipMemcpyAsync(bfm, ob, nm3 * npts * sizeof(double), hipMemcpyDeviceToHost, g.s));
   if (brm) HIPC(hipMemcpyAsync(brm, orr, nm3 * npts * sizeof(double), hipMemcpyDeviceToHost, g.s));
   HIPC(hipStreamSynchronize(g.s));
   return 0;
}

}  // namespace

extern "C" {

int bsahip_frf_apply(const bsahip_problem *p, int32_t flags, int32_t kind, int64_t n1,
                     const double *f1, int64_t n2, const double *f2, double *data)
{
   std::lock_guard<std::recursive_mutex> lk(g.mu);
   CHK(ensure_init());
   if (flags & ~BSAHIP_ONLY_DIAG) return fail(BSAHIP_EINVAL, "bsahip_frf_apply takes BSAHIP_ONLY_DIAG only");
   const bool diag = (flags & BSAHIP_ONLY_DIAG) != 0;
   if (kind < BSAHIP_FRF_PSD || kind > BSAHIP_FRF_POINTS)
      return fail(BSAHIP_EINVAL, "kind must be BSAHIP_FRF_PSD, _GRID or _POINTS");
   if (n1 < 0 || n2 < 0 || (kind == BSAHIP_FRF_GRID && n2 == 0 && n1 > 0))
      return fail(BSAHIP_EINVAL, "negative sizes, or a grid without f2");
   if (n1 == 0) return 0;
   if (!f1 || !data || (kind != BSAHIP_FRF_PSD && !f2))
      return fail(BSAHIP_EINVAL, "NULL frequencies or data");
   Bufs &B = oneshot_bufs();
   Prob D;
   CHK(load_problem(p, 0, B, D));
   const long nm = D.nm, nm2 = diag ? nm : nm * nm, nm3 = diag ? nm : nm * nm * nm;
   const long nf2 = kind == BSAHIP_FRF_GRID ? n2 : (kind == BSAHIP_FRF_POINTS ? n1 : 0);
   const double tot_d = kind == BSAHIP_FRF_PSD ? (double)n1 * nm2
                       : (kind == BSAHIP_FRF_GRID ? (double)n1 * n2 * nm3 : (double)n1 * nm3);
   if (tot_d > (double)0xffffff00L)  // one thread per value, grid within 2^32 threads
      return fail(BSAHIP_EUNSUPPORTED, "more than 2^32 spectral values in one call");
   const long total = (long)tot_d;
   std::vector<double> h1(f1, f1 + n1), h2;
   if (nf2) h2.assign(f2, f2 + nf2);
   double *d1, *d2 = nullptr, *dd;
   CHK(upload(B, "frf_f1", h1, &d1));
   if (nf2) CHK(upload(B, "frf_f2", h2, &d2));
   CHK(B.get("frf_data", (size_t)total, &dd));
   HIPC(hipMemcpyAsync(dd, data, (size_t)total * sizeof(double), hipMemcpyHostToDevice, g.s));
   k_frf_apply<<<(unsigned)((total + 255) / 256), 256, 0, g.s>>>(kind, diag ? 1 : 0, n1, d1, n2, d2,
                                                                  D.nm, D.Mm, D.Km, D.Cm, total, dd);
   HIPC(hipGetLastError());
   HIPC(hipMemcpyAsync(data, dd, (size_t)total * sizeof(double), hipMemcpyDeviceToHost, g.s));
   HIPC(hipStreamSynchronize(g.s));
   return 0;
}

int bsahip_selftest_mfma_layout(double *out512)
{
   std::lock_guard<std::recursive_mutex> lk(g.mu);
   CHK(ensure_init());
   if (!out512) return fail(BSAHIP_EINVAL, "NULL output");
   double *d;
   CHK(oneshot_bufs().get("selftest", 512, &d));
   k_mfma_f64_layout<<<1, 64, 0, g.s>>>(d);
   HIPC(hipGetLastError());
   HIPC(hipMemcpyAsync(out512, d, 512 * sizeof(double), hipMemcpyDeviceToHost, g.s));
   HIPC(hipStreamSynchronize(g.s));
   return 0;
}

int bsahip_dgemm_strided_batched(int32_t transa, int32_t transb, int32_t M, int32_t N, int32_t K,
                                 double alpha, const double *A, int64_t lda, int64_t sA,
                                 const double *B, int64_t ldb, int64_t sB, double beta, double *C,
                                 int64_t ldc, int64_t sC, int32_t batch)
{
   std::lock_guard<std::recursive_mutex> lk(g.mu);
   CHK(ensure_init());
   if (M < 0 || N < 0 || K < 0 || batch < 1 || !A || !B || !C || (transa != 0 && transa != 1) ||
       (transb != 0 && transb != 1) || lda < (transa ? K : M) || ldb < (transb ? N : K) || ldc < M ||
       sA < 0 || sB < 0 || sC < 0)
      return fail(BSAHIP_EINVAL, "bad GEMM arguments");
   // extents of the host arrays (column-major, last batch included)
   const size_t na = (size_t)sA * (batch - 1) + (size_t)lda * (transa ? M : K);
   const size_t nb = (size_t)sB * (batch - 1) + (size_t)ldb * (transb ? K : N);
   const size_t nc = (size_t)sC * (batch - 1) + (size_t)ldc * N;
   Bufs &Bf = oneshot_bufs();
   double *dA, *dB, *dC;
   CHK(Bf.get("gemm_A", std::max<size_t>(na, 1), &dA));
   CHK(Bf.get("gemm_B", std::max<size_t>(nb, 1), &dB));
   CHK(Bf.get("gemm_C", std::max<size_t>(nc, 1), &dC));
   HIPC(hipMemcpyAsync(dA, A, na * sizeof(double), hipMemcpyHostToDevice, g.s));
   HIPC(hipMemcpyAsync(dB, B, nb * sizeof(double), hipMemcpyHostToDevice, g.s));
   HIPC(hipMemcpyAsync(dC, C, nc * sizeof(double), hipMemcpyHostToDevice, g.s));
   CHK(dgemm_sb(transa, transb, M, N, K, alpha, dA, lda, sA, dB, ldb, sB, beta, dC, ldc, sC, batch));
   HIPC(hipMemcpyAsync(C, dC, nc * sizeof(double), hipMemcpyDeviceToHost, g.s));
   HIPC(hipStreamSynchronize(g.s));
   return 0;
}

int bsahip_selftest_mfma_f32_layout(float *out512)
{
   std::lock_guard<std::recursive_mutex> lk(g.mu);
   CHK(ensure_init());
   if (!out512) return fail(BSAHIP_EINVAL, "NULL output");
   float *d;
   CHK(oneshot_bufs().get("selftest32", 512, &d));
   k_mfma_f32_layout<<<1, 64, 0, g.s>>>(d);
   HIPC(hipGetLastError());
   HIPC(hipMemcpyAsync(out512, d, 512 * sizeof(float), hipMemcpyDeviceToHost, g.s));
   HIPC(hipStreamSynchronize(g.s));
   return 0;
}

int bsahip_selftest_mfma_bf16_layout(float *out768)
{
   std::lock_guard<std::recursive_mutex> lk(g.mu);
   CHK(ensure_init());
   if (!out768) return fail(BSAHIP_EINVAL, "NULL output");
   float *d;
   CHK(oneshot_bufs().get("selftest16", 768, &d));
   k_mfma_bf16_layout<<<1, 64, 0, g.s>>>(d);
   HIPC(hipGetLastError());
   HIPC(hipMemcpyAsync(out768, d, 768 * sizeof(float), hipMemcpyDeviceToHost, g.s));
   HIPC(hipStreamSynchronize(g.s));
   return 0;
}

}  // extern "C"

#include "mesher.hip"
#include "recomb.hip"
#include "mlp.hip"
#include "batch.hip"
