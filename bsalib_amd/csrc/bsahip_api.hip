// C ABI of libbsahip.so (declarations and layouts: include/bsahip.h).
//
// Host side of the MI355X BFM path: validates the reference-layout inputs,
// builds the device tables (Psi per K index, ln rho, PSD parameters, FRF),
// launches the kernels of kernels.hip on the library stream and copies the
// moments back.  There is no CPU fallback: without a usable GPU every entry
// point returns BSAHIP_EDEVICE.
#include "kernels.hip"
#include "../../include/bsahip.h"
#include "bsahip_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
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

struct Buf {
   void *p = nullptr;
   size_t n = 0;
};

struct Ctx {
   bool ready = false;
   int dev = 0;
   hipStream_t s = nullptr;
   hipEvent_t e0 = nullptr, e1 = nullptr;
   std::map<std::string, Buf> bufs;
   double last_ms = 0.0;
   int last_launches = 0;
   std::recursive_mutex mu;
};

Ctx g;

template <class T>
int dbuf(const char *name, size_t count, T **out)
{
   Buf &b = g.bufs[name];
   const size_t bytes = std::max<size_t>(count * sizeof(T), 256);
   if (b.n < bytes) {
      if (b.p) (void)hipFree(b.p);
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

template <class T>
int upload(const char *name, const std::vector<T> &h, T **out)
{
   CHK(dbuf(name, h.size(), out));
   if (!h.empty()) HIPC(hipMemcpyAsync(*out, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, g.s));
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

inline long corr_index(long ni, long nj, long tot)
{
   if (nj >= ni) return (ni - 1) * tot + nj - (ni * ni - ni) / 2;
   return (nj - 1) * tot + ni - (nj * nj - nj) / 2;
}

inline int round_up(int a, int b) { return (a + b - 1) / b * b; }

// Host-prepared, device-resident description of one structure.
struct Prep {
   int nn, nnl, nnlp, nm, ntc, K;
   std::vector<double> pa_t;   // [tc][m][nnl]
   std::vector<double> pa_k;   // [m][K]
   std::vector<double> pb_k;   // [m][K]
   std::vector<double> lnr;    // [tc][nnl(j)][nnl(i)]
   std::vector<double> prm;    // [tc][nnl][4]
   std::vector<int> formula;   // [tc]
   std::vector<double> Mm, Km, Cm;
};

int validate(const bsahip_problem *p)
{
   if (!p) return fail(BSAHIP_EINVAL, "problem pointer is NULL");
   if (p->nn <= 0 || p->nnl <= 0 || p->nm <= 0 || p->ntc <= 0 || p->ntc > 3 || p->nwz <= 0)
      return fail(BSAHIP_EINVAL, "invalid sizes (nn, nnl, nm, ntc in 1..3, nwz must be > 0)");
   if (p->nnl > p->nn) return fail(BSAHIP_EINVAL, "nnl > nn");
   if (p->nm > 64) return fail(BSAHIP_EUNSUPPORTED, "more than 64 modes");
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

// psd_mode 0: reference Classic/Mesher semantics (evalPSD per component);
// psd_mode 1: BsaCL semantics (u formula for every component, zone 1, rho clamped).
int prepare(const bsahip_problem *p, Prep &P, int psd_mode = 0)
{
   CHK(validate(p));
   P.nn = p->nn;
   P.nnl = p->nnl;
   P.nm = p->nm;
   P.ntc = p->ntc;
   P.nnlp = round_up(p->nnl, KC);
   P.K = P.ntc * P.nnlp;
   const int nm = P.nm, nnl = P.nnl, K = P.K;
   P.pa_t.assign((size_t)P.ntc * nm * nnl, 0.0);
   P.pa_k.assign((size_t)nm * K, 0.0);
   P.pb_k.assign((size_t)nm * K, 0.0);
   for (int c = 0; c < P.ntc; ++c) {
      const int t = p->tc[c];
      for (int i = 0; i < nnl; ++i)
         for (int m = 0; m < nm; ++m) {
            const double a = p->psi[m + (size_t)nm * (i + (size_t)nnl * (t - 1))];
            const double b = p->psi[m + (size_t)nm * (i + (size_t)nnl * (t + 2))];
            P.pa_t[((size_t)c * nm + m) * nnl + i] = a;
            P.pa_k[(size_t)m * K + (size_t)c * P.nnlp + i] = a;
            P.pb_k[(size_t)m * K + (size_t)c * P.nnlp + i] = b;
         }
   }
   P.lnr.assign((size_t)P.ntc * nnl * nnl, 0.0);
   for (int c = 0; c < P.ntc; ++c) {
      const int col = (p->corr_col_mode == 1 || psd_mode == 1) ? 1 : p->tc[c];
      const double *rc = p->nod_corr + (size_t)(col - 1) * p->ncorr_rows;
      double *out = P.lnr.data() + (size_t)c * nnl * nnl;
      for (int j = 0; j < nnl; ++j)
         for (int i = 0; i < nnl; ++i) {
            double r = rc[corr_index(p->n_load[j], p->n_load[i], p->nn) - 1];
            if (psd_mode == 1 && r < 2.2250738585072014e-308) r = 2.2250738585072014e-308;
            out[(size_t)j * nnl + i] = std::log(r);
         }
   }
   P.prm.assign((size_t)P.ntc * nnl * 4, 0.0);
   P.formula.assign(P.ntc, 0);
   for (int c = 0; c < P.ntc; ++c) {
      const int t = p->tc[c];
      int f = 0;
      if (psd_mode == 1) f = (p->psd_id == 5) ? 5 : (p->psd_id == 1 ? 1 : 0);
      else if (p->psd_id == 1) f = (t == 1) ? 1 : 2;
      else f = (p->psd_id == 2) ? 0 : p->psd_id;
      P.formula[c] = f;
      for (int i = 0; i < nnl; ++i) {
         const int node = p->n_load[i] - 1;
         const int wz = (psd_mode == 1) ? 0 : p->wz_node[node] - 1;
         // L(component, direction 1, zone): von Karman u reads component 1
         // (WindPSDImpl.F90:167), the other formulas component tc.
         const int lc = (psd_mode == 0 && f == 1) ? 1 : t;
         double *q = P.prm.data() + ((size_t)c * nnl + i) * 4;
         q[0] = p->turb_scales[(lc - 1) + 9 * (size_t)wz];
         q[1] = p->turb_std[(t - 1) + 3 * (size_t)wz];
         q[2] = p->u_node[node];
         q[3] = p->u_mean_ref ? p->u_mean_ref[wz] : 1.0;
      }
   }
   P.Mm.assign(p->Mm, p->Mm + nm);
   P.Km.assign(p->Km, p->Km + nm);
   P.Cm.assign(p->Cm, p->Cm + nm);
   return 0;
}

struct DevProb {
   double *pa_t, *pa_k, *pb_k, *lnr, *prm, *Mm, *Km, *Cm;
   int *formula;
};

int upload_prep(const Prep &P, DevProb &D)
{
   CHK(upload("pa_t", P.pa_t, &D.pa_t));
   CHK(upload("pa_k", P.pa_k, &D.pa_k));
   CHK(upload("pb_k", P.pb_k, &D.pb_k));
   CHK(upload("lnr", P.lnr, &D.lnr));
   CHK(upload("prm", P.prm, &D.prm));
   CHK(upload("formula", P.formula, &D.formula));
   CHK(upload("Mm", P.Mm, &D.Mm));
   CHK(upload("Km", P.Km, &D.Km));
   CHK(upload("Cm", P.Cm, &D.Cm));
   return 0;
}

// A table on a list of |f| values: U[K][ldU], rows y*rows_per_mode + q.
int build_table(const Prep &P, const DevProb &D, const std::vector<double> &fabs_list,
                int rows_per_mode, int pulsation, double **U_out, long ldU)
{
   const int nfl = (int)fabs_list.size();
   if (nfl > 65535) return fail(BSAHIP_EUNSUPPORTED, "more than 65535 distinct frequencies in one table");
   double *fl, *sq, *U;
   CHK(upload("fabs", fabs_list, &fl));
   CHK(dbuf("sq", (size_t)P.ntc * nfl * P.nnl, &sq));
   CHK(dbuf("U", (size_t)P.K * ldU, &U));
   CHK(fill(U, (size_t)P.K * ldU, 0.0));
   dim3 gp((P.nnl + 63) / 64, nfl, P.ntc);
   k_psd_sqrt<<<gp, 64, 0, g.s>>>(fl, nfl, D.prm, D.formula, P.nnl, P.ntc, pulsation, sq);
   HIPC(hipGetLastError());
   const int nq = (P.nm + 3) / 4;
#define LAUNCH_PROJ(NQ)                                                                       \
   k_project<NQ><<<gp, 256, 0, g.s>>>(D.pa_t, D.lnr, sq, fl, nfl, P.nnl, P.nm, P.nnlp,        \
                                      rows_per_mode, U, ldU)
   if (nq <= 1) LAUNCH_PROJ(1);
   else if (nq <= 2) LAUNCH_PROJ(2);
   else if (nq <= 4) LAUNCH_PROJ(4);
   else if (nq <= 8) LAUNCH_PROJ(8);
   else LAUNCH_PROJ(16);
#undef LAUNCH_PROJ
   HIPC(hipGetLastError());
   *U_out = U;
   return 0;
}

int frf_table(const Prep &P, const DevProb &D, int nft, double df, double2 **H)
{
   CHK(dbuf("H", (size_t)P.nm * nft, H));
   dim3 gh((nft + 255) / 256, P.nm);
   k_frf_table<<<gh, 256, 0, g.s>>>(D.Mm, D.Km, D.Cm, P.nm, nft, df, *H);
   HIPC(hipGetLastError());
   return 0;
}

int check_flags(int32_t flags, bool grid, bool allow_mixed)
{
   const int prec = flags & BSAHIP_PREC_MASK;
   const int quad = (flags & BSAHIP_QUAD_MASK) >> 4;
   if (prec != BSAHIP_PREC_FP64 && !(allow_mixed && prec == BSAHIP_PREC_BF16X3))
      return fail(BSAHIP_EUNSUPPORTED, "precision not available in this entry point");
   if (grid && quad > 2) return fail(BSAHIP_EINVAL, "unknown quadrature rule");
   return 0;
}

}  // namespace

extern "C" {

int bsahip_version(void) { return 100; }

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
   HIPC(hipSetDevice(device));
   HIPC(hipStreamCreateWithFlags(&g.s, hipStreamNonBlocking));
   HIPC(hipEventCreate(&g.e0));
   HIPC(hipEventCreate(&g.e1));
   g.dev = device;
   g.ready = true;
   return 0;
}

int bsahip_finalise(void)
{
   std::lock_guard<std::recursive_mutex> lk(g.mu);
   if (!g.ready) return 0;
   (void)hipStreamSynchronize(g.s);
   for (auto &kv : g.bufs)
      if (kv.second.p) (void)hipFree(kv.second.p);
   g.bufs.clear();
   (void)hipEventDestroy(g.e0);
   (void)hipEventDestroy(g.e1);
   (void)hipStreamDestroy(g.s);
   g.ready = false;
   return 0;
}

int bsahip_selftest_mfma_layout(double *out512)
{
   std::lock_guard<std::recursive_mutex> lk(g.mu);
   CHK(ensure_init());
   if (!out512) return fail(BSAHIP_EINVAL, "NULL output");
   double *d;
   CHK(dbuf("selftest", 512, &d));
   k_mfma_f64_layout<<<1, 64, 0, g.s>>>(d);
   HIPC(hipGetLastError());
   HIPC(hipMemcpyAsync(out512, d, 512 * sizeof(double), hipMemcpyDeviceToHost, g.s));
   HIPC(hipStreamSynchronize(g.s));
   return 0;
}

int bsahip_selftest_mfma_bf16_layout(float *out768)
{
   std::lock_guard<std::recursive_mutex> lk(g.mu);
   CHK(ensure_init());
   if (!out768) return fail(BSAHIP_EINVAL, "NULL output");
   float *d;
   CHK(dbuf("selftest16", 768, &d));
   k_mfma_bf16_layout<<<1, 64, 0, g.s>>>(d);
   HIPC(hipGetLastError());
   HIPC(hipMemcpyAsync(out768, d, 768 * sizeof(float), hipMemcpyDeviceToHost, g.s));
   HIPC(hipStreamSynchronize(g.s));
   return 0;
}

int bsahip_last_kernel_time(double *ms, int32_t *launches)
{
   if (ms) *ms = g.last_ms;
   if (launches) *launches = g.last_launches;
   return 0;
}

int bsahip_classic_run(const bsahip_problem *p, int32_t nfh, double df, int32_t flags,
                       double *m2mf, double *m2mr, double *m2o2mr, double *m3mf, double *m3mr)
{
   return bsahip__classic_moments(p, nfh, df, flags, 0, m2mf, m2mr, m2o2mr, m3mf, m3mr);
}

int bsahip__classic_moments(const bsahip_problem *p, int32_t nfh, double df, int32_t flags,
                            int32_t psd_mode, double *m2mf, double *m2mr, double *m2o2mr,
                            double *m3mf, double *m3mr)
{
   std::lock_guard<std::recursive_mutex> lk(g.mu);
   CHK(ensure_init());
   CHK(check_flags(flags, true, true));
   if (nfh < 2 || !(df > 0.0)) return fail(BSAHIP_EINVAL, "need nfh >= 2 and df > 0");
   const int prec = flags & BSAHIP_PREC_MASK;
   Prep P;
   CHK(prepare(p, P, psd_mode));
   DevProb D;
   CHK(upload_prep(P, D));
   const bool exact = (flags & BSAHIP_SUM_EXACT) != 0;
   const int nh = nfh, nf = 2 * nh - 1;
   const int nu = exact ? nf : nh;                 // rows of A per mode
   const int nhp = round_up(nu, 16);
   const int rows = P.nm * nhp;
   const int tile = (prec == BSAHIP_PREC_BF16X3) ? T2 : TILE;
   const int ntile = (rows + tile - 1) / tile;
   const long ldU = (long)round_up(rows, T2);
   const int nft = 2 * nu - 1;
   std::vector<double> fl(nu);
   for (int u = 0; u < nu; ++u) fl[u] = (double)u * df;
   double *U;
   CHK(build_table(P, D, fl, nhp, p->pulsation, &U, ldU));
   double2 *H;
   CHK(frf_table(P, D, nft, df, &H));
   const double d = p->pulsation ? df * 2.0 * M_PI : df;
   const bool want_psd = !(flags & BSAHIP_NO_PSD);
   const bool want_bisp = !(flags & BSAHIP_NO_BISP);
   const size_t nm3 = (size_t)P.nm * P.nm * P.nm;
   g.last_ms = 0.0;
   g.last_launches = 0;
   if (want_bisp) {
      double *m3;
      CHK(dbuf("m3", 2 * nm3, &m3));
      CHK(fill(m3, 2 * nm3, 0.0));
      GsymArgs a;
      a.U = U;
      a.ldU = ldU;
      a.pb = D.pb_k;
      a.K = P.K;
      a.nm = P.nm;
      a.nhp = nhp;
      a.nh = nh;
      a.nu = nu;
      a.ntile = ntile;
      a.npairs = (long)ntile * (ntile + 1) / 2;
      a.H = H;
      a.nft = nft;
      a.d2 = d * d;
      a.rule = (flags & BSAHIP_QUAD_MASK) >> 4;
      a.m3f = m3;
      a.m3r = m3 + nm3;
      const long nwg = a.npairs * P.nm;
      if (nwg > 0x7fffffffL) return fail(BSAHIP_EUNSUPPORTED, "grid too large for one launch");
      if (prec == BSAHIP_PREC_BF16X3) {
         GsymArgs16 b16;
         const long ldk = P.K;                       // multiple of KC2
         const int rowsP = (int)ldU;                 // ntile * 128
         float *UR, *pb32;
         uint16_t *Ch, *Cl;
         CHK(dbuf("UR32", (size_t)rowsP * ldk, &UR));
         CHK(dbuf("UCh", (size_t)rowsP * ldk, &Ch));
         CHK(dbuf("UCl", (size_t)rowsP * ldk, &Cl));
         std::vector<float> pbf(P.pb_k.begin(), P.pb_k.end());
         CHK(upload("pb32", pbf, &pb32));
         dim3 gs((rowsP + 31) / 32, (unsigned)((ldk + 31) / 32));
         k_split_table<<<gs, 256, 0, g.s>>>(U, ldU, P.K, rowsP, ldk, UR, Ch, Cl);
         HIPC(hipGetLastError());
         b16.UR = UR;
         b16.UCh = reinterpret_cast<const uint4 *>(Ch);
         b16.UCl = reinterpret_cast<const uint4 *>(Cl);
         b16.pb32 = pb32;
         b16.ldk = ldk;
         HIPC(hipEventRecord(g.e0, g.s));
         k_gsym_bf16x3<<<dim3((unsigned)nwg), 256, 0, g.s>>>(a, b16);
      } else {
         HIPC(hipEventRecord(g.e0, g.s));
         k_gsym_f64<<<dim3((unsigned)nwg), 256, 0, g.s>>>(a);
      }
      HIPC(hipGetLastError());
      HIPC(hipEventRecord(g.e1, g.s));
      if (m3mf) HIPC(hipMemcpyAsync(m3mf, m3, nm3 * sizeof(double), hipMemcpyDeviceToHost, g.s));
      if (m3mr) HIPC(hipMemcpyAsync(m3mr, m3 + nm3, nm3 * sizeof(double), hipMemcpyDeviceToHost, g.s));
   }
   if (want_psd) {
      double *m2;
      const size_t nm2 = (size_t)P.nm * P.nm;
      CHK(dbuf("m2", 3 * nm2, &m2));
      k_m2<<<P.nm * P.nm, 256, 0, g.s>>>(U, ldU, D.pa_k, P.K, P.nm, nhp, nh, H, nft, d,
                                         2.0 * M_PI * df, m2, m2 + nm2, m2 + 2 * nm2);
      HIPC(hipGetLastError());
      if (m2mf) HIPC(hipMemcpyAsync(m2mf, m2, nm2 * sizeof(double), hipMemcpyDeviceToHost, g.s));
      if (m2mr) HIPC(hipMemcpyAsync(m2mr, m2 + nm2, nm2 * sizeof(double), hipMemcpyDeviceToHost, g.s));
      if (m2o2mr) HIPC(hipMemcpyAsync(m2o2mr, m2 + 2 * nm2, nm2 * sizeof(double), hipMemcpyDeviceToHost, g.s));
   }
   HIPC(hipStreamSynchronize(g.s));
   if (want_bisp) {
      float ms = 0.f;
      HIPC(hipEventElapsedTime(&ms, g.e0, g.e1));
      g.last_ms = ms;
      g.last_launches = 1;
   }
   return 0;
}

int bsahip__points_sum(const bsahip_problem *p, int32_t npts, const double *fi, const double *fj,
                       const double *w, int32_t psd_mode, int32_t with_frf, double *m3)
{
   std::lock_guard<std::recursive_mutex> lk(g.mu);
   CHK(ensure_init());
   if (npts < 0 || (npts > 0 && (!fi || !fj || !w || !m3)))
      return fail(BSAHIP_EINVAL, "npts < 0 or NULL point/weight/output arrays");
   Prep P;
   CHK(prepare(p, P, psd_mode));
   DevProb D;
   CHK(upload_prep(P, D));
   const size_t nm3 = (size_t)P.nm * P.nm * P.nm;
   if (npts == 0) {
      std::fill(m3, m3 + nm3, 0.0);
      return 0;
   }
   std::vector<double> all;
   all.reserve(3 * (size_t)npts);
   for (int q = 0; q < npts; ++q) {
      all.push_back(std::fabs(fi[q]));
      all.push_back(std::fabs(fj[q]));
      all.push_back(std::fabs(fi[q] + fj[q]));
   }
   std::vector<double> uq = all;
   std::sort(uq.begin(), uq.end());
   uq.erase(std::unique(uq.begin(), uq.end()), uq.end());
   auto idx = [&](double v) { return (int)(std::lower_bound(uq.begin(), uq.end(), v) - uq.begin()); };
   std::vector<int> i1(npts), i2(npts), i3(npts);
   for (int q = 0; q < npts; ++q) {
      i1[q] = idx(all[3 * q]);
      i2[q] = idx(all[3 * q + 1]);
      i3[q] = idx(all[3 * q + 2]);
   }
   const int nu = (int)uq.size();
   const long ldU = (long)P.nm * nu;
   double *U;
   CHK(build_table(P, D, uq, nu, p->pulsation, &U, ldU));
   std::vector<double> f1(fi, fi + npts), f2(fj, fj + npts), ww(w, w + npts);
   int *di1, *di2, *di3;
   double *df1, *df2, *dw, *out;
   CHK(upload("i1", i1, &di1));
   CHK(upload("i2", i2, &di2));
   CHK(upload("i3", i3, &di3));
   CHK(upload("f1", f1, &df1));
   CHK(upload("f2", f2, &df2));
   CHK(upload("wpt", ww, &dw));
   CHK(dbuf("m3sum", nm3, &out));
   CHK(fill(out, nm3, 0.0));
   const int nchunk = (int)std::min<long>(1024, std::max<long>(1, npts / 16));
   dim3 gs((unsigned)((nm3 + 255) / 256), nchunk);
   k_bfm_points_sum<<<gs, 256, 0, g.s>>>(U, ldU, nu, D.pb_k, P.K, P.nm, npts, di1, di2, di3,
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
   if (nfh < 2 || !(df > 0.0)) return fail(BSAHIP_EINVAL, "need nfh >= 2 and df > 0");
   Prep P;
   CHK(prepare(p, P));
   DevProb D;
   CHK(upload_prep(P, D));
   const int nh = nfh, nf = 2 * nh - 1, c = nh - 1;
   const bool exact = (flags & BSAHIP_SUM_EXACT) != 0;
   const int nu = exact ? nf : nh;
   const long ldU = (long)P.nm * nu;
   std::vector<double> fl(nu);
   for (int u = 0; u < nu; ++u) fl[u] = (double)u * df;
   double *U;
   CHK(build_table(P, D, fl, nu, p->pulsation, &U, ldU));
   double2 *H;
   CHK(frf_table(P, D, nf, df, &H));
   const size_t nm2 = (size_t)P.nm * P.nm, nm3 = nm2 * P.nm;
   if (psd_f || psd_r) {
      double *pf, *pr;
      CHK(dbuf("psd_f", nm2 * nf, &pf));
      CHK(dbuf("psd_r", nm2 * nf, &pr));
      dim3 gq((nf + 255) / 256, (unsigned)nm2);
      k_psd_grid<<<gq, 256, 0, g.s>>>(U, ldU, D.pa_k, P.K, P.nm, nu, nf, c, H, nf, pf, pr);
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
      CHK(upload("i1", i1, &di1));
      CHK(upload("i2", i2, &di2));
      CHK(upload("i3", i3, &di3));
      CHK(upload("f1", f1, &df1));
      CHK(upload("f2", f2, &df2));
      CHK(dbuf("bfm_out", nm3 * npts, &ob));
      CHK(dbuf("brm_out", nm3 * npts, &orr));
      const long tot = (long)nm3 * npts;
      k_bfm_points<<<(unsigned)((tot + 255) / 256), 256, 0, g.s>>>(
         U, ldU, nu, D.pb_k, P.K, P.nm, (int)npts, di1, di2, di3, df1, df2, D.Mm, D.Km, D.Cm,
         ob, orr, 1, npts);
      HIPC(hipGetLastError());
      if (bfm) HIPC(hipMemcpyAsync(bfm, ob, nm3 * npts * sizeof(double), hipMemcpyDeviceToHost, g.s));
      if (brm) HIPC(hipMemcpyAsync(brm, orr, nm3 * npts * sizeof(double), hipMemcpyDeviceToHost, g.s));
   }
   HIPC(hipStreamSynchronize(g.s));
   return 0;
}

int bsahip_bfm_points(const bsahip_problem *p, int32_t npts, const double *fi, const double *fj,
                      int32_t flags, double *bfm, double *brm)
{
   std::lock_guard<std::recursive_mutex> lk(g.mu);
   CHK(ensure_init());
   CHK(check_flags(flags, false, false));
   if (npts < 0 || (npts > 0 && (!fi || !fj || !bfm)))
      return fail(BSAHIP_EINVAL, "npts < 0 or NULL point/output arrays");
   if (npts == 0) return 0;
   Prep P;
   CHK(prepare(p, P));
   DevProb D;
   CHK(upload_prep(P, D));
   // distinct |f| over f1, f2, f1+f2 of every point
   std::vector<double> all;
   all.reserve(3 * (size_t)npts);
   for (int q = 0; q < npts; ++q) {
      all.push_back(std::fabs(fi[q]));
      all.push_back(std::fabs(fj[q]));
      all.push_back(std::fabs(fi[q] + fj[q]));
   }
   std::vector<double> uq = all;
   std::sort(uq.begin(), uq.end());
   uq.erase(std::unique(uq.begin(), uq.end()), uq.end());
   auto idx = [&](double v) { return (int)(std::lower_bound(uq.begin(), uq.end(), v) - uq.begin()); };
   std::vector<int> i1(npts), i2(npts), i3(npts);
   std::vector<double> f1(fi, fi + npts), f2(fj, fj + npts);
   for (int q = 0; q < npts; ++q) {
      i1[q] = idx(all[3 * q]);
      i2[q] = idx(all[3 * q + 1]);
      i3[q] = idx(all[3 * q + 2]);
   }
   const int nu = (int)uq.size();
   const long ldU = (long)P.nm * nu;
   double *U;
   CHK(build_table(P, D, uq, nu, p->pulsation, &U, ldU));
   int *di1, *di2, *di3;
   double *df1, *df2, *ob, *orr = nullptr;
   CHK(upload("i1", i1, &di1));
   CHK(upload("i2", i2, &di2));
   CHK(upload("i3", i3, &di3));
   CHK(upload("f1", f1, &df1));
   CHK(upload("f2", f2, &df2));
   const size_t nm3 = (size_t)P.nm * P.nm * P.nm;
   CHK(dbuf("bfm_out", nm3 * npts, &ob));
   if (brm) CHK(dbuf("brm_out", nm3 * npts, &orr));
   const long tot = (long)nm3 * npts;
   k_bfm_points<<<(unsigned)((tot + 255) / 256), 256, 0, g.s>>>(
      U, ldU, nu, D.pb_k, P.K, P.nm, npts, di1, di2, di3, df1, df2, D.Mm, D.Km, D.Cm, ob, orr,
      (long)nm3, 1);
   HIPC(hipGetLastError());
   HIPC(hipMemcpyAsync(bfm, ob, nm3 * npts * sizeof(double), hipMemcpyDeviceToHost, g.s));
   if (brm) HIPC(hipMemcpyAsync(brm, orr, nm3 * npts * sizeof(double), hipMemcpyDeviceToHost, g.s));
   HIPC(hipStreamSynchronize(g.s));
   return 0;
}

}  // extern "C"
