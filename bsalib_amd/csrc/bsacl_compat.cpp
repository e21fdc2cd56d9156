// The BsaCL C ABI (include/bsacl.h) on top of libbsahip.
//
// State machine, argument checks and error codes follow the reference runtime
// BsaCL/src/bsacl.c (cited per function).  The computation of bsaclRun is the
// FP64 G-contraction of libbsahip (exact S(f1+f2), Riemann weights, BsaCL
// PSD/coherence/FRF conventions) when the computation frequencies form a
// symmetric uniform grid, and a per-point reduction otherwise.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <vector>

#include "../../include/bsacl.h"
#include "../../include/bsahip.h"
#include "bsahip_internal.h"

namespace {

// Acquired external data: raw caller pointers, no copies (bsacl.c:161-231).
struct ExtData {
   int *nodes_load = nullptr;
   unsigned NNL = 0, NN = 0, PSD_ID = 0, NLIBS = 0, NDOFS = 0, NM_EFF = 0, NDEGW = 0;
   unsigned NTC = 0, NNOD_CORR = 0, NMODES = 0, NWZ = 0;
   int *tc = nullptr, *modes_eff = nullptr, *wz = nullptr;
   double *modmat = nullptr, *natf = nullptr, *Mg = nullptr, *Cg = nullptr, *Kg = nullptr;
   double *wfc = nullptr, *phiTc = nullptr, *nod_corr = nullptr, *u_node = nullptr;
   double *wt_scl = nullptr, *wt_std = nullptr, *S_uvw = nullptr;
};

ExtData ext;
bool has_halted = false, init_called = false, devmem_called = false;
unsigned n_threads = 0;
unsigned kernel_id = 2;
unsigned device_type = BSACL_DEVICE_TYPE_DEF_;
std::vector<unsigned> nfi, nfj;
std::vector<double *> fi, fj;

// freeMem_ (bsacl.c:288-362): drop references to external memory and thread slots
void free_mem()
{
   ext = ExtData();
   nfi.clear();
   nfj.clear();
   fi.clear();
   fj.clear();
   init_called = false;
}

// abortInternal_ (bsacl.c:377-382)
void abort_internal(int ierr, const char *emsg)
{
   if (emsg && ierr != 0) std::printf("[BsaCL] ERROR: %s Aborting  (%d)\n", emsg, ierr);
   free_mem();
   devmem_called = false;
   if (ierr != 0) has_halted = true;
}

bool uniform_symmetric(const double *f, unsigned n, double &df)
{
   if (n < 3 || (n % 2) == 0) return false;
   const int c = (int)(n / 2);
   df = f[c + 1] - f[c];
   if (!(df > 0.0)) return false;
   for (unsigned k = 0; k < n; ++k)
      if (std::fabs(f[k] - (double)((int)k - c) * df) > 1e-12 * df) return false;
   return true;
}

}  // namespace

extern "C" {

int bsaclInit(unsigned n_thr)
{
   if (has_halted) return -999;
   if (n_thr == 0) {
      abort_internal(-32, "Failed to initialise thread memory.");
      return -32;
   }
   const int rc = bsahip_init(0);
   if (rc != 0) {
      std::printf("[BsaCL] ERROR: %s\n", bsahip_last_error());
      return BSACL_PLATFORM_FIND_ERROR_;
   }
   nfi.assign(n_thr, 0);
   nfj.assign(n_thr, 0);
   fi.assign(n_thr, nullptr);
   fj.assign(n_thr, nullptr);
   n_threads = n_thr;
   init_called = true;
   return 0;
}

int bsaclInitDeviceMemory(void)
{
   // bsacl.c:1099-1167: the four mandatory arrays, in this order
   if (ext.tc == nullptr) {
      abort_internal(-83, "List of turb. components was not acquired. Aborting.");
      return -83;
   }
   if (ext.nodes_load == nullptr) {
      abort_internal(-84, "List of loaded nodes was not acquired. Aborting.");
      return -84;
   }
   if (ext.phiTc == nullptr) {
      abort_internal(-85, "phiTc was not acquired. Aborting.");
      return -85;
   }
   if (ext.nod_corr == nullptr) {
      abort_internal(-86, "Nodal spatial correlation was not acquired. Aborting.");
      return -86;
   }
   devmem_called = true;
   return 0;
}

int bsaclRun(unsigned i_thread, const size_t dim, double *h_res)
{
   if (has_halted) return -1;
   int ierr = 0;
   if (!init_called) {                                   // bsacl.c:1319-1330
      n_threads = 1;
      ierr = bsaclInit(1);
      if (ierr != 0) {
         std::printf("[BsaCL] ERROR: Failed to initialise BsaCL.\n");
         return ierr;
      }
      if (!devmem_called) {
         ierr = bsaclInitDeviceMemory();
         if (ierr != 0) {
            abort_internal(ierr, "Failed to initialise BsaCL device memory.");
            return ierr;
         }
      }
   } else {                                              // bsacl.c:1331-1345
      if (i_thread >= n_threads) {
         abort_internal(-10, "Runtime has to be re-initialised due to incompatible settings.");
         return -10;
      }
      if (!devmem_called) {
         abort_internal(-101, "Device memory has not been initialised. Aborting.");
         return -101;
      }
   }
   if (i_thread >= fi.size() || fi[i_thread] == nullptr || fj[i_thread] == nullptr) {
      std::printf("[BsaCL] ERROR: No computation frequencies acquired.\n");   // bsacl.c:1356-1357
      return -20;
   }
   if (h_res == nullptr) {
      abort_internal(-123, "Host result array pointer is an invalid pointer.");
      return -123;
   }
   const unsigned nI = nfi[i_thread], nJ = nfj[i_thread];
   const double *FI = fi[i_thread], *FJ = fj[i_thread];
   if (nI < 2 || nJ < 2) {
      abort_internal(-1, "Need at least two computation frequencies per direction.");
      return -1;
   }
   const double dInfl = (FI[1] - FI[0]) * (FJ[1] - FJ[0]) * 4.0 * (M_PI * M_PI);

   // problem in libbsahip form
   const unsigned nm = ext.NM_EFF, nnl = ext.NNL;
   unsigned nn = ext.NN;
   if (nn == 0)
      for (unsigned a = 0; a < nnl; ++a) nn = std::max(nn, (unsigned)ext.nodes_load[a]);
   if (nm == 0 || nnl == 0 || ext.NTC == 0 || ext.NDEGW < 6 || !ext.Mg || !ext.Cg || !ext.Kg ||
       !ext.u_node || !ext.wt_scl || !ext.wt_std) {
      abort_internal(-1, "Incomplete acquired data (modes, nodes, components, NDEGW>=6, modal matrices, wind).");
      return -1;
   }
   std::vector<int32_t> wz1(nn, 1);
   std::vector<double> Cd(nm);
   for (unsigned m = 0; m < nm; ++m) Cd[m] = ext.Cg[(size_t)m * nm + m];
   bsahip_problem p{};
   p.nn = (int32_t)nn;
   p.nnl = (int32_t)nnl;
   p.nm = (int32_t)nm;
   p.ntc = (int32_t)ext.NTC;
   p.ndegw = (int32_t)ext.NDEGW;
   p.nwz = (int32_t)std::max(1u, ext.NWZ);
   p.psd_id = (ext.PSD_ID == 1) ? 1 : 5;
   p.pulsation = 1;
   p.ncorr_rows = (int32_t)(((long)nn * nn + nn) / 2);
   p.corr_col_mode = 1;
   p.n_load = ext.nodes_load;
   p.tc = ext.tc;
   p.psi = ext.phiTc;
   p.nod_corr = ext.nod_corr;
   p.u_node = ext.u_node;
   p.wz_node = wz1.data();
   p.turb_scales = ext.wt_scl;
   p.turb_std = ext.wt_std;
   p.u_mean_ref = nullptr;
   p.Mm = ext.Mg;
   p.Km = ext.Kg;
   p.Cm = Cd.data();

   const size_t nm3 = (size_t)nm * nm * nm;
   std::vector<double> m3f(nm3, 0.0), m3r(nm3, 0.0);
   double dfi = 0.0, dfj = 0.0;
   int rc;
   if (nI == nJ && uniform_symmetric(FI, nI, dfi) && uniform_symmetric(FJ, nJ, dfj) &&
       std::fabs(dfi - dfj) <= 1e-12 * dfi) {
      const int nfh = (int)(nI / 2) + 1;
      const int flags = BSAHIP_PREC_FP64 | BSAHIP_QUAD_RIEMANN | BSAHIP_SUM_EXACT | BSAHIP_NO_PSD;
      rc = bsahip__classic_moments(&p, nfh, dfi, flags, 1, nullptr, nullptr, nullptr,
                                   m3f.data(), m3r.data());
   } else {
      const size_t npts = (size_t)nI * nJ;
      std::vector<double> a(npts), b(npts), w(npts, dInfl);
      for (unsigned q2 = 0; q2 < nJ; ++q2)
         for (unsigned q1 = 0; q1 < nI; ++q1) {
            a[q1 + (size_t)nI * q2] = FI[q1];
            b[q1 + (size_t)nI * q2] = FJ[q2];
         }
      rc = bsahip__points_sum(&p, (int32_t)npts, a.data(), b.data(), w.data(), 1,
                              kernel_id == 2 ? 1 : 0, kernel_id == 2 ? m3r.data() : m3f.data());
   }
   if (rc != 0) {
      std::printf("[BsaCL] ERROR: %s\n", bsahip_last_error());
      abort_internal(-56, "Device computation failed.");
      return -56;
   }
   const size_t nout = std::min(dim, nm3);
   if (kernel_id == 2) {
      for (size_t i = 0; i < nout; ++i) h_res[i] = m3r[i];
      for (size_t i = nout; i < dim; ++i) h_res[i] = 0.0;
   } else {
      for (size_t i = 0; i < nout; ++i) h_res[i] = h_res[i] * dInfl + m3f[i];
      for (size_t i = nout; i < dim; ++i) h_res[i] *= dInfl;
   }
   return 0;
}

int bsaclSetKernelID(unsigned kid)
{
   if (kid < 1 || kid > 2) return 1;                     // bsacl.c:1597-1608
   kernel_id = kid;
   return 0;
}

void bsaclAcquirePSDId(const uint32_t psdid)
{
   if (has_halted) return;
   ext.PSD_ID = psdid;
}

void bsaclAcquireStructModMat(double *modmat, double *natf, const uint32_t ndofs, const uint32_t nmodes)
{
   if (has_halted) return;
   ext.modmat = modmat;
   ext.natf = natf;
   ext.NDOFS = ndofs;
   ext.NMODES = nmodes;
}

void bsaclAcquireModalMatrices(double *Mg, double *Cg, double *Kg)
{
   if (has_halted) return;
   ext.Mg = Mg;
   ext.Cg = Cg;
   ext.Kg = Kg;
}

void bsaclAcquireLoadedNodesList(int *nodes_load, const uint32_t nnodes_l)
{
   if (has_halted) return;
   if (ext.NNL != 0 && nnodes_l != ext.NNL) {
      abort_internal(BSACL_VALUE_MISMATCH_ERROR_, "N. of loaded nodes does not match.");
      return;
   }
   ext.nodes_load = nodes_load;
   ext.NNL = nnodes_l;
}

void bsaclAcquireTotalNOfNodes(const uint32_t nnodes_tot)
{
   if (has_halted) return;
   if (ext.NN != 0 && nnodes_tot != ext.NN) {
      abort_internal(BSACL_VALUE_MISMATCH_ERROR_, "N. of total nodes does not match.");
      return;
   }
   ext.NN = nnodes_tot;
}

void bsaclAcquireUsedModesList(int *modes, const uint32_t nmodes_eff)
{
   if (has_halted) return;
   if (ext.NM_EFF != 0 && nmodes_eff != ext.NM_EFF) {
      abort_internal(BSACL_VALUE_MISMATCH_ERROR_, "N. of effective modes does not match.");
      return;
   }
   ext.modes_eff = modes;
   ext.NM_EFF = nmodes_eff;
}

void bsaclAcquireWindCoeffs(double *wfc, const uint32_t nnodes_l, const uint32_t nlibs, const uint32_t ndegw)
{
   if (has_halted) return;
   if (ext.NNL != 0 && nnodes_l != ext.NNL) {
      abort_internal(BSACL_VALUE_MISMATCH_ERROR_, "N. of loaded nodes does not match.");
      return;
   }
   if (ext.NLIBS != 0 && nlibs != ext.NLIBS) {
      abort_internal(BSACL_VALUE_MISMATCH_ERROR_, "N. of LIBS does not match.");
      return;
   }
   if (ext.NDEGW != 0 && ndegw != ext.NDEGW) {
      abort_internal(BSACL_VALUE_MISMATCH_ERROR_, "DEGW does not match.");
      return;
   }
   ext.NNL = nnodes_l;
   ext.NLIBS = nlibs;
   ext.NDEGW = ndegw;
   ext.wfc = wfc;
}

void bsaclAcquireTurbComponentsList(int *tc, const uint32_t ntc)
{
   if (has_halted) return;
   ext.NTC = ntc;
   ext.tc = tc;
}

void bsaclAcquirePhiTimesCMat(double *phi_T_c, const uint32_t nmodes_eff, const uint32_t nnodes_l, const uint32_t ndegw)
{
   if (has_halted) return;
   if (ext.NNL != 0 && nnodes_l != ext.NNL) {
      abort_internal(BSACL_VALUE_MISMATCH_ERROR_, "N. of loaded nodes does not match.");
      return;
   }
   if (ext.NM_EFF != 0 && nmodes_eff != ext.NM_EFF) {
      abort_internal(BSACL_VALUE_MISMATCH_ERROR_, "N. of effective modes does not match.");
      return;
   }
   if (ext.NDEGW != 0 && ndegw != ext.NDEGW) {
      abort_internal(BSACL_VALUE_MISMATCH_ERROR_, "DEGW does not match.");
      return;
   }
   ext.NNL = nnodes_l;
   ext.NM_EFF = nmodes_eff;
   ext.NDEGW = ndegw;
   ext.phiTc = phi_T_c;
}

void bsaclAcquireNodalCorrelation(double *nod_corr, const uint32_t nnod_corr)
{
   if (has_halted) return;
   ext.nod_corr = nod_corr;
   ext.NNOD_CORR = nnod_corr;
}

void bsaclAcquireWindNodalVelocities(double *nod_vel)
{
   if (has_halted) return;
   ext.u_node = nod_vel;
}

void bsaclAcquireWindNodalWindZones(int *nod_wz)
{
   if (has_halted) return;
   ext.wz = nod_wz;
}

void bsaclAcquireWindTurbScales(double *wt_scl, const uint32_t nwz)
{
   if (has_halted) return;
   ext.NWZ = nwz;
   ext.wt_scl = wt_scl;
}

void bsaclAcquireWindTurbStd(double *wt_std, const uint32_t nwz)
{
   if (has_halted) return;
   ext.NWZ = nwz;
   ext.wt_std = wt_std;
}

void bsaclAcquireComputationFreqs(const unsigned i_thread, const uint32_t nfi_, double *fi_,
                                  const uint32_t nfj_, double *fj_)
{
   if (has_halted) return;
   if (!init_called) {                                   // bsacl.c:1857-1864
      abort_internal(-89, "This function must be called after BsaCL initialisation.");
      return;
   }
   if (n_threads <= i_thread) {
      abort_internal(-103, "Current thread index exceeds max n. of threads.");
      return;
   }
   nfi[i_thread] = nfi_;
   fi[i_thread] = fi_;
   nfj[i_thread] = nfj_;
   fj[i_thread] = fj_;
}

void bsaclAcquireBaseWindTurbPSD(double *S_uvw)
{
   if (has_halted) return;
   ext.S_uvw = S_uvw;
}

void bsaclSetDeviceType(const uint32_t itype)
{
   if (has_halted) return;
   if (itype > BSACL_DEVICE_TYPE_DEF_) {                 // bsacl.c:1912-1929
      abort_internal(BSACL_INVALID_DEVICE_TYPE_, "Invalid device type requested.");
      return;
   }
   device_type = itype;
}

void bsaclVerifyMaxAllocCondition(size_t idim, uint32_t *ican)
{
   if (has_halted || ican == nullptr) return;
   size_t free_b = 0, total_b = 0;
   if (bsahip_init(0) != 0 || hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
      *ican = 0U;
      return;
   }
   *ican = (idim <= free_b / sizeof(double)) ? 1U : 0U;
}

void bsaclAbort(const int ierr) { abort_internal(ierr, nullptr); }

void bsaclFinalise(void)
{
   // bsacl.c:1975-2012: release device resources, then abortInternal_(iret)
   abort_internal(0, nullptr);
}

}  // extern "C"
