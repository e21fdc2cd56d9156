/*
 * bsacl.h -- the BsaCL C ABI, served by libbsahip.so on MI355X (gfx950).
 *
 * Same symbol names, argument order, argument meaning and error codes as the
 * reference's BsaCL runtime (BsaCL/include/bsacl.h:28-55, implemented in
 * BsaCL/src/bsacl.c), so its Fortran binding module BsaCL
 * (BsaCL/BsaCL_Fortran/bsacl.F90:54-205 over bsacl_c.F90:5-181) links against
 * this library unchanged.  Reals are double (the reference's __real without
 * BSA_SINGLE_FLOATING_PRECISION, BsaCL/src/_base.h:22-26).
 *
 * Like the reference, every Acquire* call stores the caller's raw pointer
 * (no copy); the arrays must stay alive until bsaclRun returns.  A failure in
 * a void function makes the runtime halt: later calls return -999 (Init) or
 * -1 (Run) or do nothing, until the process ends.
 *
 * Semantics of bsaclRun (reference bfm_kernel ids 1 and 2, BsaCL/src/bsacl.cl:
 * 179-390 and 405-930, host reduction bsacl.c:1307-1592), computed in FP64:
 *   kernel 2 (default): h_res[mnp] = dInfl * sum over all (fi, fj) pairs of the
 *     modal RESPONSE bispectrum B(m,n,p) Re[H_m(fi) H_n(fj) conj H_p(fi+fj)];
 *   kernel 1: h_res[mnp] = (h_res[mnp] + sum of the modal FORCE bispectrum) * dInfl;
 *   dInfl = dfi * dfj * 4 pi^2.  S(fi+fj) is evaluated exactly.  As the
 *   reference kernels do: turbulence length scale / std of wind zone 1
 *   (L(tc,1,1), sigma(tc,1)), nodal mean speeds u_node, coherence column 1 of
 *   nod_corr for every component clamped to DBL_MIN, the von Karman u formula
 *   (psd id 1) or Davenport (psd id 5, also when no id was acquired, as the
 *   OpenCL build forces it), and Mg[m], Kg[m], Cg[m*NM+m] indexed by
 *   effective-mode position.
 */
#ifndef BSACL_H_
#define BSACL_H_

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* error codes (BsaCL/include/_errtypes.h:22-28) */
#define BSACL_VALUE_MISMATCH_ERROR_        -1
#define BSACL_PLATFORM_FIND_ERROR_         -2
#define BSACL_DEF_PLATFORM_ACQUIRE_ERROR_  -3
#define BSACL_CQUEUES_CREATION_ERROR_      -4
#define BSACL_INVALID_DEVICE_TYPE_         -5
#define BSACL_PROGRAMS_CREATION_ERROR_     -6
#define BSACL_KERNELS_CREATION_ERROR_      -7

/* device types (BsaCL/include/_devtypes.h:21-24) */
#define BSACL_DEVICE_TYPE_CPU_ 0
#define BSACL_DEVICE_TYPE_GPU_ 1
#define BSACL_DEVICE_TYPE_ACC_ 2
#define BSACL_DEVICE_TYPE_DEF_ 3

/* bsacl.h:35 / bsacl.c:1032 -- -999 if halted; -32 thread slot failure */
int  bsaclInit(unsigned n_threads);
/* bsacl.h:36 / bsacl.c:1099 -- -83 no tc, -84 no loaded nodes, -85 no phiTc, -86 no nod_corr */
int  bsaclInitDeviceMemory(void);
/* bsacl.h:37 / bsacl.c:1307 -- -1 halted, -10 i_thread >= n_threads, -101 device memory
 * not initialised, -20 no frequencies, -123 h_res NULL */
int  bsaclRun(unsigned i_thread, const size_t dim, double *h_res);
/* bsacl.c:1597 (defined, not declared, in the reference): 1 selects kernel 1, 2 kernel 2; else returns 1 */
int  bsaclSetKernelID(unsigned kid);
/* bsacl.h:38 */
void bsaclAcquirePSDId(const uint32_t psdid);
/* bsacl.h:39 */
void bsaclAcquireStructModMat(double *modmat, double *natf, const uint32_t ndofs, const uint32_t nmodes);
/* bsacl.h:40 -- Mg[nm], Cg[nm*nm], Kg[nm] in effective-mode order */
void bsaclAcquireModalMatrices(double *Mg, double *Cg, double *Kg);
/* bsacl.h:41 */
void bsaclAcquireLoadedNodesList(int *nodes_load, const uint32_t nnodes_l);
/* bsacl.c:1687 (defined, not declared, in the reference) */
void bsaclAcquireTotalNOfNodes(const uint32_t nnodes_tot);
/* bsacl.h:42 */
void bsaclAcquireUsedModesList(int *modes, const uint32_t nmodes_eff);
/* bsacl.h:43 */
void bsaclAcquireWindCoeffs(double *wfc, const uint32_t nnodes_l, const uint32_t nlibs, const uint32_t ndegw);
/* bsacl.h:44 */
void bsaclAcquireTurbComponentsList(int *tc, const uint32_t ntc);
/* bsacl.h:45 -- phi_T_c(nmodes_eff, nnodes_l, ndegw) */
void bsaclAcquirePhiTimesCMat(double *phi_T_c, const uint32_t nmodes_eff, const uint32_t nnodes_l, const uint32_t ndegw);
/* bsacl.h:46 */
void bsaclAcquireNodalCorrelation(double *nod_corr, const uint32_t nnod_corr);
/* bsacl.h:47 */
void bsaclAcquireWindNodalVelocities(double *nod_vel);
/* bsacl.h:48 */
void bsaclAcquireWindNodalWindZones(int *nod_wz);
/* bsacl.h:49 */
void bsaclAcquireWindTurbScales(double *wt_scl, const uint32_t nwz);
/* bsacl.h:50 */
void bsaclAcquireWindTurbStd(double *wt_std, const uint32_t nwz);
/* bsacl.h:56 -- -89 before bsaclInit, -103 i_thread >= n_threads (both halt) */
void bsaclAcquireComputationFreqs(const unsigned i_thread, const uint32_t nfi, double *fi,
                                  const uint32_t nfj, double *fj);
/* bsacl.h:58 (deprecated in the reference; pointer stored, unused) */
void bsaclAcquireBaseWindTurbPSD(double *S_uvw);
/* bsacl.h:59 -- invalid type halts with -5 */
void bsaclSetDeviceType(const uint32_t itype);
/* bsacl.h:60 -- *ican = 1 when idim reals fit in free device memory, else 0 */
void bsaclVerifyMaxAllocCondition(size_t idim, uint32_t *ican);
/* bsacl.h:61 */
void bsaclAbort(const int ierr);
/* bsacl.h:62 */
void bsaclFinalise(void);

#ifdef __cplusplus
}
#endif
#endif /* BSACL_H_ */
