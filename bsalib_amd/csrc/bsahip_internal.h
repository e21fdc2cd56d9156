// Library-internal entry points shared by the C ABI translation units
// (exported from libbsahip.so but not part of the public headers).
#ifndef BSAHIP_INTERNAL_H_
#define BSAHIP_INTERNAL_H_
#include "../../include/bsahip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Classic moments; psd_mode 0 = reference Classic/Mesher PSD semantics,
 * 1 = BsaCL semantics (von Karman u formula for every component, wind zone 1,
 * coherence column 1 clamped to DBL_MIN). */
int bsahip__classic_moments(const bsahip_problem *p, int32_t nfh, double df, int32_t flags,
                            int32_t psd_mode, double *m2mf, double *m2mr, double *m2o2mr,
                            double *m3mf, double *m3mr);

/* m3[mnp] = sum_q w[q] * B(m,n,p)(fi[q], fj[q]) [* Re(H_m(fi) H_n(fj) conj H_p(fi+fj))],
 * S(f1+f2) evaluated exactly. */
int bsahip__points_sum(const bsahip_problem *p, int32_t npts, const double *fi, const double *fj,
                       const double *w, int32_t psd_mode, int32_t with_frf, double *m3);

#ifdef __cplusplus
}
#endif
#endif
