/*
 * bsahip.h -- C ABI of the MI355X (gfx950) BFM hot path of BsaLib.
 *
 * One shared library, libbsahip.so, exports these entry points (plain
 * pointers, sizes and integer error codes; callable from C, C++, Fortran
 * iso_c_binding and Python ctypes) next to the BsaCL symbol set of bsacl.h.
 *
 * Array layouts are the reference's own (Fortran column-major, 1-based ids):
 *   psi          phi_times_A_ndegw_(NM, NNL, NDEGW)   src/bsa/BsaLibImpl.F90:611-671
 *   nod_corr     ((NN*NN+NN)/2, ncorr_cols)           src/utils/utility.F90:60-79
 *   turb_scales  L(3, 3, NWZ), turb_std sigma(3, NWZ) src/wind/WindType.F90:69-209
 *   m2*          (NM*NM)  index n + (p-1)*NM
 *   m3*          (NM^3)   index m + (n-1)*NM + (p-1)*NM^2
 *   bisp         (NF, NF, NM^3)                        functionsImpl.F90:1312-1587
 *
 * Every function returns 0 on success and a negative BSAHIP_E* code on
 * failure; bsahip_last_error() returns the message of the last failure on the
 * calling thread.  A missing or unusable GPU is an error, never a fallback.
 */
#ifndef BSAHIP_H_
#define BSAHIP_H_

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* error codes */
#define BSAHIP_OK              0
#define BSAHIP_EINVAL        -1   /* bad size / pointer / flag combination  */
#define BSAHIP_EDEVICE       -2   /* HIP runtime failure                     */
#define BSAHIP_ENOMEM        -3   /* device allocation failed                */
#define BSAHIP_EUNSUPPORTED  -4   /* valid request this build does not do    */
#define BSAHIP_ESTATE        -5   /* call order violated                     */

/* precision of the node contraction (flags bits 0-3) */
#define BSAHIP_PREC_FP64       0  /* FP64 MFMA, the plain form                */
#define BSAHIP_PREC_FP32       1  /* FP32 operands, FP32 accumulate on MFMA    */
#define BSAHIP_PREC_BF16X3     2  /* split-bf16 x3 products, FP32 accumulate  */
#define BSAHIP_PREC_MASK     0xF

/* quadrature over the (f1,f2) plane (flags bits 4-7) */
#define BSAHIP_QUAD_VECT     (0 << 4) /* intgSpectraVect_: edges 1/2, corners -1/2
                                         (src/bsa/classic/BsaClassicImpl.F90:567-633) */
#define BSAHIP_QUAD_RIEMANN  (1 << 4) /* every point dw^2 (BsaClassicImpl.F90:331,341) */
#define BSAHIP_QUAD_TRAPZ    (2 << 4) /* tensor trapezoid (_shared_poly2d.fi:827-832)   */
#define BSAHIP_QUAD_MASK     (0xF << 4)

/* S(f1+f2) beyond the grid (flag bit 8) */
#define BSAHIP_SUM_TRUNCATED (0 << 8) /* zero beyond +-fmax: Classic
                                         (functionsImpl.F90:1409-1413, 1474-1477) */
#define BSAHIP_SUM_EXACT     (1 << 8) /* evaluated: Mesher / BsaCL semantics       */

/* what to compute (flag bits 9-10) */
#define BSAHIP_NO_PSD        (1 << 9)  /* skip m2 (modal PSD) outputs   */
#define BSAHIP_NO_BISP       (1 << 10) /* skip m3 (bispectrum) outputs  */

/* the ONLY_DIAG model (flag bit 11; bsahip_classic_run, bsahip_classic_bispectra_sel,
 * bsahip_mesher_create; FP64): loaded nodes
 * without coherence between them, diagonal modal entries only, every m2* and
 * m3* output holds NM values (getFM_diag_tnlm_vect_cls_ / getRM_diag_vect_cls_,
 * functionsImpl.F90:1747-2101) */
#define BSAHIP_ONLY_DIAG     (1 << 11)

/* testing (flag bit 12; Classic runs and plans): every block of the plane
 * quadrature through the generic epilogue instead of the specialised ones for
 * interior blocks; the results agree to rounding.  No reference counterpart. */
#define BSAHIP_EPI_GENERIC   (1 << 12)

/* fixed-order summation (flag bit 13; Classic runs and plans, not batches):
 * every partial sum of an m3 or m2 entry is stored instead of added
 * atomically, then each entry sums its partials in one fixed order with
 * Kahan compensation (the wave reductions already run in a fixed tree), so
 * repeated runs give bitwise-equal moments.  Costs an extra buffer of
 * 12 doubles per wave flush of the G epilogue (0.36 GB at NN=1000, NM=30,
 * 1025^2 in split-bf16x3) and a reduction pass.  Without it the moments are
 * FP64 atomic sums whose order, and last bits, vary between runs.  No
 * reference counterpart (the reference's sums are sequential). */
#define BSAHIP_DETERMINISTIC (1 << 13)

/* Problem description: borrowed host pointers, read during the call only. */
typedef struct bsahip_problem {
   int32_t nn;            /* total nodes (packed coherence index base)       */
   int32_t nnl;           /* loaded nodes                                    */
   int32_t nm;            /* effective (kept) modes                          */
   int32_t ntc;           /* turbulence components                           */
   int32_t ndegw;         /* third extent of psi (6)                         */
   int32_t nwz;           /* wind zones                                      */
   int32_t psd_id;        /* 1 von Karman, 2 Kaimal(=0), 3 Davenport-Greisch, 5 Davenport-Uliege */
   int32_t pulsation;     /* 1: PSD / 4 pi and d = 2 pi df (default convention) */
   int32_t ncorr_rows;    /* rows of nod_corr ((nn*nn+nn)/2)                  */
   int32_t corr_col_mode; /* 0: column tc per component, 1: column 1 for all   */
   int32_t pod_nmodes;    /* point BFM (bsahip_bfm_points, Mesher): project     */
                          /* through the pod_nmodes largest eigenpairs of S(f)  */
                          /* (bsa_setPODNOfModesKept); 0: no truncation         */
   int32_t reserved;      /* 0                                                  */
   double  pod_trunc_lim; /* point BFM: keep the eigenpairs >= lim * largest    */
                          /* (bsa_setPODTruncationThreshold / 100); 0: off; it  */
                          /* takes precedence over pod_nmodes.  Either needs    */
                          /* corr_col_mode 1 (the POD reads column 1).          */
   const int32_t *n_load;     /* [nnl] 1-based node ids          */
   const int32_t *tc;         /* [ntc] 1-based components        */
   const double  *psi;        /* (nm, nnl, ndegw)                */
   const double  *nod_corr;   /* (ncorr_rows, >= max tc)         */
   const double  *u_node;     /* [nn]                            */
   const int32_t *wz_node;    /* [nn] 1-based                    */
   const double  *turb_scales;/* (3, 3, nwz)                     */
   const double  *turb_std;   /* (3, nwz)                        */
   const double  *u_mean_ref; /* [nwz]                           */
   const double  *Mm;         /* [nm] modal mass, kept-mode order      */
   const double  *Km;         /* [nm] modal stiffness                  */
   const double  *Cm;         /* [nm] modal damping (diagonal of Cm)   */
} bsahip_problem;

/* Library / device. */
int         bsahip_init(int device);
int         bsahip_finalise(void);
const char *bsahip_last_error(void);
int         bsahip_version(void);

/*
 * Classic full-grid run: grid f_k = k*df, k = -(nfh-1)..(nfh-1) (NF = 2nfh-1),
 * the same one computeFreqsVect_ builds (BsaClassicImpl.F90:443-560).
 * Outputs (host pointers, any may be NULL when not wanted):
 *   m2mf, m2mr, m2o2mr  (nm*nm)     -- as mainClassic_ (BsaClassicImpl.F90:137-212)
 *   m3mf, m3mr          (nm^3)
 * Replaces getBFM_vect_cls/getBRM_vect_cls + intgSpectraVect_ inside
 * mainClassic_ (BsaClassicImpl.F90:147-210; slots in src/bsa/data/BsaLibData.F90:166-204).
 */
int bsahip_classic_run(const bsahip_problem *p, int32_t nfh, double df, int32_t flags,
                       double *m2mf, double *m2mr, double *m2o2mr,
                       double *m3mf, double *m3mr);

/*
 * bsahip_classic_run without any device allocation: same arguments, flags
 * (any precision; not BSAHIP_ONLY_DIAG) and results, with every
 * device buffer of the run carved out of a static 64 MiB array of the library
 * (one per GPU, held until the process ends, its address taken by bsahip_init,
 * which must have been called).  No reference counterpart: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc).  A problem whose buffers exceed
 * the array (roughly NTC*NNL * NM*NF*8 bytes for the A table plus the bf16x3
 * planes and the coherence tables) is rejected with BSAHIP_EINVAL and a
 * message before the contraction runs.  Every call on one GPU uses that GPU's
 * array, which keeps the last call's contents (each call clears what must
 * start at zero); calls that overlapped on one GPU would race on it, which the
 * library prevents by serialising its entry points (a process-wide mutex) and
 * returning only after the device work.
 */
int bsahip_classic_run_noalloc(const bsahip_problem *p, int32_t nfh, double df, int32_t flags,
                               double *m2mf, double *m2mr, double *m2o2mr, double *m3mf,
                               double *m3mr);

/*
 * Plans: the same Classic run with the structure resident on the device.
 * bsahip_plan_create validates and uploads the problem (the host arrays may
 * be released afterwards), computes ln rho on the device and allocates every
 * buffer of the run; bsahip_plan_run executes one full pass (nodal PSD,
 * projection, node contraction and plane quadrature, modal PSD moments) and
 * copies the moments back; outputs as bsahip_classic_run.  No reference
 * counterpart: the reference recomputes its tables inside every bsa_Run.
 */
typedef struct bsahip_plan bsahip_plan;
int bsahip_plan_create(const bsahip_problem *p, int32_t nfh, double df, int32_t flags,
                       bsahip_plan **plan);
int bsahip_plan_run(bsahip_plan *plan, double *m2mf, double *m2mr, double *m2o2mr,
                    double *m3mf, double *m3mr);
int bsahip_plan_destroy(bsahip_plan *plan);

/*
 * Bispectra on the Classic grid in the reference layout bisp(NF, NF, NM^3):
 * bfm = getFM_full_tnm_vect_cls_ output, brm = after getRM_full_vect_cls_
 * (functionsImpl.F90:1312-1738).  psd_f/psd_r: (NF, NM^2) modal PSDs.
 * Any output may be NULL.  Meant for parity checks on small grids; the
 * moments path (bsahip_classic_run) never materialises them.  Full model only
 * (BSAHIP_ONLY_DIAG: BSAHIP_EUNSUPPORTED).
 */
int bsahip_classic_spectra(const bsahip_problem *p, int32_t nfh, double df, int32_t flags,
                           double *psd_f, double *psd_r, double *bfm, double *brm);

/*
 * Bispectra of selected modal entries on the Classic grid, the data of the
 * reference's visual-mode export (src/bsa/classic/BsaClassicImpl.F90:158-194,
 * exportBRM_base_internal_ in src/bsa/BsaLibImpl.F90:1876-1913):
 * bfm, brm (NF, NF, nsel), column s = bisp(:, :, sel[s]) of the reference.
 * sel is 1-based: m + (n-1) NM + (p-1) NM^2 over the kept modes (visual_idx_,
 * BsaLibImpl.F90:387-393); with BSAHIP_ONLY_DIAG the mode position 1..NM and the
 * diagonal model (getFM_diag_tnlm_vect_cls_ / getRM_diag_vect_cls_).
 * BSAHIP_SUM_EXACT as bsahip_classic_spectra (full model only).  nsel at most
 * 65535.  Precision FP64 (default), or BSAHIP_PREC_FP32 / BSAHIP_PREC_BF16X3
 * (full model): each of the three G values of a point with the operand
 * rounding and FP32 accumulation of that precision's G kernel (summed in K
 * order; the kernel's MFMA sums in another order), for error reports of the
 * mixed-precision paths.
 */
int bsahip_classic_bispectra_sel(const bsahip_problem *p, int32_t nfh, double df, int32_t flags,
                                 int32_t nsel, const int32_t *sel, double *bfm, double *brm);

/*
 * FRF weighting of given modal spectra, in place (the reference's response
 * operators getRM_full_vect_cls_, functionsImpl.F90:1598-1738,
 * getRM_full_scalar_cls_, :2329-2450, and getRM_full_scalar_msh_, :1031-1115):
 * H_m(f) = 1 / (K_m - w^2 M_m + i w C_mm), w = 2 pi f, over the kept modes.
 *   BSAHIP_FRF_PSD:    data (n1, NM^2), column n + p NM (0-based) times
 *                      Re[H_p(f1_i) conj H_n(f1_i)]; f2, n2 unused
 *   BSAHIP_FRF_GRID:   data (n1, n2, NM^3), entry m + n NM + p NM^2 at
 *                      (f1_i, f2_j) times Re[H_m(f1_i) H_n(f2_j) conj H_p(f1_i+f2_j)]
 *   BSAHIP_FRF_POINTS: data (NM^3, n1) at the points (f1_q, f2_q): the same
 * flags BSAHIP_ONLY_DIAG: the diagonal model (getRM_diag_vect_cls_ :1975-2101,
 * getRM_diag_scalar_cls_ :2570-2637, getRM_diag_scalar_msh_ :1224-1286): NM in place of
 * NM^2 and NM^3, entry m times |H_m|^2 (PSD) or Re[H_m(f1) H_m(f2) conj H_m(f1+f2)].
 * The reference's vectorised operator keeps its column counters from one call
 * to the next (trap T3: correct on the first call only); this one has no state.
 */
#define BSAHIP_FRF_PSD    0
#define BSAHIP_FRF_GRID   1
#define BSAHIP_FRF_POINTS 2
int bsahip_frf_apply(const bsahip_problem *p, int32_t flags, int32_t kind, int64_t n1,
                     const double *f1, int64_t n2, const double *f2, double *data);

/*
 * Modal bispectra at an arbitrary list of points (fi[q], fj[q]), S(f1+f2)
 * evaluated exactly: getBFM_msh / getBRM_msh (functionsImpl.F90:247-465,
 * 1031-1115).  bfm, brm: (NM^3, npts); brm may be NULL.  flags: 0, or
 * BSAHIP_ONLY_DIAG for the diagonal model (getFM_diag_tnm_scalar_msh_ /
 * getRM_diag_scalar_msh_, :1123-1216, :1224-1286): (NM, npts).
 */
int bsahip_bfm_points(const bsahip_problem *p, int32_t npts, const double *fi,
                      const double *fj, int32_t flags, double *bfm, double *brm);

/*
 * The same at points of the Classic grid of (nfh, df) with its truncation:
 * S(|f|) = 0 beyond (nfh-1) df, which reaches f1 + f2 only (the padded S of
 * getFM_full_tnm_scalar_cls_, functionsImpl.F90:2122-2318, trap T2 read as
 * zero): the point form of the reference's scalar Classic operators.  FP64.
 * flags 0: bfm, brm (NM^3, npts); BSAHIP_ONLY_DIAG: the diagonal model
 * (getFM_diag_tnlm_scalar_cls_, :2460-2563), (NM, npts).  brm may be NULL.
 */
int bsahip_bfm_points_cls(const bsahip_problem *p, int32_t nfh, double df, int32_t flags,
                          int32_t npts, const double *fi, const double *fj, double *bfm,
                          double *brm);
/*
 * Modal force PSD of the diagonal model at arbitrary frequencies, psd (NM,
 * npts): sum over components and loaded nodes of Psi_l(m)^2 S_n(f) (the psd of
 * getFM_diag_tnlm_vect_cls_ / _scalar_cls_, functionsImpl.F90:1747-1970).
 */
int bsahip_diag_psd_points(const bsahip_problem *p, int32_t npts, const double *f, double *psd);

/*
 * Modal force PSD S_Q(n, p; f) at arbitrary frequencies: psd (NM*NM, npts),
 * index n + p*NM per point, under the problem's PSD convention and coherence
 * columns.  Its diagonal at a natural frequency is getBR_SFm_val_
 * (functionsImpl.F90:2650-2743), the input of bsa_computeBRdecomp
 * (src/bsa/BsaLibImpl.F90:1292-1421).
 */
int bsahip_modal_psd_points(const bsahip_problem *p, int32_t npts, const double *f, double *psd);

/*
 * Nodal second/third moments from modal ones (modalRecombination,
 * app/bsa.F90:935-1053): phi (ndofs, nm) column-major, m2 and m2o2 (nm*nm),
 * m3 (nm^3) in the layouts of bsahip_classic_run.  ccc = 0: unilateral
 * combination (diagonal modal terms only); 1: complete cubic combination.
 * Outputs (ndofs each): m2_r, m3_r, m2o2_r (NULL, or m2o2 given), sk_r =
 * m3_r / m2_r^(3/2) (may be NULL).
 */
int bsahip_modal_recombination(int32_t ndofs, int32_t nm, const double *phi, const double *m2,
                               const double *m3, const double *m2o2, int32_t ccc, double *m2_r,
                               double *m3_r, double *m2o2_r, double *sk_r);

/*
 * Health check of the matrix-core lane layout the kernels rely on
 * (v_mfma_f64_16x16x4f64): out512[t*64+l] / out512[256+t*64+l] receive the row /
 * column index that accumulator register t of lane l holds.  No reference
 * counterpart.
 */
int bsahip_selftest_mfma_layout(double *out512);

/*
 * Same for v_mfma_f32_16x16x32_bf16 (split-bf16 path): out768 receives three
 * 4x64 register images of products with known answers (rows, columns k<16,
 * columns k>=16).  No reference counterpart.
 */
int bsahip_selftest_mfma_bf16_layout(float *out768);

/*
 * Same for v_mfma_f32_16x16x4f32 (FP32 path): out512 receives two 4x64
 * register images of products with known answers.  No reference counterpart.
 */
int bsahip_selftest_mfma_f32_layout(float *out512);

/*
 * The library's FP64 strided-batched GEMM on host arrays (column-major):
 * C_b = alpha op(A_b) op(B_b) + beta C_b, b = 0 .. batch-1, op = transpose when
 * trans* = 1, X_b = X + b sX (sX = 0 shares one matrix).  The same kernel
 * (v_mfma_f64_16x16x4f64, 64 x 64 tiles) forms the ONLY_DIAG model's per-mode
 * T_m = S W_m S^T and the truncated-POD tables on the device.  No reference
 * counterpart.
 */
int bsahip_dgemm_strided_batched(int32_t transa, int32_t transb, int32_t M, int32_t N, int32_t K,
                                 double alpha, const double *A, int64_t lda, int64_t sA,
                                 const double *B, int64_t ldb, int64_t sB, double beta, double *C,
                                 int64_t ldc, int64_t sC, int32_t batch);

/*
 * Last measured device time (milliseconds, HIP events on the library stream)
 * of the dominant kernel of the previous call, and its launch count.
 */
int bsahip_last_kernel_time(double *ms, int32_t *launches);

/* ---- Mesher: zone quadrature with the zone data resident on the device ----
 * Zone layout (src/bsa/meshing/BsaMesherImpl.F90:266-1393, 1638-1890): a
 * background zone |f| <= b (b = bkg_peakw*bkg_area_ext/2), peak zones
 * f_n +- xsi_n f_n peak_area_ext, pre-peak gaps, external padding up to
 * max(f_n)*max_area_ext, on both axes; rectangular zones are the tensor
 * products of these intervals, cut into tiles of at most tile_cells^2
 * level-0 cells.  Each tile's force bispectrum is sampled at a refinement
 * level (1, 2, 4 or 8 cells per level-0 cell), interpolated bilinearly onto
 * the tile's BRM grid, multiplied by the FRF product and integrated with
 * trapezoid weights (zones/_shared_poly2d.fi:827-832).  S(f1+f2) is evaluated
 * exactly (functionsImpl.F90:247-465). */
typedef struct bsahip_mesher_cfg {
   int32_t bkg_base_rfmnt;    /* level-0 points per side of the background zone (reference default 20) */
   int32_t tile_cells;        /* max level-0 cells per tile side (e.g. 32) */
   int32_t diag_crest;        /* 1: BRM grids also resolve the f1+f2 = +-f_n crests
                                 (BSA_PREMESH_TYPE_DIAG_CREST_YES); reference default 0 */
   int32_t brm_cells;         /* 0: BRM spacing from the FRF (above).  > 0, a power of two
                                 <= 256: every tile's BRM spacing is at most
                                 df_ref / brm_cells, BRM cells per level-0 cell the
                                 smallest power of two that gives it (exactly brm_cells
                                 on tiles whose level-0 cell is df_ref): one uniform
                                 lattice when every level-0 cell is df_ref.  No
                                 reference counterpart. */
   double bkg_peakw;          /* background peak width [Hz] (max over L/U ratios: U/L) */
   double bkg_area_ext;       /* reference default 2 */
   double peak_area_ext;      /* reference default 2 */
   double max_area_ext;       /* reference default 2 */
   const double *natf;        /* (nm) natural frequencies of the kept modes [Hz] */
   const double *xsi;         /* (nm) damping ratios of the kept modes */
} bsahip_mesher_cfg;

typedef struct bsahip_mesher bsahip_mesher;

/* Builds the layout and evaluates the level-0 BFM of every tile (pre-mesh),
 * kept resident until bsahip_mesher_destroy.  FP64 only; at most 32 modes.
 * flags may hold BSAHIP_ONLY_DIAG: the diagonal model
 * (getFM_diag_tnm_scalar_msh_ / getRM_diag_scalar_msh_, functionsImpl.F90:1123-1286,
 * with S(f1+f2) in its products as the Classic diagonal model has them), NM
 * entries per point in every Mesher output instead of NM^3; no POD. */
int bsahip_mesher_create(const bsahip_problem *p, const bsahip_mesher_cfg *cfg, int32_t flags,
                         bsahip_mesher **out);
/* The tile layout alone (host computation, no device): *ntiles always; geom
 * and rule_level as bsahip_mesher_tiles when given (max_tiles entries). */
int bsahip_mesher_layout(const bsahip_mesher_cfg *cfg, int32_t nm, int32_t max_tiles,
                         double *geom, int32_t *rule_level, int32_t *ntiles);
int bsahip_mesher_ntiles(const bsahip_mesher *m, int32_t *ntiles);
/* geom (ntiles x 10): f1a f1b f2a f2b n1 n2 policy1 policy2 sub1 sub2 -- the tile
 * rectangle, its level-0 cells per axis, the interval policies (reference MPolicy
 * ids, 100 = background) and the BRM cells per level-0 cell per axis;
 * rule_level (ntiles): the rule-based refinement level. */
int bsahip_mesher_tiles(const bsahip_mesher *m, double *geom, int32_t *rule_level);
/* stencil (ntiles x ns^2): ||B(f1,f2)||_2 over all mode triples at the level-0
 * point nearest to each node of an ns x ns stencil over the tile. */
int bsahip_mesher_features(bsahip_mesher *m, int32_t ns, double *stencil);
/* Post-mesh: levels (ntiles, values 1/2/4/8) or NULL for the rule-based levels;
 * m3mf, m3mr (NM^3); counts[3] = {level-0 BFM evaluations, refinement BFM
 * evaluations of this run, BRM points of this run}. */
int bsahip_mesher_run(bsahip_mesher *m, const int32_t *levels, double *m3mf, double *m3mr,
                      int64_t *counts);
/* A-posteriori refinement levels for a target error tol of m3mr relative to
 * max|m3mr| (a plain adaptive Mesher: no counterpart in the reference, whose
 * levels come from the policy table alone): every tile is integrated at
 * levels 1 and 2; the error of a tile's contribution at level L is estimated
 * per entry as (c_L - c_L/2) rho / (1 - rho), rho = 1/4 (bilinear
 * interpolation is second order) until the tile has two changes, then their
 * observed ratio (clipped to [1/64, 1/2]), and the error of m3mr as the sum
 * of these estimates over the tiles; while its largest entry
 * exceeds tol max|m3mr| / 2, the tiles of largest estimate go up one level
 * (two rounds at most: levels 2, 4, 8).  levels (ntiles) receives 2, 4 or 8
 * per tile; *evals the BFM evaluations of those levels (level-0 points
 * included, each point once).  Run bsahip_mesher_run with the levels for the
 * moments. */
int bsahip_mesher_adapt(bsahip_mesher *m, double tol, int32_t *levels, int64_t *evals);
/* bsahip_mesher_adapt, then the moments and counts of bsahip_mesher_run at the
 * levels it picked, from its own last integration (each refinement round runs
 * the tiles it keeps at their current level, so no further run is needed).
 * Any of m3mf, m3mr, counts may be NULL.  No reference counterpart. */
int bsahip_mesher_adapt_run(bsahip_mesher *m, double tol, int32_t *levels, int64_t *evals,
                            double *m3mf, double *m3mr, int64_t *counts);
/* bsahip_mesher_run for n Meshers (distinct, of any structures) at once:
 * levels[i] (ntiles of Mesher i, or NULL for its rule-based levels; levels
 * itself NULL: all rule-based), m3mf[i] / m3mr[i] its outputs (either may be
 * NULL), counts (3 n, or NULL) its counts as bsahip_mesher_run's.  The BFM
 * evaluations still run per Mesher; the FRF tables and the quadrature of every
 * Mesher's tiles share one launch each (per model: full or ONLY_DIAG), which
 * pays for many small structures.  No reference counterpart. */
int bsahip_mesher_run_batch(int32_t n, bsahip_mesher *const *ms, const int32_t *const *levels,
                            double *const *m3mf, double *const *m3mr, int64_t *counts);
/* As bsahip_mesher_run, plus each tile's own m3mr contribution in tile_m3mr
 * (ntiles x NM^3; m3mr is then their sum). */
int bsahip_mesher_run_tiles(bsahip_mesher *m, const int32_t *levels, double *tile_m3mr,
                            double *m3mf, double *m3mr, int64_t *counts);
/*
 * Response bispectrum at the post-mesh (BRM) points of every tile, the data of
 * the reference's Mesher bispectrum export (exportBRM_base_internal_ called per
 * zone from src/bsa/meshing/zones/_shared_poly2d.fi:1040-2142): tile k holds a
 * dims[2k] x dims[2k+1] grid, points in order f1 fastest, tiles in layout order,
 * npts in total.  fi, fj (npts) get the point frequencies; val (ncol, npts) gets
 * val(c, pt) = sum_mnp coef(mnp, c) BRM_mnp(pt) with coef (NM^3, ncol) -- a unit
 * column selects one entry (visual mode), phi(d,m) phi(d,n) phi(d,p) gives the
 * nodal bispectrum of DOF d; coef NULL means every entry (ncol = NM^3).  The BRM
 * is the integrand of bsahip_mesher_run (interpolated BFM times the FRF
 * product).  With fi, fj and val all NULL only dims and *npts are returned; else
 * cap must hold npts.  levels as bsahip_mesher_run.  No reference counterpart
 * for the coefficient form.
 */
int bsahip_mesher_export(bsahip_mesher *m, const int32_t *levels, int32_t ncol, const double *coef,
                         int64_t cap, int32_t *dims, double *fi, double *fj, double *val,
                         int64_t *npts);

/*
 * Checkpoint of the resident zones (the reference's BFM dump file,
 * src/bsa/meshing/zones/MZone.F90:158-255, BsaMesherImpl.F90:1437-1448):
 * bsahip_mesher_save writes the zone layout and the level-0 BFM to path;
 * bsahip_mesher_load builds a Mesher for the same problem and configuration
 * from it instead of evaluating the pre-mesh.  The layout the configuration
 * gives must match the file's bit for bit (BSAHIP_EINVAL otherwise).  The file
 * is this library's own format: magic "BSAHIP-MSH-DUMP", int32 version 1, NM,
 * tiles, 0; int64 level-0 nodes per axis, level-0 evaluations; the tile
 * geometry as bsahip_mesher_tiles (real64, 10 per tile); the rule levels
 * (int32); the level-0 BFM (real64, NM^3 per point, nodes^2 points).
 */
int bsahip_mesher_save(const bsahip_mesher *m, const char *path);
int bsahip_mesher_load(const bsahip_problem *p, const bsahip_mesher_cfg *cfg, int32_t flags,
                       const char *path, bsahip_mesher **out);
int bsahip_mesher_destroy(bsahip_mesher *m);

/* ---- Learned Mesher proposer: a multilayer perceptron trained and evaluated
 * on the device (no counterpart in the reference; it proposes the per-tile
 * refinement levels that the rule-based delta validation would otherwise set).
 * depth linear layers: nin -> width -> ... -> width -> 1, GELU between them,
 * FP32 parameters; inputs are normalised as (x - mean) * inv_scale.
 * Parameter vector layout: for each layer l, W_l [out][in] then b_l [out]. */
typedef struct bsahip_mlp bsahip_mlp;
int bsahip_mlp_create(int32_t nin, int32_t width, int32_t depth, uint64_t seed, bsahip_mlp **out);
int bsahip_mlp_nparams(const bsahip_mlp *mlp, int64_t *n);
/* any of params (nparams), mean, inv_scale (nin) may be NULL */
int bsahip_mlp_set_params(bsahip_mlp *mlp, const float *params, const float *mean,
                          const float *inv_scale);
int bsahip_mlp_get_params(bsahip_mlp *mlp, float *params, float *grads);
/* logits (n) for rows x (n x nin, row-major) */
int bsahip_mlp_forward(bsahip_mlp *mlp, const float *x, int32_t n, float *logits);
/* weighted binary cross-entropy on logits and its gradient (left in the MLP,
 * copied to grads when non-NULL) */
int bsahip_mlp_loss_grad(bsahip_mlp *mlp, const float *x, const float *y, const float *w,
                         int32_t n, float *loss, float *grads);
/* minibatch Adam (learning rate lr; weight_decay * p is added to each gradient)
 * over `epochs` seeded permutations of the rows; losses (epochs) = mean batch
 * loss per epoch, may be NULL */
int bsahip_mlp_train(bsahip_mlp *mlp, const float *x, const float *y, const float *w, int32_t n,
                     int32_t epochs, int32_t batch, float lr, float weight_decay, uint64_t seed,
                     float *losses);
int bsahip_mlp_destroy(bsahip_mlp *mlp);

/* ---- Batches of structures of uneven size (no counterpart in the reference) ----
 * Each structure keeps a resident plan (as bsahip_plan_create, truncated Classic
 * sum).  A step runs every structure's tables, then ONE persistent G launch
 * whose workgroups take (structure, tile pair, mode) items from a device queue
 * ordered by decreasing item cost, then the modal PSD moments; mode 0 replays
 * that step from a HIP graph, mode 1 launches the G contraction once per
 * structure instead (the comparison point).  The item list is split into one
 * queue per XCD (workgroup blockIdx % 8) with work stealing at the end. */
typedef struct bsahip_batch bsahip_batch;
int bsahip_batch_create(const bsahip_problem *const *probs, const int32_t *nfh, const double *df,
                        int32_t n, int32_t flags, bsahip_batch **out);
/* times the persistent launch for grid sizes {1,2,4} x CUs x occupancy and
 * {1,4,16} items per queue ticket on a quarter of the queue; keeps the fastest */
int bsahip_batch_autotune(bsahip_batch *b, int32_t *grid_mult, int32_t *chunk);
/* mode 0: the captured graph of the step; 1: one G launch per structure, all on
 * the library stream; 2: the step of mode 0 launched directly (the worker
 * streams then run concurrently on the hardware queues).  Per-structure outputs
 * (arrays of n host pointers, each NM_s^3 or NM_s^2; any may be NULL). */
int bsahip_batch_run(bsahip_batch *b, int32_t mode, double *const *m3mf, double *const *m3mr,
                     double *const *m2mr);
/* device time of the last step (tables + contraction + PSD moments), queue size,
 * current launch shape */
int bsahip_batch_info(const bsahip_batch *b, double *ms_step, int64_t *items, int32_t *grid_mult,
                      int32_t *chunk);
int bsahip_batch_destroy(bsahip_batch *b);

#ifdef __cplusplus
}
#endif
#endif /* BSAHIP_H_ */
