!! BsaLib front for the GPU BFM path: module BsaLib with the reference's
!! bsa_Init / bsa_Run / bsa_Finalise and the setters that feed the Classic
!! third-moment computation, with the reference's names, argument meanings and
!! abort behaviour (src/BsaLib.F90, src/bsa/BsaLibImpl.F90).  All O(NF^2)
!! work runs in libbsahip.so on the GPU (bsahip_classic_run); this module only
!! collects the inputs, does the reference's O(NDOFS*NM) host preparation
!! (mode validation, Phi*C projection, frequency-grid checks) and hands plain
!! arrays to the C ABI.  There is no CPU path: without a GPU bsa_Run aborts.
!!
!! Differences from the reference, all deliberate:
!!  * Inputs are copied at the setter call (the reference keeps pointers to
!!    the caller's arrays); L(3,3,*) and sigma(3,*) are assumed-size in the
!!    reference interface, so those two are kept as addresses and copied at
!!    bsa_Run once the number of wind zones is known.
!!  * The number of wind zones is taken from bsa_setWindZoneLimits when it
!!    was called (as the reference), else from bsa_setWZMeanWindVel, else
!!    from max(NodWZ).
!!  * Analysis type 1 runs the Classic approach, 2 the Mesher, 3 both (as
!!    BsaLibImpl.F90:412-452).  bsa_enableGPU forces Classic, as the
!!    reference's GPU path does (BsaLibImpl.F90:347-353); unlike that path,
!!    PSDs are still computed when requested.  The Mesher is the device zone
!!    Mesher of libbsahip (zone limits and policies of BsaMesherImpl.F90,
!!    tiles of at most 32 x 32 level-0 cells); its refinement levels are
!!    chosen a posteriori for a target error of m3mr (bsa_setMesherTolerance,
!!    an addition; default 1e-3 of max|m3mr|; 0 selects the rule-based levels
!!    of the policy factors); its results are m3mf_msh, m3mr_msh.  With isvd = 1 its BFM is the POD
!!    one; without truncation that is evaluated as the direct projection
!!    with coherence column 1 (equal up to rounding), with truncation through
!!    eigendecompositions on the GPU.  A threshold counts the modes as the
!!    reference's dgesvd build does (largest first); its default dsyev build
!!    scans its ascending eigenvalues from the small end and so keeps a single
!!    mode unless the second-smallest eigenvalue passes the threshold.
!!  * bsa_setComputePrecision is an addition: BSA_PREC_BF16X3 selects the
!!    split-bf16x3 node contraction and BSA_PREC_FP32 the FP32 one (FP32
!!    accumulation on matrix cores, FP64 everywhere else); the default is FP64.
!!  * Bispectrum export files: visual mode writes one entry (or
!!    the nodal bispectrum of one DOF) at the Mesher's device tiles without an
!!    earlier run's BFM dump file, which the reference's visual mode re-reads.
!!  * Reference defects that are not reproduced: the frequency convention
!!    (its vectorised Classic BFM indexes S(f1+f2) as if the grid were
!!    centred, functionsImpl.F90:1307-1309) stops bsa_Run with a message; the
!!    Mesher's diagonal model uses S(f1+f2) in its second and third products
!!    (functionsImpl.F90:1161-1163 uses S(f1) S(f2)^2 and S(f1)^2 S(f2)).
module BsaLib
   use, intrinsic :: iso_fortran_env, only: int32, int64, real32, real64, output_unit
   use, intrinsic :: iso_c_binding
   use BsaHip_C
   implicit none
   private

   integer, parameter, public :: bsa_int_t = int32
   integer, parameter, public :: bsa_real_t = real64
   ! kind names and constants of the reference's module (src/CONSTANTS/_CONSTANTS.F90)
   integer, parameter, public :: IK = int32, RK = real64
   integer(bsa_int_t), parameter, public :: BSA_PREMESH_MODE_BASE = 0
   integer(bsa_int_t), parameter, public :: BSA_PREMESH_MODE_ZONE_REFINED = 1
   character(len=*), parameter, public :: BSA_OUT_DIRNAME_DEFAULT = '.\bsaresults\'
   character(len=*), parameter, public :: BSA_OUT_FILENAME_PREFIX_DEFAULT_ = 'bsaout_def'
   character(len=*), parameter, public :: BSA_EXPORT_M2MF_CLS_FNAME = 'm2mf_cls'
   character(len=*), parameter, public :: BSA_EXPORT_M2MR_CLS_FNAME = 'm2mr_cls'
   character(len=*), parameter, public :: BSA_EXPORT_M2O2MR_CLS_FNAME = 'm2o2mr_cls'
   character(len=*), parameter, public :: BSA_EXPORT_M3MF_CLS_FNAME = 'm3mf_cls'
   character(len=*), parameter, public :: BSA_EXPORT_M3MR_CLS_FNAME = 'm3mr_cls'
   character(len=*), parameter, public :: BSA_EXPORT_M2MF_MSH_FNAME = 'm2mf_msh'
   character(len=*), parameter, public :: BSA_EXPORT_M2MR_MSH_FNAME = 'm2mr_msh'
   character(len=*), parameter, public :: BSA_EXPORT_M2O2MR_MSH_FNAME = 'm2o2mr_msh'
   character(len=*), parameter, public :: BSA_EXPORT_M3MF_MSH_FNAME = 'm3mf_msh'
   character(len=*), parameter, public :: BSA_EXPORT_M3MR_MSH_FNAME = 'm3mr_msh'
   character(len=*), parameter, public :: IO_ACCESS_DIRECT = 'DIRECT', IO_ACCESS_SEQUEN = 'SEQUENTIAL'
   character(len=*), parameter, public :: IO_ACCESS_STREAM = 'STREAM', IO_ACCESS_APPEND = 'APPEND'
   character(len=*), parameter, public :: IO_ACTION_READWRITE = 'READWRITE'
   character(len=*), parameter, public :: IO_ASYNC_YES = 'YES', IO_ASYNC_NO = 'NO'
   character(len=*), parameter, public :: IO_FORM_FORMATTED = 'FORMATTED'
   character(len=*), parameter, public :: IO_FORM_UNFORMATTED = 'UNFORMATTED', IO_FORM_BINARY = 'BINARY'
   character(len=*), parameter, public :: IO_POSITION_ASIS = 'ASIS', IO_POSITION_REWIND = 'REWIND'
   character(len=*), parameter, public :: IO_POSITION_APPEND = 'APPEND'
   character(len=*), parameter, public :: IO_STATUS_OLD = 'OLD', IO_STATUS_NEW = 'NEW'
   character(len=*), parameter, public :: IO_STATUS_SCRATCH = 'SCRATCH', IO_STATUS_REPLACE = 'REPLACE'
   character(len=*), parameter, public :: IO_STATUS_UNKNOWN = 'UNKNOWN'
   character(len=*), parameter, public :: NOTEMSG = '  --[note]   '
   real(RK), parameter, public :: MACHINE_PRECISION = 1e-12_RK
   real(RK), parameter, public :: CST_PIGREC = 4.0_RK * atan(1.0_RK)
   real(RK), parameter, public :: CST_PIt2 = CST_PIGREC * 2.0_RK, CST_PIt4 = CST_PIGREC * 4.0_RK
   real(RK), parameter, public :: CST_PId2 = CST_PIGREC / 2.0_RK, CST_PId4 = CST_PIGREC / 4.0_RK
   real(RK), parameter, public :: CST_2d3 = 2.0_RK / 3.0_RK
   real(RK), parameter, public :: CST_PIt3d2 = CST_PIGREC * 1.5_RK

   ! user callback for bispectrum exports (src/CONSTANTS/_CONSTANTS.F90:171-185)
   abstract interface
      subroutine exportInterf_vect_(f1, f2, brm, pdata)
         import :: bsa_real_t
         real(bsa_real_t), intent(in) :: f1(:)
         real(bsa_real_t), intent(in) :: f2(:)
         real(bsa_real_t), intent(in) :: brm(:, :)
         class(*), pointer, intent(in) :: pdata
      end subroutine
   end interface
   public :: exportInterf_vect_
   ! re-exported, as the reference module (which is public) re-exports them
   public :: int32, real64

   integer(bsa_int_t), parameter, public :: BSA_SPATIAL_SYM_NONE = 0
   integer(bsa_int_t), parameter, public :: BSA_SPATIAL_SYM_HALF = 2
   integer(bsa_int_t), parameter, public :: BSA_SPATIAL_SYM_FOUR = 4
   integer(bsa_int_t), parameter, public :: BSA_CLASSIC_MODE_VECTOR = 0
   integer(bsa_int_t), parameter, public :: BSA_CLASSIC_MODE_SCALAR = 1
   integer(bsa_int_t), parameter, public :: BSA_PSD_CONVENTION_FREQ = 0
   integer(bsa_int_t), parameter, public :: BSA_PSD_CONVENTION_PULS = 1
   integer(bsa_int_t), parameter, public :: BSA_WIND_PSD_VONKARMAN = 1
   integer(bsa_int_t), parameter, public :: BSA_WIND_PSD_KAIMAL = 2
   integer(bsa_int_t), parameter, public :: BSA_WIND_PSD_DAVENPORT = 5
   integer(bsa_int_t), parameter, public :: BSA_PREC_FP64 = 0
   integer(bsa_int_t), parameter, public :: BSA_PREC_FP32 = 1
   integer(bsa_int_t), parameter, public :: BSA_PREC_BF16X3 = 2

   integer(bsa_int_t), parameter, public :: BSA_EXPORT_FORMAT_FORMATTED = 0
   integer(bsa_int_t), parameter, public :: BSA_EXPORT_FORMAT_UNFORMATTED = 1
   integer(bsa_int_t), parameter, public :: BSA_EXPORT_MODE_REPLACE = 0
   integer(bsa_int_t), parameter, public :: BSA_EXPORT_MODE_APPEND = 1
   integer(bsa_int_t), parameter, public :: BSA_EXPORT_BRM_MODE_NONE = 0
   integer(bsa_int_t), parameter, public :: BSA_EXPORT_BRM_MODE_BASE = 1
   integer(bsa_int_t), parameter, public :: BSA_EXPORT_BRM_MODE_USR = 9
   integer(bsa_int_t), parameter, public :: BSA_WIND_VERT_PROFILE_POWER = 1
   integer(bsa_int_t), parameter, public :: BSA_WIND_VERT_PROFILE_LOG = 2
   ! input file names and the constants programs written against the reference use
   character(len=*), parameter, public :: BSA_DATA_FNAME = 'bsa.bsadata'
   character(len=*), parameter, public :: EXT_DATA_FNAME = 'bsa.extdata'
   ! bispectrum export files (src/CONSTANTS/_CONSTANTS.F90:165-168)
   character(len=*), parameter, public :: BRM_EXPORT_FNAME_CLS = 'bsaexport_cls.brm'
   character(len=*), parameter, public :: BRN_EXPORT_FNAME_CLS = 'bsaexport_cls.brn'
   character(len=*), parameter, public :: BRM_EXPORT_FNAME_MSH = 'bsaexport_msh.brm'
   character(len=*), parameter, public :: BRN_EXPORT_FNAME_MSH = 'bsaexport_msh.brn'
   character(len=0), parameter, public :: BSA_FILE_NAME_CL_SUFFIX = ''
   character(len=*), parameter, public :: IO_ACTION_READ = 'READ'
   character(len=*), parameter, public :: IO_ACTION_WRITE = 'WRITE'
   real(bsa_real_t), parameter, public :: CST_3d2 = 3.0_bsa_real_t / 2.0_bsa_real_t

   character(len=*), parameter, public :: INFOMSG = '  --[info]   '
   character(len=*), parameter, public :: WARNMSG = '  --[warn]   '
   character(len=*), parameter, public :: ERRMSG = '  --[error]  '
   character(len=*), parameter, public :: MSGCONT = '             '
   character(len=*), parameter, public :: DBGMSG = '  --[debug]  '

   public :: bsa_Init, bsa_Run, bsa_Finalise, bsa_isCleaned, bsa_Abort
   public :: bsa_enableGPU, bsa_setAnalysisType, bsa_setClassicMode, bsa_setScalingConv
   public :: bsa_setSpectraComputation, bsa_setSpectraExtension, bsa_setTestMode
   public :: bsa_setSpatialSymmetry, bsa_doValidateModalData, bsa_setupClassic
   public :: bsa_setWindDirections, bsa_setWindTurbComps, bsa_setPSDType, bsa_setWindZoneLimits
   public :: bsa_setWZMeanWindVel, bsa_setTurbWindScales, bsa_setTurbWindSDT
   public :: bsa_setNodalVel, bsa_setNodalWindZones, bsa_setSpatialNodalCorr
   public :: bsa_setWindFCoeffs, bsa_setPhitimesC
   public :: bsa_setTotalNumOfDOFs, bsa_setNumOfNodalDOFs, bsa_setTotalNOfNodes
   public :: bsa_setLoadedNodalDOFs, bsa_setLoadedNodes, bsa_setModalInfo
   public :: bsa_setKeptModalShapes, bsa_setModalMatrices, bsa_setTotDamping
   public :: bsa_getUsedModeShapes, bsa_setComputePrecision
   public :: bsa_setupMesher, bsa_setPremeshScheme, bsa_setMaxBkgPeakRestriction
   public :: bsa_setPODTruncationThreshold, bsa_setPODNOfModesKept
   public :: bsa_computeBRdecomp, bsa_computePeakFactors, bsa_modalRecombination
   ! settings that do not change the device path, kept for the input files
   public :: bsa_setVersion, bsa_setSpectraSymmetries, bsa_setNodalCoords, bsa_setWindVertProf
   public :: bsa_setWindAltDir, bsa_setAirDensity, bsa_setGlobalRotMatW2G, bsa_setWZRefAlt
   public :: bsa_setWindCorrCoeffs, bsa_setWindCorrExpnts, bsa_setWZRotMatW2G
   public :: bsa_setIncidenceAngles, bsa_setNodalWindAltitudes, bsa_setBfmMLR
   public :: bsa_doValidateZoneDeltas, bsa_forceBsaClsExecution, bsa_generateBSAInputFiles
   ! exports and run-time I/O settings
   public :: bsa_printBSAHeader, bsa_setExportInCurrDir, bsa_setExportDirectory
   public :: bsa_closeUnitsAtEnd, bsa_setExportFileFormat, bsa_setExportAppendMode
   public :: bsa_setOutFileName, bsa_setBRMExportDefaultMode, bsa_enableVisualMode
   public :: bsa_setVisualIndexes, bsa_exportMomentToFile, bsa_exportSkewness
   public :: bsa_exportBRdecomp, bsa_exportExtremeValuesToFile, bsa_saveCoordinatesToFile
   public :: bsa_setPremeshMode, bsa_enableOnlyPremesh, bsa_setPolicyIDValidationValues
   public :: bsa_setMesherTolerance
   public :: bsa_isFullComp, bsa_setOutputDirectory, bsa_setOutUnit, bsa_exportPSDToFile
   public :: bsa_exportBispToFile, bsa_exportModalData, bsa_setBispExportCallback

   interface bsa_exportSkewness
      module procedure exportSkewnessCompute_, exportSkewnessGiven_
   end interface
   interface bsa_exportBRdecomp
      module procedure exportBR_
   end interface
   integer(bsa_int_t), parameter, public :: BSA_PREMESH_TYPE_DIAG_CREST_NO = 0
   integer(bsa_int_t), parameter, public :: BSA_PREMESH_TYPE_DIAG_CREST_YES = 1

   ! ---- settings (defaults as src/settings/Settings.F90:33-84) ----
   integer(bsa_int_t) :: isuban_ = 2, iclassic_ = 0, iconv_ = 1
   integer(bsa_int_t) :: icomp_psd_ = 1, icomp_bisp_ = 1, ionly_diag_ = 0, itest_ = 0
   integer(bsa_int_t) :: isym_ = 0, nfreqs_ = 0, iprec_ = BSA_PREC_FP64
   real(bsa_real_t) :: df_ = 0
   ! Mesher settings (src/settings/Settings.F90:86-116)
   integer(bsa_int_t) :: bkgrfmt_ = 20, premesh_type_ = 0, isvd_ = 1
   real(bsa_real_t) :: bkgaext_ = 2, genpaext_ = 2, maxaext_ = 2
   logical :: restrict_bkg_ = .false.
   ! POD truncation (src/bsa/data/BsaLibData.F90:276-285)
   logical :: do_trunc_pod_ = .false., npod_set_ = .false.
   real(bsa_real_t) :: pod_trunc_lim_ = 0
   integer(bsa_int_t) :: npod_ = 0
   logical :: validate_modal_ = .true., gpu_enabled_ = .false.
   logical :: cleaned_ = .true., device_ready_ = .false.

   ! ---- structure ----
   integer(bsa_int_t) :: nn_ = 0, nlibs_ = 0, ndofs_ = 0, nm_ = 0
   integer(bsa_int_t), allocatable, target :: n_load_(:), libs_load_(:), modes_user_(:), modes_(:)
   real(bsa_real_t), allocatable :: phi_(:, :), natf_(:), xsi_(:), Cg_(:, :)
   real(bsa_real_t), allocatable :: Mg_(:), Kg_(:)

   ! ---- wind ----
   integer(bsa_int_t) :: ipsd_ = 0, nz_lim_ = 0
   integer(bsa_int_t), allocatable, target :: tc_(:), dirs_(:), wz_node_(:)
   real(bsa_real_t), allocatable, target :: u_node_(:), umref_(:), corr_(:, :), wfc_(:, :, :)
   real(bsa_real_t), allocatable, target :: psi_user_(:, :, :)
   type(c_ptr) :: tscl_addr_ = c_null_ptr, tstd_addr_ = c_null_ptr

   ! ---- inputs the device path does not read, kept for bsa_generateBSAInputFiles
   !      and the coordinates export (defaults as src/settings/Settings.F90) ----
   integer(bsa_int_t) :: ivers_ = 2, ispctrsym_ = 0, iwprof_ = BSA_WIND_VERT_PROFILE_POWER
   integer(bsa_int_t) :: ivert_ = 3, ifcov_ = 1, idumpmod_ = 0
   real(bsa_real_t) :: aird_ = 1.25_bsa_real_t
   real(bsa_real_t) :: rotW2G_(3, 3) = reshape([1, 0, 0, 0, 1, 0, 0, 0, 1], [3, 3]) * 1.0_bsa_real_t
   real(bsa_real_t), allocatable :: coords_(:, :), zref_(:), incang_(:), walt_(:), wz_lims_(:)
   type(c_ptr) :: ccoef_addr_ = c_null_ptr, cexp_addr_ = c_null_ptr, rotlw_addr_ = c_null_ptr
   logical :: gen_inputs_ = .false., run_after_gen_ = .true.

   ! ---- exports ----
   character(len=:), allocatable :: exp_dir_
   logical :: exp_binary_ = .false., exp_append_ = .false., visual_ = .false.
   ! visual mode: indexes and modal (BRM) or nodal (BRN) export
   ! (src/bsa/data/BsaLibData.F90:87-100)
   integer(bsa_int_t) :: vis_idx_(3) = 0
   logical :: vis_modal_ = .true.
   ! bispectrum export of Mesher runs (bsa_setBRMExportDefaultMode), a user
   ! callback in place of the export files (bsa_setBispExportCallback)
   logical :: brm_base_ = .false.
   procedure(exportInterf_vect_), pointer :: brm_cb_ => null()
   class(*), pointer :: null_pdata_ => null()
   ! Mesher pre-mesh only, pre-mesh mode, modal data file at bsa_Run
   logical :: only_premesh_ = .false., export_modal_ = .false.
   ! Mesher refinement: target error of m3mr relative to its maximum (0: the
   ! rule-based levels); levels of the last run, used by its exports
   real(bsa_real_t) :: mesher_tol_ = 1.0e-3_bsa_real_t
   integer(c_int32_t), allocatable, target :: msh_levels_(:)
   integer(bsa_int_t) :: premesh_mode_ = BSA_PREMESH_MODE_ZONE_REFINED

   ! ---- problem of the last bsa_Run, kept for the post-processing calls ----
   type(bsahip_problem_t), save :: lp_
   logical :: lp_ready_ = .false.
   real(bsa_real_t), allocatable, target :: lp_psi_(:, :, :), lp_Mk_(:), lp_Kk_(:), lp_Ck_(:)
   real(bsa_real_t), allocatable, target :: lp_tscl_(:, :, :), lp_tstd_(:, :), lp_umref_(:)
   integer(bsa_int_t), allocatable, target :: lp_tcs_(:)
   ! the Mesher's problem (POD settings), Classic grid of the last Classic run
   type(bsahip_problem_t), save :: lp_msh_
   integer(bsa_int_t) :: lp_nfh_ = 0
   real(bsa_real_t) :: lp_df_ = 0
   logical :: lp_diag_ = .false.

   ! ---- operator slots (src/bsa/data/BsaLibData.F90:166-310) ----
   ! The reference's spectral operators behind procedure pointers, bound by
   ! bsa_Run (BsaLibImpl.F90:313-372): the Mesher ones always, the Classic ones
   ! once a Classic run has fixed the grid, for the model of that run (full:
   ! NM^2 PSD and NM^3 bispectrum entries; ONLY_DIAG: NM of each).  Each
   ! evaluates on the device from the problem of the last bsa_Run: the nodal
   ! wind PSD arguments (Suvw, Suvw_pad) are taken for the interface and
   ! checked for shape only, and the scalar BFM is not pure (it calls the
   ! library).  Thread-safe: the library serialises calls.
   abstract interface
      subroutine getBFMClsVect(f, Suvw, psd, bisp)
         import :: bsa_real_t
         real(bsa_real_t), intent(in) :: f(:)
         real(bsa_real_t), intent(in) :: Suvw(:, :)
         real(bsa_real_t), allocatable, intent(inout) :: psd(:, :), bisp(:, :, :)
      end subroutine

      subroutine getBRMClsVect(f, psd, bisp)
         import :: bsa_real_t
         real(bsa_real_t), intent(in) :: f(:)
         real(bsa_real_t), allocatable, intent(inout) :: psd(:, :), bisp(:, :, :)
      end subroutine

      subroutine getBFMClsScalar(ii, ij, fi, fj, Suvw, Suvw_pad, psd, bisp)
         import :: bsa_int_t, bsa_real_t
         integer(bsa_int_t), intent(in) :: ii, ij
         real(bsa_real_t), intent(in) :: fi, fj
         real(bsa_real_t), intent(in) :: Suvw(:, :), Suvw_pad(:)
         real(bsa_real_t), intent(inout) :: psd(:), bisp(:)
      end subroutine

      subroutine getBRMClsScalar(ii, ij, fi, fj, psdin, psdout, bispin, bispout)
         import :: bsa_int_t, bsa_real_t
         integer(bsa_int_t), intent(in) :: ii, ij
         real(bsa_real_t), intent(in) :: fi, fj
         real(bsa_real_t), intent(in) :: psdin(:), bispin(:)
         real(bsa_real_t), intent(out) :: psdout(:), bispout(:)
      end subroutine

      subroutine getMshBFM(bfm, fi, fj)
         import :: bsa_real_t
         real(bsa_real_t), intent(inout), contiguous :: bfm(:, :)
         real(bsa_real_t), intent(in), contiguous :: fi(:), fj(:)
      end subroutine

      subroutine getMshBRM(brm, fi, fj, bfm)
         import :: bsa_real_t
         real(bsa_real_t), intent(inout), contiguous :: brm(:, :)
         real(bsa_real_t), intent(in), contiguous :: fi(:), fj(:)
         real(bsa_real_t), intent(in), contiguous :: bfm(:, :)
      end subroutine
   end interface
   public :: getBFMClsVect, getBRMClsVect, getBFMClsScalar, getBRMClsScalar, getMshBFM, getMshBRM
   procedure(getBFMClsVect), pointer, public :: getBFM_vect_cls => null()
   procedure(getBRMClsVect), pointer, public :: getBRM_vect_cls => null()
   procedure(getBFMClsScalar), pointer, public :: getBFM_scalar_cls => null()
   procedure(getBRMClsScalar), pointer, public :: getBRM_scalar_cls => null()
   procedure(getMshBFM), pointer, public :: getBFM_msh => null()
   procedure(getMshBRM), pointer, public :: getBRM_msh => null()

contains

   ! ======================================================================
   !  general
   ! ======================================================================

   !> Prints emsg and stops (src/bsa/data/BsaLibDataImpl.F90:132).
   subroutine bsa_Abort(emsg)
      character(len=*), intent(in), optional :: emsg
      if (present(emsg)) then
         write(output_unit, '(1x, 2a)') ERRMSG, emsg
      else
         write(output_unit, '(1x, 2a)') ERRMSG, 'BsaLib aborted.'
      end if
      flush(output_unit)
      error stop 1
   end subroutine

   !> src/bsa/BsaLibImpl.F90:61-87 (allocates the global singletons there).
   subroutine bsa_Init()
      cleaned_ = .false.
   end subroutine

   !> src/bsa/BsaLibImpl.F90:683-685: releases every input and the device.
   subroutine bsa_Finalise()
      integer(c_int) :: rc
      if (allocated(n_load_)) deallocate(n_load_)
      if (allocated(libs_load_)) deallocate(libs_load_)
      if (allocated(modes_user_)) deallocate(modes_user_)
      if (allocated(modes_)) deallocate(modes_)
      if (allocated(phi_)) deallocate(phi_)
      if (allocated(natf_)) deallocate(natf_)
      if (allocated(xsi_)) deallocate(xsi_)
      if (allocated(Mg_)) deallocate(Mg_)
      if (allocated(Kg_)) deallocate(Kg_)
      if (allocated(Cg_)) deallocate(Cg_)
      if (allocated(tc_)) deallocate(tc_)
      if (allocated(dirs_)) deallocate(dirs_)
      if (allocated(wz_node_)) deallocate(wz_node_)
      if (allocated(u_node_)) deallocate(u_node_)
      if (allocated(umref_)) deallocate(umref_)
      if (allocated(corr_)) deallocate(corr_)
      if (allocated(wfc_)) deallocate(wfc_)
      if (allocated(psi_user_)) deallocate(psi_user_)
      tscl_addr_ = c_null_ptr
      tstd_addr_ = c_null_ptr
      if (allocated(lp_psi_)) deallocate(lp_psi_, lp_Mk_, lp_Kk_, lp_Ck_, lp_tscl_, lp_tstd_, lp_umref_, lp_tcs_)
      if (allocated(coords_)) deallocate(coords_)
      if (allocated(zref_)) deallocate(zref_)
      if (allocated(incang_)) deallocate(incang_)
      if (allocated(walt_)) deallocate(walt_)
      if (allocated(wz_lims_)) deallocate(wz_lims_)
      ccoef_addr_ = c_null_ptr
      cexp_addr_ = c_null_ptr
      rotlw_addr_ = c_null_ptr
      gen_inputs_ = .false.
      run_after_gen_ = .true.
      lp_ready_ = .false.
      call unbindSlots_()
      nn_ = 0; nlibs_ = 0; ndofs_ = 0; nm_ = 0; nz_lim_ = 0; ipsd_ = 0
      if (device_ready_) rc = bsahip_finalise()
      device_ready_ = .false.
      cleaned_ = .true.
   end subroutine

   logical pure function bsa_isCleaned()
      bsa_isCleaned = cleaned_
   end function

   !> Marks the run as GPU-offloaded (src/bsa/BsaLibImpl.F90:54-56); as in
   !> the reference's GPU path it forces the Classic approach.
   subroutine bsa_enableGPU()
      gpu_enabled_ = .true.
   end subroutine

   !> Extension (no counterpart in the reference): node-contraction precision,
   !> BSA_PREC_FP64 (default), BSA_PREC_FP32 or BSA_PREC_BF16X3.
   subroutine bsa_setComputePrecision(iprec)
      integer(bsa_int_t), value :: iprec
      if (iprec /= BSA_PREC_FP64 .and. iprec /= BSA_PREC_FP32 .and. iprec /= BSA_PREC_BF16X3) &
         call bsa_Abort('Unknown compute precision (use BSA_PREC_FP64, BSA_PREC_FP32 or BSA_PREC_BF16X3).')
      iprec_ = iprec
   end subroutine

   ! ======================================================================
   !  settings (src/settings/SettingsImpl.F90)
   ! ======================================================================

   subroutine bsa_setAnalysisType(isuban)
      integer(bsa_int_t), value :: isuban
      isuban_ = isuban
   end subroutine

   subroutine bsa_setClassicMode(i_mode)
      integer(bsa_int_t), value :: i_mode
      iclassic_ = i_mode
   end subroutine

   subroutine bsa_setScalingConv(iconv)
      integer(bsa_int_t), value :: iconv
      iconv_ = iconv
   end subroutine

   subroutine bsa_setSpectraComputation(ipsd, ibisp)
      integer(bsa_int_t), value :: ipsd, ibisp
      icomp_psd_ = ipsd
      icomp_bisp_ = ibisp
   end subroutine

   subroutine bsa_setSpectraExtension(ionlydiag)
      integer(bsa_int_t), value :: ionlydiag
      ionly_diag_ = ionlydiag
   end subroutine

   subroutine bsa_setTestMode(itest)
      integer(bsa_int_t), value :: itest
      itest_ = itest
   end subroutine

   !> Accepted values and the fallback to HALF for any other value as
   !> src/bsa/BsaLibImpl.F90:750-767 (bsa_setSpatialSymmetry).  The device
   !> paths do not use it; it is written to the input files and to the header
   !> of the bispectrum export file.
   subroutine bsa_setSpatialSymmetry(isym)
      integer(bsa_int_t), value :: isym
      if (isym == BSA_SPATIAL_SYM_NONE .or. isym == BSA_SPATIAL_SYM_FOUR) then
         isym_ = isym
      else
         if (isym /= BSA_SPATIAL_SYM_HALF) then
            write(output_unit, '(1x, 2a, i0, a)') WARNMSG, 'Invalid   ', isym, '  spatial symmetry value.'
            write(output_unit, '(1x, 2a)') MSGCONT, 'Setting default (HALF).'
         end if
         isym_ = BSA_SPATIAL_SYM_HALF
      end if
   end subroutine

   !> src/settings/SettingsImpl.F90:107-119.  isvd = 1 (the default) selects the
   !> POD BFM (getFM_full_tm_scalar_msh_POD_, BsaLibImpl.F90:364-368), which
   !> reads coherence column 1 for every turbulence component; 0 the direct
   !> BFM with column tc.  ifcov and idumpmod are accepted for interface
   !> compatibility (the device Mesher covers the full domain and keeps its
   !> zones in memory).
   subroutine bsa_setupMesher(isvd, bkgrfmt, bkgaext, genpaext, maxaext, ifcov, idumpmod)
      integer(bsa_int_t), value :: isvd, bkgrfmt, ifcov, idumpmod
      real(bsa_real_t), value :: bkgaext, genpaext, maxaext
      isvd_ = isvd
      bkgrfmt_ = bkgrfmt
      bkgaext_ = bkgaext
      genpaext_ = genpaext
      maxaext_ = maxaext
   end subroutine

   !> src/bsa/BsaLibImpl.F90:100-107: keep the POD modes whose eigenvalue is at
   !> least rval % of the largest (Mesher, POD BFM only).
   subroutine bsa_setPODTruncationThreshold(rval)
      real(bsa_real_t), value :: rval
      if (rval > 0.0_bsa_real_t) then
         do_trunc_pod_ = .true.
         pod_trunc_lim_ = rval / 100.0_bsa_real_t
      end if
   end subroutine

   !> src/bsa/BsaLibImpl.F90:110-115: keep the nmodes largest POD modes.
   subroutine bsa_setPODNOfModesKept(nmodes)
      integer(bsa_int_t), value :: nmodes
      npod_set_ = .true.
      npod_ = nmodes
   end subroutine

   !> BSA_PREMESH_TYPE_DIAG_CREST_YES also resolves the f1+f2 = +-f_n crests.
   subroutine bsa_setPremeshScheme(itype)
      integer(bsa_int_t), value :: itype
      premesh_type_ = itype
   end subroutine

   !> Background peak width from the first direction only (BsaMesherImpl.F90:298-302).
   subroutine bsa_setMaxBkgPeakRestriction(bool)
      logical, intent(in) :: bool
      restrict_bkg_ = bool
   end subroutine

   subroutine bsa_doValidateModalData(bool)
      logical, value :: bool
      validate_modal_ = bool
   end subroutine

   subroutine bsa_setupClassic(nfreqs, df)
      integer(bsa_int_t), value :: nfreqs
      real(bsa_real_t), value :: df
      nfreqs_ = nfreqs
      df_ = df
   end subroutine

   ! ======================================================================
   !  wind (src/wind/WindSetImpl.F90)
   ! ======================================================================

   subroutine bsa_setWindDirections(dirs, ndirs)
      integer(bsa_int_t), intent(in) :: dirs(:)
      integer(bsa_int_t), value, optional :: ndirs
      integer(bsa_int_t) :: n
      n = size(dirs)
      if (present(ndirs)) n = ndirs
      dirs_ = dirs(1:n)
   end subroutine

   subroutine bsa_setWindTurbComps(tc, ntc)
      integer(bsa_int_t), intent(in) :: tc(:)
      integer(bsa_int_t), value, optional :: ntc
      integer(bsa_int_t) :: n
      n = size(tc)
      if (present(ntc)) n = ntc
      tc_ = tc(1:n)
   end subroutine

   subroutine bsa_setPSDType(ipsd)
      integer(bsa_int_t), value :: ipsd
      ipsd_ = ipsd
   end subroutine

   subroutine bsa_setWindZoneLimits(lim, ilim)
      real(bsa_real_t), intent(in), target :: lim(:)
      integer(bsa_int_t), intent(in), optional :: ilim(:)
      if (nz_lim_ /= 0 .and. nz_lim_ /= size(lim) - 1) &
         call bsa_Abort('N. of wind zones differs from the one previously set.')
      nz_lim_ = size(lim) - 1
      wz_lims_ = lim
   end subroutine

   subroutine bsa_setWZMeanWindVel(mat)
      real(bsa_real_t), target, intent(in) :: mat(:)
      umref_ = mat
   end subroutine

   !> L(3, 3, NWZ): turbulence length scales per wind zone.
   subroutine bsa_setTurbWindScales(L)
      real(bsa_real_t), target, intent(in) :: L(3, 3, *)
      tscl_addr_ = c_loc(L(1, 1, 1))
   end subroutine

   !> sigma(3, NWZ): turbulence standard deviations per wind zone.
   subroutine bsa_setTurbWindSDT(sigma)
      real(bsa_real_t), target, intent(in) :: sigma(3, *)
      tstd_addr_ = c_loc(sigma(1, 1))
   end subroutine

   subroutine bsa_setNodalVel(Unod)
      real(bsa_real_t), target, intent(in) :: Unod(:)
      u_node_ = Unod
   end subroutine

   subroutine bsa_setNodalWindZones(NodWZ)
      integer(bsa_int_t), target, intent(in) :: NodWZ(:)
      wz_node_ = NodWZ
   end subroutine

   !> Packed lower triangle ((NN^2+NN)/2, 3) of the nodal coherence, one column
   !> per turbulence component.
   subroutine bsa_setSpatialNodalCorr(nodCorr)
      real(bsa_real_t), target, intent(in) :: nodCorr(:, :)
      corr_ = nodCorr
   end subroutine

   !> wfc(NLIBSL, NDEGW, NNL): wind force coefficients.
   subroutine bsa_setWindFCoeffs(wfc)
      real(bsa_real_t), target, intent(in) :: wfc(:, :, :)
      wfc_ = wfc
   end subroutine

   !> phiTc(NM_EFF, NNL, NDEGW): projected coefficients given by the caller
   !> (otherwise bsa_Run builds them from Phi and wfc).
   subroutine bsa_setPhitimesC(phiTc)
      real(bsa_real_t), target, intent(in) :: phiTc(:, :, :)
      psi_user_ = phiTc
   end subroutine

   ! ======================================================================
   !  structure (src/structure/StructureImpl.F90)
   ! ======================================================================

   subroutine bsa_setTotalNumOfDOFs(ndofs)
      integer(bsa_int_t), value :: ndofs
      ndofs_ = ndofs
   end subroutine

   subroutine bsa_setNumOfNodalDOFs(nndofs)
      integer(bsa_int_t), value :: nndofs
      nlibs_ = nndofs
   end subroutine

   subroutine bsa_setTotalNOfNodes(nn)
      integer(bsa_int_t), value :: nn
      nn_ = nn
   end subroutine

   subroutine bsa_setLoadedNodalDOFs(libs_l, nlibs_l)
      integer(bsa_int_t), intent(in), target, allocatable :: libs_l(:)
      integer(bsa_int_t), value, optional :: nlibs_l
      integer(bsa_int_t) :: n
      n = size(libs_l)
      if (present(nlibs_l)) n = nlibs_l
      libs_load_ = libs_l(1:n)
   end subroutine

   subroutine bsa_setLoadedNodes(nodes_l, nn_l)
      integer(bsa_int_t), intent(in), target, allocatable :: nodes_l(:)
      integer(bsa_int_t), value, optional :: nn_l
      integer(bsa_int_t) :: n
      n = size(nodes_l)
      if (present(nn_l)) n = nn_l
      n_load_ = nodes_l(1:n)
   end subroutine

   subroutine bsa_setModalInfo(ndofs, nm, Phi, natf)
      integer(bsa_int_t), value :: ndofs, nm
      real(bsa_real_t), intent(in), target :: Phi(ndofs, nm), natf(nm)
      if (nm_ /= 0 .and. nm_ /= nm) &
         call bsa_Abort('You passed a value of "nm" which differs from previously set.')
      nm_ = nm
      phi_ = Phi
      natf_ = natf
   end subroutine

   subroutine bsa_setKeptModalShapes(modes)
      integer(bsa_int_t), intent(in) :: modes(:)
      modes_user_ = modes
   end subroutine

   !> Generalised mass, stiffness (NM) and damping (NM, NM) matrices; the FRF
   !> uses the diagonal of Cgen.
   subroutine bsa_setModalMatrices(nm, Mgen, Kgen, Cgen)
      integer(bsa_int_t), value :: nm
      real(bsa_real_t), intent(in), target :: Mgen(nm), Kgen(nm), Cgen(nm, nm)
      if (nm_ /= 0 .and. nm_ /= nm) &
         call bsa_Abort('You passed a value of "nm" which differs from previously set.')
      nm_ = nm
      Mg_ = Mgen
      Kg_ = Kgen
      Cg_ = Cgen
   end subroutine

   subroutine bsa_setTotDamping(xsi)
      real(bsa_real_t), intent(in) :: xsi(:)
      if (nm_ == 0) call bsa_Abort('NM == 0 when setting Damping info.')
      xsi_ = xsi(1:nm_)
   end subroutine

   !> Modes used by the last run (after validation).
   pure function bsa_getUsedModeShapes() result(modes)
      integer(bsa_int_t), allocatable :: modes(:)
      if (allocated(modes_)) then
         modes = modes_
      else
         allocate(modes(0))
      end if
   end function

   ! ======================================================================
   !  run
   ! ======================================================================

   !> Kept modes as validateModalInfo_ (src/bsa/BsaLibImpl.F90:545-599):
   !> with validation on, modes whose max|phi| is not exactly 1 are dropped
   !> and, if any is dropped, the kept list is every normalised mode.
   subroutine selectModes_()
      real(bsa_real_t) :: mx
      integer(bsa_int_t) :: i, j, nk
      logical :: keep(nm_)
      if (validate_modal_) then
         do i = 1, nm_
            mx = maxval(abs(phi_(:, i)))
            keep(i) = (mx == 1.0_bsa_real_t)
            if (.not. keep(i)) write(output_unit, '(1x, 2a, i3, a)') &
               WARNMSG, 'Mode  ', i, '  is not 1-normalised. Ignoring it..'
         end do
         nk = count(keep)
         if (nk < nm_) then
            allocate(modes_(nk))
            j = 0
            do i = 1, nm_
               if (keep(i)) then
                  j = j + 1
                  modes_(j) = i
               end if
            end do
            return
         end if
      end if
      if (allocated(modes_user_)) then
         modes_ = modes_user_
      else
         modes_ = [(i, i = 1, nm_)]
      end if
   end subroutine

   !> Phi*C projection as setPhitimesCLocalInstance_ (src/bsa/BsaLibImpl.F90:611-671):
   !> psi(im, in, id) = sum_l wfc(l, id, in) * phi((n_load(in)-1)*NLIBS + libs_load(l), modes(im)).
   subroutine buildPsi_(psi)
      real(bsa_real_t), allocatable, intent(out) :: psi(:, :, :)
      integer(bsa_int_t) :: nlibl, ndegw, nnl, nme, in, id, im, skip
      nlibl = size(wfc_, 1)
      ndegw = size(wfc_, 2)
      nnl = size(wfc_, 3)
      nme = size(modes_)
      if (nnl /= size(n_load_)) call bsa_Abort('N. of loaded nodes does not match.')
      if (nlibl /= size(libs_load_)) call bsa_Abort('N. of loaded libs  does not match.')
      if (nme == 0) call bsa_Abort('Effective number of modes to be kept is 0. Validate modal info before this step.')
      allocate(psi(nme, nnl, ndegw))
      do in = 1, nnl
         skip = (n_load_(in) - 1) * nlibs_
         do id = 1, ndegw
            do im = 1, nme
               psi(im, in, id) = sum(wfc_(:, id, in) * phi_(skip + libs_load_, modes_(im)))
            end do
         end do
      end do
   end subroutine

   !> Frequency grid as computeFreqsVect_ (src/bsa/classic/BsaClassicImpl.F90:437-560):
   !> outside test mode, df and the number of points are raised so that the
   !> grid reaches the highest natural frequency with df <= min(xsi*f_nat)/5.
   !> Returns the number of non-negative frequencies (pulsation convention).
   subroutine classicGrid_(nfh, df)
      integer(bsa_int_t), intent(out) :: nfh
      real(bsa_real_t), intent(out) :: df
      real(bsa_real_t) :: df_ref, max_freq, max_freq_ref
      integer(bsa_int_t) :: nf1
      logical :: df_big
      if (nfreqs_ == 0 .or. df_ == 0.0_bsa_real_t) call bsa_Abort('Either NFREQs or DF are == 0.')
      df = df_
      nf1 = nfreqs_ - 1
      if (itest_ == 0) then
         if (.not. allocated(xsi_)) &
            call bsa_Abort('Damping ratios are needed to check the frequency grid (bsa_setTotDamping).')
         df_ref = minval(xsi_ * natf_) / 5
         df_big = df > df_ref
         max_freq = df * nf1
         max_freq_ref = maxval(natf_)
         if (max_freq < max_freq_ref) then
            if (df_big) then
               max_freq = df_ref * nf1
               if (max_freq < max_freq_ref) nf1 = ceiling(max_freq_ref / df_ref)
               df = df_ref
            else
               nf1 = ceiling(max_freq_ref / df)
            end if
            write(output_unit, '(1x, 2a, i0, a, es12.5)') INFOMSG, 'Classic grid adjusted: nfreqs= ', &
               nf1 + 1, '  df= ', df
         end if
      end if
      nfh = nf1 + 1
   end subroutine

   !> Classic moments on the GPU (the Classic branch of src/bsa/BsaLibImpl.F90:153-462).
   !> m2* have NM_EFF^2 entries, m3* NM_EFF^3 (first index fastest); the
   !> Mesher outputs m3mf_msh, m3mr_msh are left unallocated.
   subroutine bsa_Run(m2mf_cls, m2mr_cls, m2o2mr_cls, m3mf_msh, m3mr_msh, m3mf_cls, m3mr_cls)
      real(bsa_real_t), target, allocatable, dimension(:) :: &
         m2mf_cls, m2mr_cls, m2o2mr_cls, m3mf_msh, m3mr_msh, m3mf_cls, m3mr_cls
      type(bsahip_problem_t) :: p
      real(bsa_real_t), allocatable, target :: psi(:, :, :), Mk(:), Kk(:), Ck(:)
      real(bsa_real_t), allocatable, target :: tscl(:, :, :), tstd(:, :), umref(:)
      integer(bsa_int_t), allocatable, target :: tcs(:)
      real(bsa_real_t), pointer :: fp3(:, :, :), fp2(:, :)
      real(bsa_real_t) :: df
      integer(bsa_int_t) :: nfh, nme, nwz, i, iclassic, n2, n3
      logical :: do_cls, do_msh
      integer(c_int32_t) :: flags
      integer(c_int) :: rc
      type(c_ptr) :: o_m2mf, o_m2mr, o_m2o2mr, o_m3mf, o_m3mr

      ! input files of the standalone program (src/bsa/BsaLibImpl.F90:156-161)
      if (gen_inputs_) then
         write(output_unit, '(/ 1x, 2a)') INFOMSG, 'Generating BSA-compatible input files.'
         call generateInputFiles_()
      end if
      if (.not. run_after_gen_) return
      if (icomp_psd_ == 0 .and. icomp_bisp_ == 0) then
         write(output_unit, '(1x, 2a)') WARNMSG, 'Both  PSD  and  BISP  computation are disabled.'
         return
      end if
      ! the reference's GPU switch forces the Classic approach outside visual mode
      ! (src/bsa/BsaLibImpl.F90:406-412)
      if (gpu_enabled_ .and. .not. visual_) isuban_ = 1
      if (isuban_ < 1 .or. isuban_ > 3) call bsa_Abort('Analysis type must be 1, 2 or 3.')
      do_cls = isuban_ == 1 .or. isuban_ == 3
      do_msh = isuban_ == 2 .or. isuban_ == 3
      if (iconv_ /= BSA_PSD_CONVENTION_PULS) &
         call bsa_Abort('Only the pulsation convention (BSA_PSD_CONVENTION_PULS) is available in this library.')

      ! ---- structure checks (validateAll_, BsaLibImpl.F90:468-541) ----
      if (.not. allocated(phi_)) call bsa_Abort('Modal matrix was not acquired. Aborting.')
      if (ndofs_ == 0) ndofs_ = nn_ * nlibs_
      if (visual_ .and. .not. vis_modal_) then         ! BsaLibImpl.F90:202-207
         if (vis_idx_(1) > nn_) call bsa_Abort('Node index exceeds max n. of nodes')
         if (vis_idx_(2) > nlibs_) call bsa_Abort('DOF index exceeds max n. of DOFs per node')
      end if
      if (ndofs_ /= size(phi_, 1)) call bsa_Abort('NDOFs does not match modal matrix size.')
      if (.not. allocated(n_load_)) n_load_ = [(i, i = 1, nn_)]
      if (.not. allocated(libs_load_)) libs_load_ = [(i, i = 1, nlibs_)]
      if (.not. (allocated(Mg_) .and. allocated(Kg_) .and. allocated(Cg_))) &
         call bsa_Abort('Modal matrices were not acquired (bsa_setModalMatrices).')
      if (allocated(modes_)) deallocate(modes_)
      call selectModes_()
      if (export_modal_) call exportModalData_()
      nme = size(modes_)
      if (nme == 0) call bsa_Abort('Effective number of modes to be kept is 0.')
      if (any(modes_ < 1) .or. any(modes_ > nm_)) call bsa_Abort('Kept mode index out of range.')
      Mk = Mg_(modes_)
      Kk = Kg_(modes_)
      allocate(Ck(nme))
      do i = 1, nme
         Ck(i) = Cg_(modes_(i), modes_(i))
      end do

      ! ---- wind (turbulence components default as SetTurbCompsAndDirsDefault) ----
      if (allocated(tc_) .and. allocated(dirs_)) then
         tcs = tc_
      else
         tcs = [1]
      end if
      if (.not. allocated(wfc_)) call bsa_Abort('Wind force coefficients were not acquired. Aborting.')
      if (allocated(psi_user_)) then
         if (size(psi_user_, 1) /= nme .or. size(psi_user_, 2) /= size(n_load_)) &
            call bsa_Abort('Phi*C sizes do not match the kept modes and loaded nodes.')
         psi = psi_user_
      else
         call buildPsi_(psi)
      end if
      if (.not. (allocated(u_node_) .and. allocated(wz_node_) .and. allocated(corr_))) &
         call bsa_Abort('Nodal wind data missing (velocities, wind zones or nodal correlation).')
      if (.not. (c_associated(tscl_addr_) .and. c_associated(tstd_addr_))) &
         call bsa_Abort('Turbulence scales and standard deviations were not acquired.')
      if (size(corr_, 2) < maxval(tcs)) call bsa_Abort('Nodal correlation needs one column per turbulence component.')
      if (nz_lim_ > 0) then
         nwz = nz_lim_
      else if (allocated(umref_)) then
         nwz = size(umref_)
      else
         nwz = maxval(wz_node_)
      end if
      call c_f_pointer(tscl_addr_, fp3, [3, 3, nwz])
      call c_f_pointer(tstd_addr_, fp2, [3, nwz])
      tscl = fp3
      tstd = fp2
      if (allocated(umref_)) then
         umref = umref_
      else
         allocate(umref(nwz), source=0.0_bsa_real_t)
      end if

      p%nn = nn_
      p%nnl = size(n_load_)
      p%nm = nme
      p%ntc = size(tcs)
      p%ndegw = size(psi, 3)
      p%nwz = nwz
      p%psd_id = ipsd_
      p%pulsation = 1
      p%ncorr_rows = size(corr_, 1)
      p%corr_col_mode = 0
      p%n_load = c_loc(n_load_)
      p%tc = c_loc(tcs)
      p%psi = c_loc(psi)
      p%nod_corr = c_loc(corr_)
      p%u_node = c_loc(u_node_)
      p%wz_node = c_loc(wz_node_)
      p%turb_scales = c_loc(tscl)
      p%turb_std = c_loc(tstd)
      p%u_mean_ref = c_loc(umref)
      p%Mm = c_loc(Mk)
      p%Km = c_loc(Kk)
      p%Cm = c_loc(Ck)
      ! keep a copy for bsa_computeBRdecomp (module arrays outlive this call)
      lp_psi_ = psi
      lp_Mk_ = Mk
      lp_Kk_ = Kk
      lp_Ck_ = Ck
      lp_tscl_ = tscl
      lp_tstd_ = tstd
      lp_umref_ = umref
      lp_tcs_ = tcs
      lp_ = p
      lp_%psi = c_loc(lp_psi_)
      lp_%tc = c_loc(lp_tcs_)
      lp_%turb_scales = c_loc(lp_tscl_)
      lp_%turb_std = c_loc(lp_tstd_)
      lp_%u_mean_ref = c_loc(lp_umref_)
      lp_%Mm = c_loc(lp_Mk_)
      lp_%Km = c_loc(lp_Kk_)
      lp_%Cm = c_loc(lp_Ck_)
      lp_ready_ = .true.
      lp_msh_ = meshProblem_(lp_)
      lp_diag_ = ionly_diag_ /= 0
      call unbindSlots_()
      getBFM_msh => opBFMMsh_
      getBRM_msh => opBRMMsh_

      if (.not. device_ready_) then
         rc = bsahip_init(0_c_int)
         if (rc /= 0) call bsa_Abort('GPU initialisation failed: '//bsahip_last_error_f())
         device_ready_ = .true.
      end if
      if (allocated(m3mf_msh)) deallocate(m3mf_msh)
      if (allocated(m3mr_msh)) deallocate(m3mr_msh)
      if (do_msh) call runMesher_(p, m3mf_msh, m3mr_msh, tscl, umref)
      if (.not. do_cls) return

      call classicGrid_(nfh, df)
      lp_nfh_ = nfh
      lp_df_ = df
      getBFM_vect_cls => opBFMVectCls_
      getBRM_vect_cls => opBRMVectCls_
      getBFM_scalar_cls => opBFMScalarCls_
      getBRM_scalar_cls => opBRMScalarCls_

      ! Classic version as checkMaxAllocation_ (BsaClassicImpl.F90:384-434): outside
      ! test mode a vectorised bispectrum above 1e9 elements switches to the scalar
      ! version, whose third moments are plain Riemann sums (every point times
      ! dw^2, BsaClassicImpl.F90:331,341); the vectorised version uses the
      ! edge/corner-corrected rule of intgSpectraVect_.  The element count is
      ! formed in 64-bit here (the reference forms it in 32-bit, which wraps).
      ! ONLY_DIAG model: NM diagonal entries per output (dimM_psd_ = dimM_bisp_ = NM)
      n2 = nme**2
      n3 = nme**3
      if (ionly_diag_ /= 0) then
         n2 = nme
         n3 = nme
      end if
      iclassic = iclassic_
      if (itest_ == 0 .and. icomp_bisp_ /= 0 .and. iclassic == BSA_CLASSIC_MODE_VECTOR) then
         if (int(2 * nfh - 1, int64)**2 * int(n3, int64) > 1000000000_int64) then
            write(output_unit, '(1x, 2a)') WARNMSG, 'Too high allocation size for VECTORISED BSA version: using SCALAR.'
            iclassic = BSA_CLASSIC_MODE_SCALAR
         end if
      end if

      if (iclassic == BSA_CLASSIC_MODE_SCALAR) then
         flags = int(iprec_, c_int32_t) + BSAHIP_QUAD_RIEMANN
      else
         flags = int(iprec_, c_int32_t) + BSAHIP_QUAD_VECT
      end if
      if (ionly_diag_ /= 0) flags = ior(flags, BSAHIP_ONLY_DIAG)
      o_m2mf = c_null_ptr; o_m2mr = c_null_ptr; o_m2o2mr = c_null_ptr
      o_m3mf = c_null_ptr; o_m3mr = c_null_ptr
      if (icomp_psd_ /= 0) then
         call realloc_(m2mf_cls, n2)
         call realloc_(m2mr_cls, n2)
         call realloc_(m2o2mr_cls, n2)
         o_m2mf = c_loc(m2mf_cls)
         o_m2mr = c_loc(m2mr_cls)
         o_m2o2mr = c_loc(m2o2mr_cls)
      else
         flags = ior(flags, BSAHIP_NO_PSD)
      end if
      if (icomp_bisp_ /= 0) then
         call realloc_(m3mf_cls, n3)
         call realloc_(m3mr_cls, n3)
         o_m3mf = c_loc(m3mf_cls)
         o_m3mr = c_loc(m3mr_cls)
      else
         flags = ior(flags, BSAHIP_NO_BISP)
      end if

      rc = bsahip_classic_run(p, int(nfh, c_int32_t), real(df, c_double), flags, &
                              o_m2mf, o_m2mr, o_m2o2mr, o_m3mf, o_m3mr)
      if (rc /= 0) call bsa_Abort('bsahip_classic_run failed: '//bsahip_last_error_f())
      write(output_unit, '(1x, 2a, i0, a, i0, a, a)') INFOMSG, 'Classic run on the GPU: ', &
         nme, ' modes, ', 2 * nfh - 1, '^2 frequency points, ', &
         merge('SCALAR    ', 'VECTORISED', iclassic == BSA_CLASSIC_MODE_SCALAR)
      if (visual_) then
         ! the reference exports the modal BRM of the vectorised version only
         ! (BsaClassicImpl.F90:158); its other visual cases re-read an empty
         ! export file (writeNTotPtsToBispUn_, BsaLibImpl.F90:1832-1845)
         if (.not. vis_modal_) then
            write(output_unit, '(1x, 2a)') WARNMSG, &
               'Nodal (BRN) bispectrum export is not defined for the Classic approach; no file written.'
         else if (icomp_bisp_ == 0 .or. iclassic /= BSA_CLASSIC_MODE_VECTOR) then
            write(output_unit, '(1x, 2a)') WARNMSG, &
               'BRM export needs the bispectrum and the VECTORISED Classic version; no file written.'
         else
            call exportVisualCls_(p, nfh, df, nme)
         end if
      end if
   contains
      subroutine runMesher_(p, m3mf, m3mr, tscl, umref)
         type(bsahip_problem_t), intent(in) :: p
         real(bsa_real_t), allocatable, target, intent(inout) :: m3mf(:), m3mr(:)
         real(bsa_real_t), intent(in) :: tscl(:, :, :), umref(:)
         type(bsahip_mesher_cfg_t) :: cfg
         type(bsahip_problem_t) :: pm
         type(c_ptr) :: msh
         real(bsa_real_t), allocatable, target :: fk(:), xk(:)
         real(bsa_real_t) :: v(3, 3), bw
         integer(c_int64_t), target :: counts(3), adapt_evals
         integer(bsa_int_t) :: z, nmk, nc
         integer(c_int32_t) :: mflags, ntl
         integer(c_int) :: rc2
         if (.not. allocated(xsi_)) &
            call bsa_Abort('The Mesher needs the modal damping ratios (bsa_setTotDamping).')
         ! background peak width: 1 / max_z(L/U), maximised over (component, direction)
         ! (validateAll_, BsaLibImpl.F90:513-527; StructureImpl.F90:225-260)
         v = 0
         do z = 1, size(umref)
            if (umref(z) > 0) v = max(v, tscl(:, :, z) / umref(z))
         end do
         where (v > 0) v = 1 / v
         if (restrict_bkg_) then
            bw = maxval(v(:, 1))
         else
            bw = maxval(v)
         end if
         if (.not. bw > 0) call bsa_Abort('Background peak width is 0 (turbulence scales or mean speeds).')
         nmk = size(modes_)
         fk = natf_(modes_)
         xk = xsi_(modes_)
         cfg%bkg_base_rfmnt = bkgrfmt_
         cfg%tile_cells = 32
         cfg%diag_crest = merge(1, 0, premesh_type_ == BSA_PREMESH_TYPE_DIAG_CREST_YES)
         cfg%bkg_peakw = bw
         cfg%bkg_area_ext = bkgaext_
         cfg%peak_area_ext = genpaext_
         cfg%max_area_ext = maxaext_
         cfg%natf = c_loc(fk)
         cfg%xsi = c_loc(xk)
         ! POD BFM (isvd = 1): coherence column 1 for every component (the POD
         ! path passes idir = 1, functionsImpl.F90:628-629, 770); truncation
         ! settings validated as validateAll_ (BsaLibImpl.F90:479-486)
         pm = meshProblem_(p)
         if (isvd_ == 1 .and. npod_set_) &
            write(output_unit, '(/ 1x, 2a, i0, a /)') INFOMSG, 'Using  ', npod_, '  POD modes.'
         ! ONLY_DIAG: NM entries (dimM_bisp_ = NM), no POD (BsaMesherImpl.F90:68)
         nc = nmk**3
         mflags = 0
         if (ionly_diag_ /= 0) then
            nc = nmk
            mflags = BSAHIP_ONLY_DIAG
         end if
         rc2 = bsahip_mesher_create(pm, cfg, mflags, msh)
         if (rc2 /= 0) call bsa_Abort('bsahip_mesher_create failed: '//bsahip_last_error_f())
         if (only_premesh_) then
            write(output_unit, '(1x, 2a)') NOTEMSG, 'Skipping Post-meshing phase!'
            rc2 = bsahip_mesher_destroy(msh)
            return
         end if
         allocate(m3mf(nc), m3mr(nc))
         if (allocated(msh_levels_)) deallocate(msh_levels_)
         if (mesher_tol_ > 0.0_bsa_real_t) then
            ! a-posteriori levels; the moments come from the same integration
            rc2 = bsahip_mesher_ntiles(msh, ntl)
            allocate(msh_levels_(ntl))
            rc2 = bsahip_mesher_adapt_run(msh, real(mesher_tol_, c_double), c_loc(msh_levels_), &
                                          c_loc(adapt_evals), c_loc(m3mf), c_loc(m3mr), c_loc(counts))
            if (rc2 /= 0) call bsa_Abort('bsahip_mesher_adapt_run failed: '//bsahip_last_error_f())
         else
            rc2 = bsahip_mesher_run(msh, c_null_ptr, c_loc(m3mf), c_loc(m3mr), c_loc(counts))
            if (rc2 /= 0) call bsa_Abort('bsahip_mesher_run failed: '//bsahip_last_error_f())
         end if
         if (visual_ .or. brm_base_) call exportMesher_(msh, nmk)
         rc2 = bsahip_mesher_destroy(msh)
         write(output_unit, '(1x, 2a, i0, a, i0, a, i0, a)') INFOMSG, 'Mesher on the GPU: ', &
            counts(1) + counts(2), ' BFM points (', counts(1), ' level-0), ', counts(3), ' BRM points.'
      end subroutine

      subroutine realloc_(a, n)
         real(bsa_real_t), allocatable, intent(inout) :: a(:)
         integer(bsa_int_t), intent(in) :: n
         if (allocated(a)) deallocate(a)
         allocate(a(n), source=0.0_bsa_real_t)
      end subroutine
   end subroutine bsa_Run

   !> The Mesher's problem: with the POD BFM (isvd = 1) coherence column 1 for
   !> every component (the POD path passes idir = 1, functionsImpl.F90:628-629,
   !> 770) and the truncation settings validated as validateAll_
   !> (BsaLibImpl.F90:479-486).
   function meshProblem_(p) result(pm)
      type(bsahip_problem_t), intent(in) :: p
      type(bsahip_problem_t) :: pm
      pm = p
      if (isvd_ /= 1) return
      pm%corr_col_mode = 1
      if (do_trunc_pod_ .and. (pod_trunc_lim_ == 0.0_bsa_real_t .or. &
                               pod_trunc_lim_ == 1.0_bsa_real_t)) do_trunc_pod_ = .false.
      if (npod_set_ .and. (npod_ <= 0 .or. npod_ >= p%nnl)) npod_set_ = .false.
      if (do_trunc_pod_) pm%pod_trunc_lim = pod_trunc_lim_
      if (npod_set_) pm%pod_nmodes = npod_
   end function

   subroutine unbindSlots_()
      nullify(getBFM_vect_cls, getBRM_vect_cls, getBFM_scalar_cls, getBRM_scalar_cls)
      nullify(getBFM_msh, getBRM_msh)
   end subroutine

   subroutine slotCheck_(rc, what)
      integer(c_int), intent(in) :: rc
      character(len=*), intent(in) :: what
      if (rc /= 0) call bsa_Abort(what//' failed: '//bsahip_last_error_f())
   end subroutine

   !> Half-length and spacing of a symmetric Classic grid f(i) = (i - nfh) df.
   subroutine slotGrid_(f, nfh, df)
      real(bsa_real_t), intent(in) :: f(:)
      integer(bsa_int_t), intent(out) :: nfh
      real(bsa_real_t), intent(out) :: df
      integer(bsa_int_t) :: nf
      nf = size(f)
      if (nf < 3 .or. mod(nf, 2) == 0) call bsa_Abort('The Classic grid needs an odd number (>= 3) of frequencies.')
      nfh = (nf + 1) / 2
      df = f(2) - f(1)
      if (.not. df > 0 .or. abs(f(nfh)) > 1e-9_bsa_real_t * df .or. &
          abs(f(1) + f(nf)) > 1e-9_bsa_real_t * df * nf) &
         call bsa_Abort('The Classic grid must be symmetric about f = 0 with a constant step.')
   end subroutine

   !> Entries per point of the modal PSD and bispectrum of the bound model.
   subroutine slotDims_(n2, n3, flags)
      integer(bsa_int_t), intent(out) :: n2, n3
      integer(c_int32_t), intent(out) :: flags
      n2 = lp_%nm**2
      n3 = lp_%nm**3
      flags = 0
      if (lp_diag_) then
         n2 = lp_%nm
         n3 = lp_%nm
         flags = BSAHIP_ONLY_DIAG
      end if
   end subroutine

   !> getFM_full_tnm_vect_cls_ (functionsImpl.F90:1312-1587) or, ONLY_DIAG,
   !> getFM_diag_tnlm_vect_cls_ (:1747-1970): modal PSDs psd(NF, E2) and, when
   !> bispectra are computed, bisp(NF, NF, E3) on the grid f, S(f1 + f2) zero
   !> beyond the grid; allocated here as there.
   subroutine opBFMVectCls_(f, Suvw, psd, bisp)
      real(bsa_real_t), intent(in) :: f(:)
      real(bsa_real_t), intent(in) :: Suvw(:, :)
      real(bsa_real_t), allocatable, intent(inout) :: psd(:, :), bisp(:, :, :)
      integer(bsa_int_t) :: nfh, nf, n2, n3, i
      integer(c_int32_t) :: fl
      real(bsa_real_t) :: df
      call slotGrid_(f, nfh, df)
      call slotDims_(n2, n3, fl)
      nf = size(f)
      if (size(Suvw, 1) /= nf) call bsa_Abort('getBFM_vect_cls: Suvw needs one row per frequency.')
      if (allocated(psd)) deallocate(psd)
      if (allocated(bisp)) deallocate(bisp)
      allocate(psd(nf, n2))
      if (icomp_bisp_ /= 0) allocate(bisp(nf, nf, n3))
      if (.not. lp_diag_) then
         if (icomp_bisp_ /= 0) then
            call spectra_(psd, bisp)
         else
            call psdOnly_(psd)
         end if
      else
         call diagPsd_(f, psd)
         if (icomp_bisp_ /= 0) call diagBisp_(bisp, [(i, i = 1, n3)])
      end if
   contains
      subroutine spectra_(ps, bs)
         real(bsa_real_t), target, intent(inout) :: ps(nf, n2), bs(nf, nf, n3)
         call slotCheck_(bsahip_classic_spectra(lp_, int(nfh, c_int32_t), df, 0_c_int32_t, &
                                                c_loc(ps), c_null_ptr, c_loc(bs), c_null_ptr), &
                         'getBFM_vect_cls')
      end subroutine
      subroutine psdOnly_(ps)
         real(bsa_real_t), target, intent(inout) :: ps(nf, n2)
         call slotCheck_(bsahip_classic_spectra(lp_, int(nfh, c_int32_t), df, 0_c_int32_t, &
                                                c_loc(ps), c_null_ptr, c_null_ptr, c_null_ptr), &
                         'getBFM_vect_cls')
      end subroutine
      subroutine diagPsd_(fr, ps)
         real(bsa_real_t), target, intent(in) :: fr(nf)
         real(bsa_real_t), intent(inout) :: ps(nf, n2)
         real(bsa_real_t), allocatable, target :: pt(:, :)
         allocate(pt(n2, nf))
         call slotCheck_(bsahip_diag_psd_points(lp_, int(nf, c_int32_t), c_loc(fr), c_loc(pt)), &
                         'getBFM_vect_cls')
         ps = transpose(pt)
      end subroutine
      subroutine diagBisp_(bs, sel)
         real(bsa_real_t), target, intent(inout) :: bs(nf, nf, n3)
         integer(c_int32_t), target, intent(in) :: sel(n3)
         call slotCheck_(bsahip_classic_bispectra_sel(lp_, int(nfh, c_int32_t), df, BSAHIP_ONLY_DIAG, &
                                                      int(n3, c_int32_t), c_loc(sel), c_loc(bs), &
                                                      c_null_ptr), 'getBFM_vect_cls')
      end subroutine
   end subroutine

   !> getRM_full_vect_cls_ (functionsImpl.F90:1598-1738) or, ONLY_DIAG,
   !> getRM_diag_vect_cls_ (:1975-2101): psd times Re[H_p conj H_n] at f (|H_m|^2),
   !> bisp(i, j, e) times Re[H_m(f_i) H_n(f_j) conj H_p(f_i + f_j)] (m = n = p
   !> for the diagonal model), in place; no state kept from one call to the next.
   subroutine opBRMVectCls_(f, psd, bisp)
      real(bsa_real_t), intent(in) :: f(:)
      real(bsa_real_t), allocatable, intent(inout) :: psd(:, :), bisp(:, :, :)
      integer(bsa_int_t) :: nf, n2, n3
      integer(c_int32_t) :: fl
      nf = size(f)
      call slotDims_(n2, n3, fl)
      if (allocated(psd)) then
         if (size(psd, 1) /= nf .or. size(psd, 2) /= n2) call bsa_Abort('getBRM_vect_cls: psd is not (NF, NM^2) (NM: ONLY_DIAG).')
         call frfPsd_(f, psd)
      end if
      if (allocated(bisp) .and. icomp_bisp_ /= 0) then
         if (size(bisp, 1) /= nf .or. size(bisp, 2) /= nf .or. size(bisp, 3) /= n3) &
            call bsa_Abort('getBRM_vect_cls: bisp is not (NF, NF, NM^3) (NM: ONLY_DIAG).')
         call frfGrid_(f, bisp)
      end if
   contains
      subroutine frfPsd_(fr, ps)
         real(bsa_real_t), target, intent(in) :: fr(nf)
         real(bsa_real_t), target, intent(inout) :: ps(nf, n2)
         call slotCheck_(bsahip_frf_apply(lp_, fl, BSAHIP_FRF_PSD, int(nf, c_int64_t), &
                                          c_loc(fr), 0_c_int64_t, c_null_ptr, c_loc(ps)), &
                         'getBRM_vect_cls')
      end subroutine
      subroutine frfGrid_(fr, bs)
         real(bsa_real_t), target, intent(in) :: fr(nf)
         real(bsa_real_t), target, intent(inout) :: bs(nf, nf, n3)
         call slotCheck_(bsahip_frf_apply(lp_, fl, BSAHIP_FRF_GRID, int(nf, c_int64_t), &
                                          c_loc(fr), int(nf, c_int64_t), c_loc(fr), c_loc(bs)), &
                         'getBRM_vect_cls')
      end subroutine
   end subroutine

   !> getFM_full_tnm_scalar_cls_ (functionsImpl.F90:2122-2318) or, ONLY_DIAG,
   !> getFM_diag_tnlm_scalar_cls_ (:2460-2563) at (fi, fj), a point of the grid
   !> of the last Classic run: bisp(E3) with S(fi + fj) zero beyond that grid,
   !> psd(E2) at fi when ij = 1 and zero otherwise.
   subroutine opBFMScalarCls_(ii, ij, fi, fj, Suvw, Suvw_pad, psd, bisp)
      integer(bsa_int_t), intent(in) :: ii, ij
      real(bsa_real_t), intent(in) :: fi, fj
      real(bsa_real_t), intent(in) :: Suvw(:, :), Suvw_pad(:)
      real(bsa_real_t), intent(inout) :: psd(:), bisp(:)
      integer(bsa_int_t) :: n2, n3
      integer(c_int32_t) :: fl
      call slotDims_(n2, n3, fl)
      if (size(psd) < n2 .or. size(bisp) < n3) &
         call bsa_Abort('getBFM_scalar_cls: psd needs NM^2 and bisp NM^3 entries (NM: ONLY_DIAG).')
      if (size(Suvw, 2) /= size(Suvw_pad) .or. ii < 1 .or. ij < 1) &
         call bsa_Abort('getBFM_scalar_cls: bad grid indexes or Suvw / Suvw_pad shapes.')
      psd = 0
      bisp = 0
      if (icomp_bisp_ /= 0) call point_(bisp)
      if (ij == 1) call psdAt_(psd)
   contains
      subroutine point_(bs)
         real(bsa_real_t), target, intent(inout) :: bs(n3)
         real(bsa_real_t), target :: a(1), b(1)
         a(1) = fi
         b(1) = fj
         call slotCheck_(bsahip_bfm_points_cls(lp_, int(lp_nfh_, c_int32_t), lp_df_, fl, 1_c_int32_t, &
                                               c_loc(a), c_loc(b), c_loc(bs), c_null_ptr), &
                         'getBFM_scalar_cls')
      end subroutine
      subroutine psdAt_(ps)
         real(bsa_real_t), target, intent(inout) :: ps(n2)
         real(bsa_real_t), target :: a(1)
         a(1) = fi
         if (lp_diag_) then
            call slotCheck_(bsahip_diag_psd_points(lp_, 1_c_int32_t, c_loc(a), c_loc(ps)), &
                            'getBFM_scalar_cls')
         else
            call slotCheck_(bsahip_modal_psd_points(lp_, 1_c_int32_t, c_loc(a), c_loc(ps)), &
                            'getBFM_scalar_cls')
         end if
      end subroutine
   end subroutine

   !> getRM_full_scalar_cls_ (functionsImpl.F90:2329-2450) or, ONLY_DIAG,
   !> getRM_diag_scalar_cls_ (:2570-2637): psdout at fi when ij = 1 (zero
   !> otherwise), bispout at (fi, fj), each psdin / bispin times the FRF
   !> products of getBRM_vect_cls.
   subroutine opBRMScalarCls_(ii, ij, fi, fj, psdin, psdout, bispin, bispout)
      integer(bsa_int_t), intent(in) :: ii, ij
      real(bsa_real_t), intent(in) :: fi, fj
      real(bsa_real_t), intent(in) :: psdin(:), bispin(:)
      real(bsa_real_t), intent(out) :: psdout(:), bispout(:)
      integer(bsa_int_t) :: n2, n3
      integer(c_int32_t) :: fl
      call slotDims_(n2, n3, fl)
      if (size(psdin) < n2 .or. size(psdout) < n2 .or. size(bispin) < n3 .or. &
          size(bispout) < n3 .or. ii < 1 .or. ij < 1) &
         call bsa_Abort('getBRM_scalar_cls: psd needs NM^2 and bisp NM^3 entries (NM: ONLY_DIAG).')
      psdout = 0
      bispout = 0
      if (ij == 1) then
         psdout(1:n2) = psdin(1:n2)
         call frfPsd_(psdout)
      end if
      if (icomp_bisp_ /= 0) then
         bispout(1:n3) = bispin(1:n3)
         call frfPt_(bispout)
      end if
   contains
      subroutine frfPsd_(ps)
         real(bsa_real_t), target, intent(inout) :: ps(n2)
         real(bsa_real_t), target :: a(1)
         a(1) = fi
         call slotCheck_(bsahip_frf_apply(lp_, fl, BSAHIP_FRF_PSD, 1_c_int64_t, c_loc(a), &
                                          0_c_int64_t, c_null_ptr, c_loc(ps)), 'getBRM_scalar_cls')
      end subroutine
      subroutine frfPt_(bs)
         real(bsa_real_t), target, intent(inout) :: bs(n3)
         real(bsa_real_t), target :: a(1), b(1)
         a(1) = fi
         b(1) = fj
         call slotCheck_(bsahip_frf_apply(lp_, fl, BSAHIP_FRF_POINTS, 1_c_int64_t, c_loc(a), &
                                          1_c_int64_t, c_loc(b), c_loc(bs)), 'getBRM_scalar_cls')
      end subroutine
   end subroutine

   !> getBFM_msh (functionsImpl.F90:247-465, the POD form at :582-1020 when
   !> isvd = 1; ONLY_DIAG: getFM_diag_tnm_scalar_msh_, :1123-1216): bfm(E3, npts)
   !> at the points (fi(q), fj(q)), S(fi + fj) evaluated exactly.
   subroutine opBFMMsh_(bfm, fi, fj)
      real(bsa_real_t), intent(inout), contiguous :: bfm(:, :)
      real(bsa_real_t), intent(in), contiguous :: fi(:), fj(:)
      integer(bsa_int_t) :: n, n2, n3
      integer(c_int32_t) :: fl
      call slotDims_(n2, n3, fl)
      n = size(fi)
      if (size(fj) /= n .or. size(bfm, 1) /= n3 .or. size(bfm, 2) < n) &
         call bsa_Abort('getBFM_msh: bfm must be (NM^3, npts) (NM: ONLY_DIAG) with one fj per fi.')
      if (n == 0) return
      call pts_(bfm, fi, fj)
   contains
      subroutine pts_(b, a1, a2)
         real(bsa_real_t), target, intent(inout) :: b(n3, n)
         real(bsa_real_t), target, intent(in) :: a1(n), a2(n)
         call slotCheck_(bsahip_bfm_points(lp_msh_, int(n, c_int32_t), c_loc(a1), c_loc(a2), &
                                           fl, c_loc(b), c_null_ptr), 'getBFM_msh')
      end subroutine
   end subroutine

   !> getBRM_msh (functionsImpl.F90:1031-1115; ONLY_DIAG: getRM_diag_scalar_msh_,
   !> :1224-1286): brm = bfm times Re[H_m(fi) H_n(fj) conj H_p(fi + fj)] at each point.
   subroutine opBRMMsh_(brm, fi, fj, bfm)
      real(bsa_real_t), intent(inout), contiguous :: brm(:, :)
      real(bsa_real_t), intent(in), contiguous :: fi(:), fj(:)
      real(bsa_real_t), intent(in), contiguous :: bfm(:, :)
      integer(bsa_int_t) :: n, n2, n3
      integer(c_int32_t) :: fl
      call slotDims_(n2, n3, fl)
      n = size(fi)
      if (size(fj) /= n .or. size(brm, 1) /= n3 .or. size(bfm, 1) /= n3 .or. &
          size(brm, 2) < n .or. size(bfm, 2) < n) &
         call bsa_Abort('getBRM_msh: brm and bfm must be (NM^3, npts) (NM: ONLY_DIAG) with one fj per fi.')
      if (n == 0) return
      brm(:, 1:n) = bfm(:, 1:n)
      call frf_(brm, fi, fj)
   contains
      subroutine frf_(b, a1, a2)
         real(bsa_real_t), target, intent(inout) :: b(n3, n)
         real(bsa_real_t), target, intent(in) :: a1(n), a2(n)
         call slotCheck_(bsahip_frf_apply(lp_, fl, BSAHIP_FRF_POINTS, int(n, c_int64_t), &
                                          c_loc(a1), int(n, c_int64_t), c_loc(a2), c_loc(b)), &
                         'getBRM_msh')
      end subroutine
   end subroutine

   ! ======================================================================
   !  post-processing
   ! ======================================================================

   !> src/bsa/BsaLibImpl.F90:1292-1421: background m2mf(m,m)/K_m^2 and resonant
   !> pi * 2 pi * f_m * S_Fm(f_m) / (2 xsi_m K_m^2) parts of each kept mode's
   !> variance, with S_Fm the modal force PSD at the natural frequency
   !> (getBR_SFm_val_, functionsImpl.F90:2650-2743), evaluated on the GPU.
   !> Needs a previous bsa_Run (its problem) and the damping ratios.
   subroutine bsa_computeBRdecomp(m2mf, bkg, res)
      real(bsa_real_t), intent(in) :: m2mf(:)
      real(bsa_real_t), allocatable, intent(out) :: bkg(:), res(:)
      real(bsa_real_t), allocatable, target :: f(:), psd(:)
      real(bsa_real_t), parameter :: pi_ = acos(-1.0_bsa_real_t)
      real(bsa_real_t) :: k2, sfm
      integer(bsa_int_t) :: nm, im, m
      integer(c_int) :: rc
      if (.not. lp_ready_) call bsa_Abort('bsa_computeBRdecomp needs a previous bsa_Run.')
      if (.not. allocated(xsi_)) call bsa_Abort('BR decomposition needs the modal damping ratios.')
      nm = size(modes_)
      if (size(m2mf) < nm * nm) call bsa_Abort('m2mf has fewer than NM*NM entries.')
      allocate(bkg(nm), res(nm))
      f = natf_(modes_)
      allocate(psd(nm * nm * nm))
      rc = bsahip_modal_psd_points(lp_, int(nm, c_int32_t), c_loc(f), c_loc(psd))
      if (rc /= 0) call bsa_Abort('bsahip_modal_psd_points failed: '//bsahip_last_error_f())
      do im = 1, nm
         m = modes_(im)
         k2 = Kg_(m) * Kg_(m)
         bkg(im) = m2mf((im - 1) * nm + im) / k2
         sfm = psd((im - 1) * nm * nm + (im - 1) * nm + im)
         res(im) = pi_ * 2 * pi_ * f(im) * sfm / (2 * xsi_(m) * k2)
      end do
   end subroutine

   !> src/bsa/BsaLibImpl.F90:1429-1519: Gaussian peak factor (Davenport) and,
   !> given a skewness, the non-Gaussian ones of the Hermite model (Kwon and
   !> Kareem), with the reference's early returns and warnings.
   subroutine bsa_computePeakFactors(m2, m2o2, obs_time, peak_g, sk, peak_ng_pos, peak_ng_neg)
      real(bsa_real_t), intent(in) :: m2(:), m2o2(:)
      real(bsa_real_t), intent(in) :: obs_time
      real(bsa_real_t), allocatable, intent(inout) :: peak_g(:)
      real(bsa_real_t), intent(in), allocatable :: sk(:)
      real(bsa_real_t), allocatable, intent(inout) :: peak_ng_pos(:)
      real(bsa_real_t), allocatable, intent(inout), optional :: peak_ng_neg(:)
      real(bsa_real_t), parameter :: gamma_ = 0.5772d0
      real(bsa_real_t), parameter :: pi_ = acos(-1.0_bsa_real_t)
      real(bsa_real_t), allocatable :: beta(:), g4(:), h40(:), h3(:), h4(:), k(:)
      real(bsa_real_t), allocatable :: b2(:), b3(:), r(:), neg(:)
      if (all(m2o2 == 0)) then
         write(output_unit, '(/ 1x, 2a)') WARNMSG, '"m2_ord2"  is null. Cannot compute extremes. Skipping.'
         return
      end if
      beta = sqrt(m2o2 / m2) / (2 * pi_) * obs_time
      if (any(beta < 1.d0)) then
         write(output_unit, '(1x, a, f7.2, a)') WARNMSG, 'Observation time of ', obs_time, ' sec.  is too short.'
         return
      end if
      beta = sqrt(2 * log(beta))
      peak_g = gamma_ / beta + beta
      if (.not. allocated(sk)) return
      g4 = 1.5d0 * sk * sk
      h40 = sqrt(1.d0 + 1.5d0 * g4)
      h3 = sk / (4.d0 + 2.d0 * h40)
      h4 = (h40 - 1.d0) / 18.0d0
      k = 1.d0 / sqrt(1 + 2.d0 * h3 * h3 + 6.d0 * h4 * h4)
      b2 = beta * beta
      b3 = b2 * beta
      peak_ng_pos = 5.44d0 / b3
      peak_ng_pos = peak_ng_pos + 3.d0 / beta * (pi_ * pi_ / 6.d0 - gamma_ + gamma_**2)
      peak_ng_pos = peak_ng_pos + (b3 + 3 * beta * (gamma_ - 1))
      peak_ng_pos = peak_ng_pos * h4
      neg = peak_ng_pos
      r = h3 * (b2 + 2.d0 * gamma_ - 1.d0 + 1.98d0 / b2)
      peak_ng_pos = (peak_ng_pos + r + peak_g) * k
      if (present(peak_ng_neg)) peak_ng_neg = (neg - r + peak_g) * k
   end subroutine

   !> Addition (the reference does this in its application, app/bsa.F90:935-1053):
   !> nodal second and third moments and skewness from the modal ones, by the
   !> unilateral combination (diagonal modal terms) and the complete cubic
   !> combination, on the GPU.  phi(ndofs, nm) holds the kept modes' shapes at
   !> the DOFs of interest; m2o2mr may be unallocated.
   subroutine bsa_modalRecombination(phi, m2mr, m3mr, m2o2mr, m2_diag, m3_diag, m2o2_diag, &
                                     sk_diag, m2_full, m3_full, m2o2_full, sk_full)
      real(bsa_real_t), intent(in) :: phi(:, :), m2mr(:), m3mr(:)
      real(bsa_real_t), intent(in), allocatable :: m2o2mr(:)
      real(bsa_real_t), allocatable, target, intent(out) :: m2_diag(:), m3_diag(:), m2o2_diag(:), &
         sk_diag(:), m2_full(:), m3_full(:), m2o2_full(:), sk_full(:)
      real(bsa_real_t), allocatable, target :: ph(:, :), c2(:), c3(:), c22(:)
      integer(bsa_int_t) :: nd, nm
      type(c_ptr) :: p22, o22
      integer(c_int) :: rc
      nd = size(phi, 1)
      nm = size(phi, 2)
      if (size(m2mr) /= nm * nm .or. size(m3mr) /= nm**3) &
         call bsa_Abort('Modal recombination needs NM*NM second and NM^3 third moments.')
      ph = phi                   ! contiguous copies for the C ABI
      c2 = m2mr
      c3 = m3mr
      allocate(m2_diag(nd), m3_diag(nd), m2o2_diag(nd), sk_diag(nd))
      allocate(m2_full(nd), m3_full(nd), m2o2_full(nd), sk_full(nd))
      m2o2_diag = 0
      m2o2_full = 0
      p22 = c_null_ptr
      if (allocated(m2o2mr)) then
         c22 = m2o2mr
         p22 = c_loc(c22)
      end if
      o22 = c_null_ptr
      if (allocated(m2o2mr)) o22 = c_loc(m2o2_diag)
      rc = bsahip_modal_recombination(int(nd, c_int32_t), int(nm, c_int32_t), c_loc(ph), c_loc(c2), &
         c_loc(c3), p22, 0_c_int32_t, c_loc(m2_diag), c_loc(m3_diag), o22, c_loc(sk_diag))
      if (rc /= 0) call bsa_Abort('bsahip_modal_recombination failed: '//bsahip_last_error_f())
      if (allocated(m2o2mr)) o22 = c_loc(m2o2_full)
      rc = bsahip_modal_recombination(int(nd, c_int32_t), int(nm, c_int32_t), c_loc(ph), c_loc(c2), &
         c_loc(c3), p22, 1_c_int32_t, c_loc(m2_full), c_loc(m3_full), o22, c_loc(sk_full))
      if (rc /= 0) call bsa_Abort('bsahip_modal_recombination failed: '//bsahip_last_error_f())
   end subroutine

   ! ======================================================================
   !  settings kept for the input files (the device path does not read them)
   ! ======================================================================

   subroutine bsa_setVersion(ivers)
      integer(bsa_int_t), value :: ivers
      ivers_ = ivers
   end subroutine

   !> Tensor symmetries: the device path computes the unique G values and
   !> mirrors them, whatever the setting; the value is kept for the input files.
   subroutine bsa_setSpectraSymmetries(ispctrsym)
      integer(bsa_int_t), value :: ispctrsym
      ispctrsym_ = ispctrsym
   end subroutine

   subroutine bsa_setNodalCoords(nn, coords)
      integer(bsa_int_t), value :: nn
      real(bsa_real_t), target, contiguous :: coords(:, :)
      if (size(coords, 2) /= nn) call bsa_Abort('Nodal coordinates must be (3, NN).')
      coords_ = coords
   end subroutine

   subroutine bsa_setWindVertProf(iwprof)
      integer(bsa_int_t), value :: iwprof
      iwprof_ = iwprof
   end subroutine

   subroutine bsa_setWindAltDir(ivert)
      integer(bsa_int_t), value :: ivert
      ivert_ = ivert
   end subroutine

   subroutine bsa_setAirDensity(aird)
      real(bsa_real_t), value :: aird
      aird_ = aird
   end subroutine

   subroutine bsa_setGlobalRotMatW2G(rotW2G)
      real(bsa_real_t), intent(in) :: rotW2G(3, 3)
      rotW2G_ = rotW2G
   end subroutine

   subroutine bsa_setWZRefAlt(Zref)
      real(bsa_real_t), target, intent(in) :: Zref(:)
      zref_ = Zref
   end subroutine

   !> Assumed-size in the reference interface: kept as addresses and read when
   !> the number of wind zones is known (as L and sigma).
   subroutine bsa_setWindCorrCoeffs(ccoeffs)
      real(bsa_real_t), target, intent(in) :: ccoeffs(3, 3, *)
      ccoef_addr_ = c_loc(ccoeffs)
   end subroutine

   subroutine bsa_setWindCorrExpnts(cexpn)
      real(bsa_real_t), target, intent(in) :: cexpn(3, 3, *)
      cexp_addr_ = c_loc(cexpn)
   end subroutine

   subroutine bsa_setWZRotMatW2G(rotW2G_L)
      real(bsa_real_t), target, intent(in) :: rotW2G_L(3, 3, *)
      rotlw_addr_ = c_loc(rotW2G_L)
   end subroutine

   subroutine bsa_setIncidenceAngles(incang)
      real(bsa_real_t), target, intent(in) :: incang(:)
      incang_ = incang
   end subroutine

   subroutine bsa_setNodalWindAltitudes(WnodAlt)
      real(bsa_real_t), target, intent(in) :: WnodAlt(:)
      walt_ = WnodAlt
   end subroutine

   !> Mesher options of the reference's zone refinement (multi-level
   !> refinement, zone-delta validation) and the switch that forces Classic
   !> under the GPU flag: accepted; the device Mesher has its own refinement
   !> and the Classic path always runs on the device.
   subroutine bsa_setBfmMLR(bool)
      logical, intent(in) :: bool
   end subroutine

   subroutine bsa_doValidateZoneDeltas(bool)
      logical, intent(in) :: bool
   end subroutine

   subroutine bsa_forceBsaClsExecution(bool)
      logical, intent(in) :: bool
   end subroutine

   !> src/bsa/BsaLibImpl.F90:126-131: bsa_Run writes bsa.bsadata and
   !> bsa.extdata in the current directory, then runs unless run is false.
   subroutine bsa_generateBSAInputFiles(run)
      logical, value :: run
      gen_inputs_ = .true.
      run_after_gen_ = run
   end subroutine

   !> The two input files of the standalone program, in the layouts its
   !> reader expects (bsa.bsadata: labelled sections, one value per line;
   !> bsa.extdata: stream of int32 / real64 arrays), from the setters' state.
   !> Reals are written with full precision.  Data never set (coordinates,
   !> reference altitudes, ...) is written as zeros of the right size.
   subroutine generateInputFiles_()
      integer :: iun, ios, i, nwz, nnl, nlibsl, ncorr, nd
      real(bsa_real_t), pointer :: p3(:, :, :), p2(:, :)
      real(bsa_real_t), allocatable :: corr3(:, :), z3(:, :, :), xs(:)
      integer(bsa_int_t), allocatable :: tcs(:), drs(:)
      character(len=*), parameter :: fi = '(i8)', fr = '(es25.17e3)'

      if (allocated(tc_)) then
         tcs = tc_
      else
         tcs = [1]
      end if
      if (allocated(dirs_)) then
         drs = dirs_
      else
         drs = [1]
      end if
      open(newunit=iun, file=BSA_DATA_FNAME, iostat=ios, form='formatted', status='replace', &
           action='write')
      if (ios /= 0) call bsa_Abort('Error opening file  '//BSA_DATA_FNAME)
      write(iun, '(a)') '-GENERAL:'
      write(iun, fi) isuban_
      write(iun, fi) ivers_
      write(iun, fi) iconv_
      write(iun, fi) icomp_psd_
      write(iun, fi) icomp_bisp_
      write(iun, fi) ionly_diag_
      write(iun, fi) isym_
      write(iun, fi) ispctrsym_
      write(iun, fi) itest_
      write(iun, '(a)') '-CLASSIC:'
      write(iun, fi) iclassic_
      write(iun, fi) nfreqs_
      write(iun, fr) df_
      write(iun, '(a)') '-MESHER:'
      write(iun, fi) isvd_
      write(iun, fi) bkgrfmt_
      write(iun, fr) bkgaext_
      write(iun, fr) genpaext_
      write(iun, fr) maxaext_
      write(iun, fi) ifcov_
      write(iun, fi) idumpmod_
      write(iun, '(a)') '-DIRECTIONS:'
      write(iun, fi) size(drs)
      do i = 1, size(drs)
         write(iun, fi) drs(i)
      end do
      write(iun, '(a)') '-TURBULENCE:'
      write(iun, fi) size(tcs)
      do i = 1, size(tcs)
         write(iun, fi) tcs(i)
      end do
      close(iun)

      if (.not. (allocated(n_load_) .and. allocated(phi_) .and. allocated(wfc_) .and. &
                 allocated(corr_) .and. allocated(u_node_) .and. allocated(wz_node_))) &
         call bsa_Abort('Input files need the structure, modal and wind data to be set.')
      if (.not. (c_associated(tscl_addr_) .and. c_associated(tstd_addr_))) &
         call bsa_Abort('Input files need the turbulence scales and standard deviations.')
      if (nz_lim_ > 0) then
         nwz = nz_lim_
      else if (allocated(umref_)) then
         nwz = size(umref_)
      else
         nwz = maxval(wz_node_)
      end if
      nnl = size(n_load_)
      nlibsl = 1
      if (allocated(libs_load_)) nlibsl = size(libs_load_)
      ncorr = (nn_ * nn_ + nn_) / 2
      allocate(corr3(ncorr, 3), source=0.0_bsa_real_t)
      corr3(:, 1:min(3, size(corr_, 2))) = corr_(:, 1:min(3, size(corr_, 2)))
      nd = size(phi_, 1)
      open(newunit=iun, file=EXT_DATA_FNAME, iostat=ios, form='unformatted', access='stream', &
           status='replace', action='write')
      if (ios /= 0) call bsa_Abort('Error opening file  '//EXT_DATA_FNAME)
      write(iun) int(nn_, int32), int(nlibs_, int32), int(nnl, int32), int(nlibsl, int32)
      write(iun) int(n_load_, int32)
      if (allocated(libs_load_)) then
         write(iun) int(libs_load_, int32)
      else
         write(iun) 1_int32
      end if
      if (allocated(coords_)) then
         write(iun) coords_
      else
         write(iun) spread(spread(0.0_bsa_real_t, 1, 3), 2, nn_)
      end if
      write(iun) int(iwprof_, int32), int(ipsd_, int32), int(ivert_, int32), &
                 int(size(wfc_, 2) - 3, int32)
      write(iun) aird_
      write(iun) rotW2G_
      write(iun) int(nwz, int32)
      call zoneVector_(zref_)
      call zoneVector_(umref_)
      call c_f_pointer(tscl_addr_, p3, [3, 3, nwz])
      write(iun) p3
      call c_f_pointer(tstd_addr_, p2, [3, nwz])
      write(iun) p2
      allocate(z3(3, 3, nwz), source=0.0_bsa_real_t)
      call zoneTensor_(ccoef_addr_)
      call zoneTensor_(cexp_addr_)
      if (allocated(wz_lims_)) then
         write(iun) wz_lims_
      else
         write(iun) spread(0.0_bsa_real_t, 1, nwz + 1)
      end if
      call zoneTensor_(rotlw_addr_)
      call zoneVector_(incang_)
      write(iun) int(wz_node_, int32)
      if (allocated(walt_)) then
         write(iun) walt_
      else
         write(iun) spread(0.0_bsa_real_t, 1, nn_)
      end if
      write(iun) u_node_
      write(iun) corr3
      write(iun) wfc_
      write(iun) int(nd, int32), int(nm_, int32)
      write(iun) natf_
      write(iun) phi_
      write(iun) Mg_
      write(iun) Kg_
      write(iun) Cg_
      if (allocated(xsi_)) then
         xs = xsi_
      else
         allocate(xs(nm_), source=0.0_bsa_real_t)
      end if
      write(iun) xs
      close(iun)
   contains
      subroutine zoneVector_(v)
         real(bsa_real_t), allocatable, intent(in) :: v(:)
         if (allocated(v)) then
            if (size(v) /= nwz) call bsa_Abort('Per-zone array size differs from the number of wind zones.')
            write(iun) v
         else
            write(iun) spread(0.0_bsa_real_t, 1, nwz)
         end if
      end subroutine
      subroutine zoneTensor_(addr)
         type(c_ptr), intent(in) :: addr
         real(bsa_real_t), pointer :: t3(:, :, :)
         if (c_associated(addr)) then
            call c_f_pointer(addr, t3, [3, 3, nwz])
            write(iun) t3
         else
            write(iun) z3
         end if
      end subroutine
   end subroutine

   ! ======================================================================
   !  exports and run-time I/O settings (src/IO/BsaIO.F90, BsaLibImpl.F90:1564-1760)
   ! ======================================================================

   subroutine bsa_printBSAHeader()
      write(output_unit, '(/ 1x, a)') 'BsaLib -- bispectral stochastic analysis, MI355X device path'
      write(output_unit, '(1x, a /)') '          (bsalib_amd: modal force and response bispectra on the GPU)'
   end subroutine

   !> Exports go to the current directory (the default here).
   subroutine bsa_setExportInCurrDir()
      if (allocated(exp_dir_)) deallocate(exp_dir_)
   end subroutine

   subroutine bsa_setExportDirectory(dirname)
      character(len=*), intent(in) :: dirname
      exp_dir_ = trim(dirname)
      if (len(exp_dir_) > 0) then
         if (exp_dir_(len(exp_dir_):) /= '/') exp_dir_ = exp_dir_//'/'
      end if
   end subroutine

   !> Units are closed after every export here, so this only exists for callers.
   subroutine bsa_closeUnitsAtEnd()
   end subroutine

   subroutine bsa_setExportFileFormat(iform)
      integer(bsa_int_t), value :: iform
      if (iform == BSA_EXPORT_FORMAT_FORMATTED) exp_binary_ = .false.
      if (iform == BSA_EXPORT_FORMAT_UNFORMATTED) exp_binary_ = .true.
   end subroutine

   subroutine bsa_setExportAppendMode(imode)
      integer(bsa_int_t), value :: imode
      if (imode == BSA_EXPORT_MODE_REPLACE) exp_append_ = .false.
      if (imode == BSA_EXPORT_MODE_APPEND) exp_append_ = .true.
   end subroutine

   !> Accepted; the library writes no log file of its own.
   subroutine bsa_setOutFileName(fname)
      character(len=*), intent(in) :: fname
   end subroutine

   !> BSA_EXPORT_BRM_MODE_BASE (and any other value, with a warning) makes a
   !> Mesher run write every bispectrum entry at every post-mesh point to
   !> BRM_EXPORT_FNAME_MSH; NONE leaves the setting as it is
   !> (src/bsa/BsaLibImpl.F90:1917-1932).  The Classic approach exports only in
   !> visual mode, as in the reference.
   subroutine bsa_setBRMExportDefaultMode(imode)
      integer(bsa_int_t), value :: imode
      if (imode == BSA_EXPORT_BRM_MODE_NONE) return
      if (imode /= BSA_EXPORT_BRM_MODE_BASE) &
         write(output_unit, '(1x, 2a)') WARNMSG, 'Unknown BRM export mode. Setting default  (base).'
      brm_base_ = .true.
   end subroutine

   !> Visual mode (src/bsa/BsaLibImpl.F90:119-122): bsa_Run also writes one
   !> entry of the response bispectrum, the modal entry or the nodal bispectrum
   !> given by bsa_setVisualIndexes: over the whole Classic plane to
   !> BRM_EXPORT_FNAME_CLS (modal entries only, as the reference), at the Mesher's
   !> post-mesh points to BRM_EXPORT_FNAME_MSH / BRN_EXPORT_FNAME_MSH.  The
   !> reference's Mesher re-reads the BFM dump of an earlier run for it; the
   !> device Mesher evaluates its pre-mesh in the same run instead.
   subroutine bsa_enableVisualMode()
      visual_ = .true.
   end subroutine

   !> Accepted values as src/bsa/BsaLibImpl.F90:813-827.  The device Mesher lays
   !> out its zones the same way in either mode.
   subroutine bsa_setPremeshMode(imode)
      integer(bsa_int_t), value :: imode
      if (imode == BSA_PREMESH_MODE_BASE) then
         premesh_mode_ = imode
      else
         if (imode /= BSA_PREMESH_MODE_ZONE_REFINED) then
            write(output_unit, '(1x, 2a, i0, a)') WARNMSG, 'Invalid   ', imode, '  pre-meshing mode value.'
            write(output_unit, '(1x, 2a)') MSGCONT, 'Setting default (ZONE REFINED).'
         end if
         premesh_mode_ = BSA_PREMESH_MODE_ZONE_REFINED
      end if
   end subroutine

   !> The Mesher stops after its pre-mesh (src/bsa/BsaLibImpl.F90:146-148,
   !> BsaMesherImpl.F90:113-119): the level-0 BFM of every zone is evaluated on
   !> the device and no post-mesh runs, so m3mf_msh and m3mr_msh stay
   !> unallocated.
   subroutine bsa_enableOnlyPremesh()
      only_premesh_ = .true.
   end subroutine

   !> Extension (no counterpart in the reference): target error of the
   !> Mesher's m3mr relative to its largest entry; the refinement level of
   !> every tile is chosen a posteriori for it (bsahip_mesher_adapt).  0 keeps
   !> the rule-based levels of the policy factors.  Default 1e-3.
   subroutine bsa_setMesherTolerance(tol)
      real(bsa_real_t), value :: tol
      if (tol < 0.0_bsa_real_t) call bsa_Abort('The Mesher tolerance must be >= 0.')
      mesher_tol_ = tol
   end subroutine

   !> Accepted for interface compatibility (src/bsa/BsaLibImpl.F90:847-865):
   !> the device Mesher's rule-based levels use the built-in policy factors.
   subroutine bsa_setPolicyIDValidationValues(id, i_bfm, j_bfm, i_brm, j_brm)
      integer(int32), value :: id, i_bfm, j_bfm, i_brm, j_brm
      if (id < 0 .or. id > 9) write(output_unit, '(1x, 2a, i0)') WARNMSG, 'Invalid policy ID  ', id
   end subroutine

   !> .true. unless the ONLY_DIAG model is selected (src/bsa/BsaLibImpl.F90:902-906).
   logical function bsa_isFullComp()
      bsa_isFullComp = ionly_diag_ == 0
   end function

   !> Output directory (src/bsa/BsaLibImpl.F90:706-710): the exports go there
   !> unless bsa_setExportDirectory or bsa_setExportInCurrDir chose otherwise.
   subroutine bsa_setOutputDirectory(dirname)
      character(len=*), intent(in) :: dirname
      if (.not. allocated(exp_dir_)) call bsa_setExportDirectory(dirname)
   end subroutine

   !> The reference's debug-log unit (src/bsa/BsaLibImpl.F90:726-730); this
   !> library writes no debug log, so the unit is not used.
   subroutine bsa_setOutUnit(iunit)
      integer(bsa_int_t), value :: iunit
   end subroutine

   !> PSD series to a text file as src/bsa/BsaLibImpl.F90:1980-2012: the two
   !> sizes, then one row per frequency (f first when given).
   subroutine bsa_exportPSDToFile(fname, psd, f)
      character(len=*), intent(in) :: fname
      real(bsa_real_t), intent(in) :: psd(:, :)
      real(bsa_real_t), intent(in), optional :: f(:)
      integer :: iun, j
      iun = openText_(fname)
      write(iun, *) size(psd, 1)
      write(iun, *) size(psd, 2)
      do j = 1, size(psd, 1)
         if (present(f)) then
            write(iun, '(*(g0, 1x))') f(j), psd(j, :)
         else
            write(iun, '(*(g0, 1x))') psd(j, :)
         end if
      end do
      close(iun)
   end subroutine

   !> Bispectrum series to a text file as src/bsa/BsaLibImpl.F90:2017-2040: the
   !> three sizes, then bisp(:, i, j) as one row per (i, j).
   subroutine bsa_exportBispToFile(fname, bisp)
      character(len=*), intent(in) :: fname
      real(bsa_real_t), intent(in) :: bisp(:, :, :)
      integer :: iun, i, j
      if (size(bisp, 2) /= size(bisp, 1)) call bsa_Abort('First two dimensions of bisp do not match.')
      iun = openText_(fname)
      write(iun, *) size(bisp, 1)
      write(iun, *) size(bisp, 2)
      write(iun, *) size(bisp, 3)
      do j = 1, size(bisp, 3)
         do i = 1, size(bisp, 2)
            write(iun, '(*(g0, 1x))') bisp(:, i, j)
         end do
      end do
      close(iun)
   end subroutine

   !> bsa_Run writes modal.txt (nodes, DOFs per node, modes, then per mode its
   !> natural frequency and shape) as src/bsa/BsaLibImpl.F90:1972-1974, 2045-2057.
   subroutine bsa_exportModalData()
      export_modal_ = .true.
   end subroutine

   !> A user routine receives the bispectra of visual mode and of
   !> bsa_setBRMExportDefaultMode in place of the export files: per zone (the
   !> Classic plane, or a Mesher tile) the point frequencies f1, f2 and the
   !> values brm(:, point), with a null pdata.  The reference resets its
   !> callback at every bsa_Run (BsaLibImpl.F90:182-188), so there the user
   !> routine never runs.
   subroutine bsa_setBispExportCallback(fptr)
      procedure(exportInterf_vect_), pointer, intent(in) :: fptr
      brm_cb_ => fptr
   end subroutine

   !> The user callback on one rectangular zone: points (f1axis(i), f2axis(j)),
   !> i fastest, values vals(:, point).
   subroutine callbackZone_(ni, nj, f1axis, f2axis, vals)
      integer, intent(in) :: ni, nj
      real(bsa_real_t), intent(in) :: f1axis(ni), f2axis(nj), vals(:, :)
      real(bsa_real_t), allocatable :: f1(:), f2(:)
      integer :: i, j
      allocate(f1(ni * nj), f2(ni * nj))
      do j = 1, nj
         do i = 1, ni
            f1(i + (j - 1) * ni) = f1axis(i)
            f2(i + (j - 1) * ni) = f2axis(j)
         end do
      end do
      call brm_cb_(f1, f2, vals, null_pdata_)
   end subroutine

   !> modal.txt as src/bsa/BsaLibImpl.F90:2045-2057.
   subroutine exportModalData_()
      integer :: iun, j
      open(newunit=iun, file='modal.txt', form='formatted', action='write', status='replace')
      write(iun, *) nn_
      write(iun, *) nlibs_
      write(iun, *) nm_
      do j = 1, nm_
         write(iun, '(1x, g20.10)') natf_(j)
         write(iun, '(1x, g20.10)') phi_(:, j)
      end do
      close(iun)
   end subroutine

   integer function openText_(fname) result(iun)
      character(len=*), intent(in) :: fname
      character(len=:), allocatable :: path
      integer :: ios
      path = fname
      if (allocated(exp_dir_)) path = exp_dir_//fname
      open(newunit=iun, file=path, form='formatted', status='replace', action='write', iostat=ios)
      if (ios /= 0) call bsa_Abort('Error opening file  '//path)
   end function

   !> Modal (i, j, k) kept-mode positions (modal = .true.), or node and DOF
   !> (modal = .false.) as src/bsa/BsaLibImpl.F90:134-141.
   subroutine bsa_setVisualIndexes(indexes, modal)
      integer(bsa_int_t), intent(in) :: indexes(3)
      logical, value :: modal
      vis_idx_ = indexes
      vis_modal_ = modal
   end subroutine

   !> Bispectrum export of a Mesher run, in the reference's stream layout
   !> (exportBRM_baseHeaderWriter_internal_ / exportBRM_base_internal_,
   !> src/bsa/BsaLibImpl.F90:1849-1913, one zone per device tile):
   !> * visual mode, modal indexes (i, j, k): BRM_EXPORT_FNAME_MSH with header
   !>   3, i, j, k, 1 and the BRM entry (i, j, k) per point;
   !> * visual mode, node and DOF: BRN_EXPORT_FNAME_MSH with header 2, node,
   !>   DOF, 1 and the nodal bispectrum sum_mnp phi_m phi_n phi_p BRM_mnp of that
   !>   DOF per point (getBRN_full_scalar_, BsaLibImpl.F90:1781-1803);
   !> * otherwise (bsa_setBRMExportDefaultMode): BRM_EXPORT_FNAME_MSH with
   !>   header NM, the kept modes, NM^3 and every entry per point
   !>   (BsaLibImpl.F90:235-240).
   !> Then the spatial-symmetry setting, the number of zones and of points; per
   !> zone 1 (rectangle), nI, nJ and per point real32 f1, f2 and the values.
   subroutine exportMesher_(msh, nme)
      type(c_ptr), intent(in) :: msh
      integer(bsa_int_t), intent(in) :: nme
      integer(c_int32_t), allocatable, target :: dims(:, :)
      real(bsa_real_t), allocatable, target :: coef(:, :), fi(:), fj(:), val(:, :)
      real(real32), allocatable :: rec(:, :)
      integer(int32), allocatable :: hidx(:)
      integer(c_int64_t) :: npts
      integer(c_int32_t) :: ncol, nt
      integer(c_int) :: rc
      integer(int64) :: off, q
      integer :: iun, ios, k, m, n, o, d, ncomb, nc
      character(len=:), allocatable :: fname
      type(c_ptr) :: pcoef, plev
      rc = bsahip_mesher_ntiles(msh, nt)
      allocate(dims(2, nt))
      pcoef = c_null_ptr
      ! entries per point: NM^3, or NM with the ONLY_DIAG model, whose visual
      ! index is the mode (BsaLibImpl.F90:387-396)
      nc = nme**3
      if (ionly_diag_ /= 0) nc = nme
      if (visual_ .and. vis_modal_) then
         ncol = 1
         allocate(coef(nc, 1), source=0.0_bsa_real_t)
         if (ionly_diag_ /= 0) then
            if (vis_idx_(1) < 1 .or. vis_idx_(1) > nme) call bsa_Abort('Visual mode index out of the kept modes.')
            coef(vis_idx_(1), 1) = 1
         else
            if (any(vis_idx_ < 1) .or. any(vis_idx_ > nme)) call bsa_Abort('Visual mode indexes out of the kept modes.')
            coef((vis_idx_(3) - 1) * nme * nme + (vis_idx_(2) - 1) * nme + vis_idx_(1), 1) = 1
         end if
         fname = BRM_EXPORT_FNAME_MSH
         hidx = int(vis_idx_, int32)
         ncomb = 1
      else if (visual_) then
         ! nodal bispectrum of one DOF: getBRN_full_scalar_ / getBRN_diag_scalar_
         ! (BsaLibImpl.F90:1760-1803)
         ncol = 1
         d = (vis_idx_(1) - 1) * nlibs_ + vis_idx_(2)
         if (vis_idx_(1) < 1 .or. vis_idx_(2) < 1 .or. d > size(phi_, 1)) &
            call bsa_Abort('Visual mode node / DOF out of range.')
         allocate(coef(nc, 1))
         if (ionly_diag_ /= 0) then
            do m = 1, nme
               coef(m, 1) = phi_(d, modes_(m))**3
            end do
         else
            do o = 1, nme
               do n = 1, nme
                  do m = 1, nme
                     coef(m + (n - 1) * nme + (o - 1) * nme * nme, 1) = &
                        phi_(d, modes_(m)) * phi_(d, modes_(n)) * phi_(d, modes_(o))
                  end do
               end do
            end do
         end if
         fname = BRN_EXPORT_FNAME_MSH
         hidx = int(vis_idx_(1:2), int32)
         ncomb = 1
      else
         ncol = int(nc, c_int32_t)
         fname = BRM_EXPORT_FNAME_MSH
         hidx = int(modes_, int32)
         ncomb = nc
      end if
      if (allocated(coef)) pcoef = c_loc(coef)
      plev = c_null_ptr
      if (allocated(msh_levels_)) plev = c_loc(msh_levels_)
      rc = bsahip_mesher_export(msh, plev, 0_c_int32_t, c_null_ptr, 0_c_int64_t, c_loc(dims), &
                                c_null_ptr, c_null_ptr, c_null_ptr, npts)
      if (rc /= 0) call bsa_Abort('bsahip_mesher_export failed: '//bsahip_last_error_f())
      allocate(fi(npts), fj(npts), val(ncol, npts))
      rc = bsahip_mesher_export(msh, plev, ncol, pcoef, npts, c_loc(dims), c_loc(fi), c_loc(fj), &
                                c_loc(val), npts)
      if (rc /= 0) call bsa_Abort('bsahip_mesher_export failed: '//bsahip_last_error_f())
      if (associated(brm_cb_)) then
         off = 0
         do k = 1, nt
            call brm_cb_(fi(off + 1:off + int(dims(1, k), int64) * dims(2, k)), &
                         fj(off + 1:off + int(dims(1, k), int64) * dims(2, k)), &
                         val(:, off + 1:off + int(dims(1, k), int64) * dims(2, k)), null_pdata_)
            off = off + int(dims(1, k), int64) * dims(2, k)
         end do
         return
      end if
      write(output_unit, '( /, 1x, 4a, i0, a, i0, a )') INFOMSG, 'Exporting  bispectrum  to  ', &
         fname, ' (', nt, ' zones, ', npts, ' points)'
      open(newunit=iun, file=fname, form='unformatted', access='stream', status='replace', &
           action='readwrite', iostat=ios)
      if (ios /= 0) call bsa_Abort('Error opening file  '//fname)
      write(iun) int(size(hidx), int32), hidx, int(ncomb, int32), int(isym_, int32), int(nt, int32), &
         int(npts, int32)
      off = 0
      do k = 1, nt
         write(iun) 1_int32, dims(1, k), dims(2, k)
         allocate(rec(2 + ncol, int(dims(1, k), int64) * dims(2, k)))
         do q = 1, size(rec, 2, int64)
            rec(1, q) = real(fi(off + q), real32)
            rec(2, q) = real(fj(off + q), real32)
            rec(3:, q) = real(val(:, off + q), real32)
         end do
         write(iun) rec
         off = off + size(rec, 2, int64)
         deallocate(rec)
      end do
      close(iun)
   end subroutine

   !> Visual-mode export of the Classic approach (BsaClassicImpl.F90:158-194
   !> with exportBRM_baseHeaderWriter_internal_ / exportBRM_base_internal_ and
   !> writeNTotPtsToBispUn_, BsaLibImpl.F90:1832-1913): an unformatted stream of
   !> int32 header values -- 3, the three visual indexes, 1 (combinations), the
   !> spatial symmetry setting, 1 (zones), NF^2 (points) -- the zone header
   !> 1 (rectangle), NF, NF, then for j = 1..NF, i = 1..NF the real32 triple
   !> f(i), f(j), BRM(f(i), f(j)) of the modal entry (its diagonal-model entry
   !> with the ONLY_DIAG model).  The values come from the device in FP64.
   subroutine exportVisualCls_(p, nfh, df, nme)
      type(bsahip_problem_t), intent(in) :: p
      integer(bsa_int_t), intent(in) :: nfh, nme
      real(bsa_real_t), intent(in) :: df
      integer(c_int32_t), target :: sel(1)
      real(bsa_real_t), allocatable, target :: brm(:, :)
      real(real32), allocatable :: rec(:, :)
      integer(c_int32_t) :: fl
      integer(c_int) :: rc
      integer :: iun, ios, i, j, nf
      nf = 2 * nfh - 1
      fl = 0
      if (ionly_diag_ /= 0) then
         fl = BSAHIP_ONLY_DIAG
         if (vis_idx_(1) < 1 .or. vis_idx_(1) > nme) call bsa_Abort('Visual mode index out of the kept modes.')
         sel(1) = int(vis_idx_(1), c_int32_t)
      else
         if (any(vis_idx_ < 1) .or. any(vis_idx_ > nme)) call bsa_Abort('Visual mode indexes out of the kept modes.')
         sel(1) = int((vis_idx_(3) - 1) * nme * nme + (vis_idx_(2) - 1) * nme + vis_idx_(1), c_int32_t)
      end if
      allocate(brm(nf, nf))
      rc = bsahip_classic_bispectra_sel(p, int(nfh, c_int32_t), real(df, c_double), fl, 1_c_int32_t, &
                                        c_loc(sel), c_null_ptr, c_loc(brm))
      if (rc /= 0) call bsa_Abort('bsahip_classic_bispectra_sel failed: '//bsahip_last_error_f())
      if (associated(brm_cb_)) then
         call callbackZone_(nf, nf, [(real(i - nfh, bsa_real_t) * df, i = 1, nf)], &
                            [(real(i - nfh, bsa_real_t) * df, i = 1, nf)], reshape(brm, [1, nf * nf]))
         return
      end if
      write(output_unit, '( /, 1x, 2a, i0, " x ", i0, a )') INFOMSG, 'Exporting  BRM  to file...  (', nf, nf, ')'
      open(newunit=iun, file=BRM_EXPORT_FNAME_CLS, form='unformatted', access='stream', &
           status='replace', action='readwrite', iostat=ios)
      if (ios /= 0) call bsa_Abort('Error opening file  '//BRM_EXPORT_FNAME_CLS)
      write(iun) 3_int32, int(vis_idx_, int32), 1_int32, int(isym_, int32), 1_int32, int(nf * nf, int32)
      write(iun) 1_int32, int(nf, int32), int(nf, int32)
      allocate(rec(3, nf))
      do i = 1, nf
         rec(1, i) = real(real(i - nfh, bsa_real_t) * df, real32)
      end do
      do j = 1, nf
         rec(2, :) = real(real(j - nfh, bsa_real_t) * df, real32)
         rec(3, :) = real(brm(:, j), real32)
         write(iun) rec
      end do
      close(iun)
   end subroutine

   function openExport_(fname) result(iun)
      character(len=*), intent(in) :: fname
      integer :: iun, ios
      character(len=:), allocatable :: path, pos, st
      path = fname
      if (allocated(exp_dir_)) path = exp_dir_//fname
      pos = merge('append', 'asis  ', exp_append_)
      st = merge('unknown', 'replace', exp_append_)
      if (exp_binary_) then
         open(newunit=iun, file=path, form='unformatted', access='stream', action='write', &
              status=trim(st), position=trim(pos), iostat=ios)
      else
         open(newunit=iun, file=path, form='formatted', action='write', status=trim(st), &
              position=trim(pos), iostat=ios)
      end if
      if (ios /= 0) call bsa_Abort('Cannot open export file  '//path)
   end function

   !> src/bsa/BsaLibImpl.F90:1590-1604: size, then one value per line.
   subroutine bsa_exportMomentToFile(fname, vec)
      character(len=*), intent(in) :: fname
      real(bsa_real_t), intent(in) :: vec(:)
      call writeVector_(fname, vec, .false.)
   end subroutine

   !> BsaLibImpl.F90:1745-1760.
   subroutine bsa_exportExtremeValuesToFile(fname, rvar)
      character(len=*), intent(in) :: fname
      real(bsa_real_t), intent(in) :: rvar(:)
      call writeVector_(fname, rvar, .false.)
   end subroutine

   !> BsaLibImpl.F90:1608-1613, 1715-1736: with the modal header (number of
   !> used modes and their indices) when the array holds NM^3 values.
   subroutine exportSkewnessGiven_(fname, sk)
      character(len=*), intent(in) :: fname
      real(bsa_real_t), intent(in) :: sk(:)
      logical :: modal
      modal = .false.
      if (allocated(modes_)) modal = size(sk) == size(modes_)**3
      call writeVector_(fname, sk, modal)
   end subroutine

   !> BsaLibImpl.F90:1617-1706: sk(i,j,k) = m3(i,j,k) / (sigma_i sigma_j sigma_k),
   !> sigma_i = sqrt(m2(i,i)) (the reference's indexing of the m2 diagonal).
   subroutine exportSkewnessCompute_(fname, dim, m2, m3)
      character(len=*), intent(in) :: fname
      integer(bsa_int_t), value :: dim
      real(bsa_real_t), intent(in) :: m2(:), m3(:)
      real(bsa_real_t), allocatable :: sk(:)
      integer :: i, j, k, pm3
      if (ionly_diag_ /= 0) then
         sk = m3 / m2**CST_3d2
      else
         allocate(sk(size(m3)))
         pm3 = 1
         do k = 1, dim
            do j = 1, dim
               do i = 1, dim
                  sk(pm3) = m3(pm3) / (sqrt(m2((k - 1) * dim + k)) * sqrt(m2((j - 1) * dim + j)) &
                                       * sqrt(m2((i - 1) * dim + i)))
                  pm3 = pm3 + 1
               end do
            end do
         end do
      end if
      call exportSkewnessGiven_(fname, sk)
   end subroutine

   !> BsaLibImpl.F90:1564-1585: used modes, damping ratios, then bkg and res per mode.
   subroutine exportBR_(fname, bkg, res, xsi)
      character(len=*), intent(in) :: fname
      real(bsa_real_t), intent(in) :: bkg(:), res(:), xsi(:)
      integer :: iun, im
      iun = openExport_(fname)
      if (exp_binary_) then
         write(iun) int(size(modes_), int32), int(modes_, int32), xsi, bkg, res
      else
         write(iun, *) size(modes_)
         write(iun, *) modes_
         write(iun, *) xsi
         do im = 1, size(modes_)
            write(iun, *) bkg(im), res(im)
         end do
      end if
      close(iun)
   end subroutine

   !> BsaLibImpl.F90:1937-1967: number of nodes, then (x, y, z) per node.
   subroutine bsa_saveCoordinatesToFile(fname, coords)
      character(len=*), intent(in) :: fname
      real(bsa_real_t), intent(in), target, optional :: coords(:, :)
      integer :: iun, ios, i
      if (.not. present(coords) .and. .not. allocated(coords_)) then
         write(output_unit, '(1x, 2a)') WARNMSG, 'Cannot save coordinates to file. Data not provided. Skipping.'
         return
      end if
      open(newunit=iun, file=fname, form='formatted', action='write', status='replace', iostat=ios)
      if (ios /= 0) return
      if (present(coords)) then
         write(iun, *) size(coords, 2)
         do i = 1, size(coords, 2)
            write(iun, *) coords(:, i)
         end do
      else
         write(iun, *) size(coords_, 2)
         do i = 1, size(coords_, 2)
            write(iun, *) coords_(:, i)
         end do
      end if
      close(iun)
   end subroutine

   subroutine writeVector_(fname, vec, modal)
      character(len=*), intent(in) :: fname
      real(bsa_real_t), intent(in) :: vec(:)
      logical, intent(in) :: modal
      integer :: iun, i
      iun = openExport_(fname)
      if (exp_binary_) then
         if (modal) write(iun) int(size(modes_), int32), int(modes_, int32)
         write(iun) int(size(vec), int32), vec
      else
         if (modal) then
            write(iun, *) size(modes_)
            write(iun, *) modes_
         end if
         write(iun, *) size(vec)
         do i = 1, size(vec)
            write(iun, *) vec(i)
         end do
      end if
      close(iun)
   end subroutine

end module BsaLib
