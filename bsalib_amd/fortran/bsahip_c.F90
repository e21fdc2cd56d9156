!! Thin iso_c_binding layer over libbsahip.so (C declarations: include/bsahip.h).
!!
!! bsahip_problem_t is interoperable with the C struct bsahip_problem; the
!! interfaces bind the C entry points by name.  Nothing here computes.
module BsaHip_C
   use, intrinsic :: iso_c_binding
   implicit none
   private

   integer(c_int32_t), parameter, public :: BSAHIP_PREC_FP64 = 0
   integer(c_int32_t), parameter, public :: BSAHIP_PREC_FP32 = 1
   integer(c_int32_t), parameter, public :: BSAHIP_PREC_BF16X3 = 2
   integer(c_int32_t), parameter, public :: BSAHIP_QUAD_VECT = 0
   integer(c_int32_t), parameter, public :: BSAHIP_QUAD_RIEMANN = 16
   integer(c_int32_t), parameter, public :: BSAHIP_QUAD_TRAPZ = 32
   integer(c_int32_t), parameter, public :: BSAHIP_SUM_EXACT = 256
   integer(c_int32_t), parameter, public :: BSAHIP_NO_PSD = 512
   integer(c_int32_t), parameter, public :: BSAHIP_NO_BISP = 1024
   integer(c_int32_t), parameter, public :: BSAHIP_ONLY_DIAG = 2048
   integer(c_int32_t), parameter, public :: BSAHIP_FRF_PSD = 0
   integer(c_int32_t), parameter, public :: BSAHIP_FRF_GRID = 1
   integer(c_int32_t), parameter, public :: BSAHIP_FRF_POINTS = 2

   type, bind(c), public :: bsahip_problem_t
      integer(c_int32_t) :: nn = 0, nnl = 0, nm = 0, ntc = 0, ndegw = 0, nwz = 0
      integer(c_int32_t) :: psd_id = 0, pulsation = 1, ncorr_rows = 0, corr_col_mode = 0
      integer(c_int32_t) :: pod_nmodes = 0, reserved = 0
      real(c_double) :: pod_trunc_lim = 0
      type(c_ptr) :: n_load = c_null_ptr, tc = c_null_ptr, psi = c_null_ptr
      type(c_ptr) :: nod_corr = c_null_ptr, u_node = c_null_ptr, wz_node = c_null_ptr
      type(c_ptr) :: turb_scales = c_null_ptr, turb_std = c_null_ptr, u_mean_ref = c_null_ptr
      type(c_ptr) :: Mm = c_null_ptr, Km = c_null_ptr, Cm = c_null_ptr
   end type bsahip_problem_t

   type, bind(c), public :: bsahip_mesher_cfg_t
      integer(c_int32_t) :: bkg_base_rfmnt = 20, tile_cells = 32, diag_crest = 0, brm_cells = 0
      real(c_double) :: bkg_peakw = 0, bkg_area_ext = 2, peak_area_ext = 2, max_area_ext = 2
      type(c_ptr) :: natf = c_null_ptr, xsi = c_null_ptr
   end type bsahip_mesher_cfg_t

   public :: bsahip_init, bsahip_finalise, bsahip_classic_run, bsahip_bfm_points
   public :: bsahip_classic_bispectra_sel
   public :: bsahip_mesher_create, bsahip_mesher_run, bsahip_mesher_destroy, bsahip_mesher_export
   public :: bsahip_mesher_adapt, bsahip_mesher_adapt_run
   public :: bsahip_mesher_ntiles
   public :: bsahip_last_error_f, bsahip_modal_psd_points, bsahip_modal_recombination
   public :: bsahip_classic_spectra, bsahip_frf_apply, bsahip_bfm_points_cls, bsahip_diag_psd_points

   interface
      integer(c_int) function bsahip_init(device) bind(c, name="bsahip_init")
         import c_int
         integer(c_int), value :: device
      end function

      integer(c_int) function bsahip_finalise() bind(c, name="bsahip_finalise")
         import c_int
      end function

      integer(c_int) function bsahip_classic_run(p, nfh, df, flags, m2mf, m2mr, m2o2mr, m3mf, m3mr) &
            bind(c, name="bsahip_classic_run")
         import c_int, c_int32_t, c_double, c_ptr, bsahip_problem_t
         type(bsahip_problem_t), intent(in) :: p
         integer(c_int32_t), value :: nfh, flags
         real(c_double), value :: df
         type(c_ptr), value :: m2mf, m2mr, m2o2mr, m3mf, m3mr
      end function

      integer(c_int) function bsahip_classic_bispectra_sel(p, nfh, df, flags, nsel, sel, bfm, brm) &
            bind(c, name="bsahip_classic_bispectra_sel")
         import c_int, c_int32_t, c_double, c_ptr, bsahip_problem_t
         type(bsahip_problem_t), intent(in) :: p
         integer(c_int32_t), value :: nfh, flags, nsel
         real(c_double), value :: df
         type(c_ptr), value :: sel, bfm, brm
      end function

      integer(c_int) function bsahip_bfm_points(p, npts, fi, fj, flags, bfm, brm) &
            bind(c, name="bsahip_bfm_points")
         import c_int, c_int32_t, c_ptr, bsahip_problem_t
         type(bsahip_problem_t), intent(in) :: p
         integer(c_int32_t), value :: npts, flags
         type(c_ptr), value :: fi, fj, bfm, brm
      end function

      integer(c_int) function bsahip_bfm_points_cls(p, nfh, df, flags, npts, fi, fj, bfm, brm) &
            bind(c, name="bsahip_bfm_points_cls")
         import c_int, c_int32_t, c_double, c_ptr, bsahip_problem_t
         type(bsahip_problem_t), intent(in) :: p
         integer(c_int32_t), value :: nfh, flags, npts
         real(c_double), value :: df
         type(c_ptr), value :: fi, fj, bfm, brm
      end function

      integer(c_int) function bsahip_diag_psd_points(p, npts, f, psd) &
            bind(c, name="bsahip_diag_psd_points")
         import c_int, c_int32_t, c_ptr, bsahip_problem_t
         type(bsahip_problem_t), intent(in) :: p
         integer(c_int32_t), value :: npts
         type(c_ptr), value :: f, psd
      end function

      integer(c_int) function bsahip_classic_spectra(p, nfh, df, flags, psd_f, psd_r, bfm, brm) &
            bind(c, name="bsahip_classic_spectra")
         import c_int, c_int32_t, c_double, c_ptr, bsahip_problem_t
         type(bsahip_problem_t), intent(in) :: p
         integer(c_int32_t), value :: nfh, flags
         real(c_double), value :: df
         type(c_ptr), value :: psd_f, psd_r, bfm, brm
      end function

      integer(c_int) function bsahip_frf_apply(p, flags, kind, n1, f1, n2, f2, data) &
            bind(c, name="bsahip_frf_apply")
         import c_int, c_int32_t, c_int64_t, c_ptr, bsahip_problem_t
         type(bsahip_problem_t), intent(in) :: p
         integer(c_int32_t), value :: flags, kind
         integer(c_int64_t), value :: n1, n2
         type(c_ptr), value :: f1, f2, data
      end function

      integer(c_int) function bsahip_mesher_create(p, cfg, flags, mesher) &
            bind(c, name="bsahip_mesher_create")
         import c_int, c_int32_t, c_ptr, bsahip_problem_t, bsahip_mesher_cfg_t
         type(bsahip_problem_t), intent(in) :: p
         type(bsahip_mesher_cfg_t), intent(in) :: cfg
         integer(c_int32_t), value :: flags
         type(c_ptr), intent(out) :: mesher
      end function

      integer(c_int) function bsahip_mesher_run(mesher, levels, m3mf, m3mr, counts) &
            bind(c, name="bsahip_mesher_run")
         import c_int, c_ptr
         type(c_ptr), value :: mesher, levels, m3mf, m3mr, counts
      end function

      integer(c_int) function bsahip_mesher_adapt(mesher, tol, levels, evals) &
            bind(c, name="bsahip_mesher_adapt")
         import c_int, c_double, c_ptr
         type(c_ptr), value :: mesher, levels, evals
         real(c_double), value :: tol
      end function

      integer(c_int) function bsahip_mesher_adapt_run(mesher, tol, levels, evals, m3mf, m3mr, &
                                                      counts) bind(c, name="bsahip_mesher_adapt_run")
         import c_int, c_double, c_ptr
         type(c_ptr), value :: mesher, levels, evals, m3mf, m3mr, counts
         real(c_double), value :: tol
      end function

      integer(c_int) function bsahip_modal_psd_points(p, npts, f, psd) &
            bind(c, name="bsahip_modal_psd_points")
         import c_int, c_int32_t, c_ptr, bsahip_problem_t
         type(bsahip_problem_t), intent(in) :: p
         integer(c_int32_t), value :: npts
         type(c_ptr), value :: f, psd
      end function

      integer(c_int) function bsahip_modal_recombination(ndofs, nm, phi, m2, m3, m2o2, ccc, &
            m2_r, m3_r, m2o2_r, sk_r) bind(c, name="bsahip_modal_recombination")
         import c_int, c_int32_t, c_ptr
         integer(c_int32_t), value :: ndofs, nm, ccc
         type(c_ptr), value :: phi, m2, m3, m2o2, m2_r, m3_r, m2o2_r, sk_r
      end function

      integer(c_int) function bsahip_mesher_export(mesher, levels, ncol, coef, cap, dims, fi, fj, &
                                                   val, npts) bind(c, name="bsahip_mesher_export")
         import c_int, c_int32_t, c_int64_t, c_ptr
         type(c_ptr), value :: mesher, levels, coef, dims, fi, fj, val
         integer(c_int32_t), value :: ncol
         integer(c_int64_t), value :: cap
         integer(c_int64_t), intent(out) :: npts
      end function

      integer(c_int) function bsahip_mesher_ntiles(mesher, ntiles) bind(c, name="bsahip_mesher_ntiles")
         import c_int, c_int32_t, c_ptr
         type(c_ptr), value :: mesher
         integer(c_int32_t), intent(out) :: ntiles
      end function

      integer(c_int) function bsahip_mesher_destroy(mesher) bind(c, name="bsahip_mesher_destroy")
         import c_int, c_ptr
         type(c_ptr), value :: mesher
      end function

      type(c_ptr) function bsahip_last_error_c() bind(c, name="bsahip_last_error")
         import c_ptr
      end function
   end interface

contains

   !> Message of the last failed libbsahip call on this thread.
   function bsahip_last_error_f() result(msg)
      character(len=:), allocatable :: msg
      character(kind=c_char), pointer :: s(:)
      type(c_ptr) :: p
      integer :: n
      p = bsahip_last_error_c()
      msg = ''
      if (.not. c_associated(p)) return
      call c_f_pointer(p, s, [4096])
      n = 0
      do while (n < 4096)
         if (s(n + 1) == c_null_char) exit
         n = n + 1
      end do
      allocate(character(len=n) :: msg)
      do n = 1, len(msg)
         msg(n:n) = s(n)
      end do
   end function

end module BsaHip_C
