!! Test driver of the Fortran BsaLib front (module BsaLib in bsalib_amd/fortran):
!! a program written the way a user of the reference writes one (setters,
!! bsa_Run, bsa_Finalise), reading its inputs from a raw stream file.
!!
!! Usage:  bsalib_driver <input.bin> <output.bin> [fp64|bf16x3] [grid] [scalar] [mesher|both]
!!                        [pod] [podn=<modes>] [podt=<percent>] [post] [gen] [diag]
!!                        [visual=<i>,<j>,<k>] [visualn=<node>,<dof>] [sym=<n>] [brmbase]
!!                        [mtol=<tolerance>]
!!                        [cb] [premesh] [modal] [slots]
!! The input layout is the one tests write for the reference harness (header
!! of 14 int32, then integer and real64 arrays in Fortran order).  "grid"
!! leaves test mode off and passes damping ratios, so bsa_Run applies the
!! frequency-grid checks; "scalar" selects the scalar Classic version;
!! "mesher" runs analysis type 2 (Mesher only), "both" type 3; "pod" sets
!! isvd = 1 (POD BFM) in bsa_setupMesher, "podn=" / "podt=" call
!! bsa_setPODNOfModesKept / bsa_setPODTruncationThreshold; "post" runs the
!! post-processing after bsa_Run (bsa_computeBRdecomp, bsa_modalRecombination
!! on the used modes at every DOF, bsa_computePeakFactors over 600 s) and
!! appends bkg, res, the unilateral and complete m2, m3, m2o2, sk and the
!! Gaussian / positive / negative peak factors of both.  "gen" calls
!! bsa_generateBSAInputFiles(.false.): bsa_Run writes bsa.bsadata and
!! bsa.extdata in the current directory and returns (no device needed);
!! "diag" selects the ONLY_DIAG model (bsa_setSpectraExtension(1)); "visual="
!! enables visual mode with modal indexes (i, j, k), "visualn=" with a node and
!! DOF (bsa_enableVisualMode, bsa_setVisualIndexes); "sym=" calls
!! bsa_setSpatialSymmetry; "brmbase" calls
!! bsa_setBRMExportDefaultMode(BSA_EXPORT_BRM_MODE_BASE); "cb" registers a
!! bispectrum export callback (bsa_setBispExportCallback) that prints, at the
!! end, "CALLBACK <zones> <points> <sum of values>"; "premesh" calls
!! bsa_enableOnlyPremesh, "modal" bsa_exportModalData; "mtol=" sets the
!! Mesher's target error (bsa_setMesherTolerance; 0: rule-based levels).
!! "slots" (with "both", and "diag" for the ONLY_DIAG operators) calls the
!! operator slots bsa_Run bound (getBFM_vect_cls ... getBRM_msh) on the Classic
!! grid, prints
!! "SLOT <check> <max relative difference>" for the scalar and Mesher
!! operators against the vectorised ones at a strided set of grid points,
!! and appends the vectorised psd_f, psd_r, bisp_f, bisp_r.
!! Output: nme (int32), the used modes, then for each of m2mf, m2mr, m2o2mr,
!! m3mf_cls, m3mr_cls, m3mf_msh, m3mr_msh: its size (int32, 0 when bsa_Run
!! left it unallocated) and its values (real64).
! the export callback lives in a module: a procedure pointer to an internal
! procedure would need an executable stack
module driver_cb
   use, intrinsic :: iso_fortran_env, only: int64, real64
   implicit none
   integer :: cb_zones = 0
   integer(int64) :: cb_pts = 0
   real(real64) :: cb_sum = 0
contains
   subroutine cb_(f1, f2, brm, pdata)
      real(real64), intent(in) :: f1(:), f2(:), brm(:, :)
      class(*), pointer, intent(in) :: pdata
      if (size(f1) /= size(f2) .or. size(brm, 2) /= size(f1) .or. associated(pdata)) &
         error stop 'callback arguments'
      cb_zones = cb_zones + 1
      cb_pts = cb_pts + size(f1)
      cb_sum = cb_sum + sum(brm)
   end subroutine
end module driver_cb

program bsalib_driver
   use, intrinsic :: iso_fortran_env, only: int32, int64, real64
   use BsaLib
   use driver_cb
   implicit none

   integer(int32) :: hdr(14)
   integer(int32) :: nn, nlibs, nnl, nlibsl, nm, nme, ntc, ndegw, nwz, psdid, nfh, ncorr, npts, i
   integer(int32) :: vis(3), isym
   integer(int32), allocatable :: n_load(:), libs_load(:), modes(:), tcs(:), wz(:), used(:)
   real(real64) :: df
   real(real64), allocatable, target :: phi(:, :), natf(:), Mm(:), Km(:), Cm(:, :), unod(:)
   real(real64), allocatable, target :: tscl(:, :, :), tstd(:, :), umref(:), corr(:, :)
   real(real64), allocatable, target :: wfc(:, :, :), psi(:, :, :), xsi(:)
   real(real64), allocatable :: m2mf(:), m2mr(:), m2o2mr(:), m3mf_msh(:), m3mr_msh(:)
   real(real64), allocatable :: m3mf(:), m3mr(:)
   real(real64), allocatable :: bkg(:), res(:), m2d(:), m3d(:), m22d(:), skd(:), m2c(:), m3c(:)
   real(real64), allocatable :: m22c(:), skc(:), pgd(:), ppd(:), pnd(:), pgc(:), ppc(:), pnc(:)
   logical :: post, slots, diagm
   real(real64), allocatable :: sl_psdf(:, :), sl_psdr(:, :), sl_bf(:, :, :), sl_br(:, :, :)
   character(len=512) :: fin, fout, arg
   logical :: grid_checks
   integer :: uin, uout, isuban, isvd, npod
   real(real64) :: podt
   procedure(exportInterf_vect_), pointer :: pcb

   call get_command_argument(1, fin)
   call get_command_argument(2, fout)
   grid_checks = .false.
   post = .false.
   slots = .false.
   diagm = .false.
   isuban = 1
   isvd = 0

   open(newunit=uin, file=trim(fin), access='stream', form='unformatted', status='old')
   read(uin) hdr
   nn = hdr(1); nlibs = hdr(2); nnl = hdr(3); nlibsl = hdr(4); nm = hdr(5); nme = hdr(6)
   ntc = hdr(7); ndegw = hdr(8); nwz = hdr(9); psdid = hdr(10); nfh = hdr(11)
   npts = hdr(13); ncorr = hdr(14)
   allocate(n_load(nnl), libs_load(nlibsl), modes(nme), tcs(ntc), wz(nn))
   read(uin) n_load, libs_load, modes, tcs, wz
   allocate(phi(nn*nlibs, nm), natf(nm), Mm(nm), Km(nm), Cm(nm, nm), unod(nn))
   allocate(tscl(3, 3, nwz), tstd(3, nwz), umref(nwz), corr(ncorr, 3))
   allocate(wfc(nlibsl, ndegw, nnl), psi(nme, nnl, ndegw))
   read(uin) df
   read(uin) phi, natf, Mm, Km, Cm, unod, tscl, tstd, umref, corr, wfc, psi
   close(uin)

   call bsa_Init()
   call bsa_setClassicMode(BSA_CLASSIC_MODE_VECTOR)
   call bsa_setScalingConv(BSA_PSD_CONVENTION_PULS)
   call bsa_setSpectraComputation(1, 1)
   call bsa_setupClassic(nfh, df)
   call bsa_setTestMode(1)
   do i = 3, command_argument_count()
      call get_command_argument(i, arg)
      if (trim(arg) == 'bf16x3') call bsa_setComputePrecision(BSA_PREC_BF16X3)
      if (trim(arg) == 'grid') grid_checks = .true.
      if (trim(arg) == 'scalar') call bsa_setClassicMode(BSA_CLASSIC_MODE_SCALAR)
      if (trim(arg) == 'mesher') isuban = 2
      if (trim(arg) == 'both') isuban = 3
      if (trim(arg) == 'pod') isvd = 1
      if (trim(arg) == 'post') post = .true.
      if (trim(arg) == 'slots') slots = .true.
      if (trim(arg) == 'gen') call bsa_generateBSAInputFiles(.false.)
      if (trim(arg) == 'diag') then
         call bsa_setSpectraExtension(1)
         diagm = .true.
      end if
      if (arg(1:7) == 'visual=') then
         read(arg(8:), *) vis
         call bsa_enableVisualMode()
         call bsa_setVisualIndexes(vis, .true.)
      end if
      if (arg(1:8) == 'visualn=') then
         vis = 0
         read(arg(9:), *) vis(1:2)
         call bsa_enableVisualMode()
         call bsa_setVisualIndexes(vis, .false.)
      end if
      if (trim(arg) == 'brmbase') call bsa_setBRMExportDefaultMode(BSA_EXPORT_BRM_MODE_BASE)
      if (trim(arg) == 'cb') then
         pcb => cb_
         call bsa_setBispExportCallback(pcb)
      end if
      if (trim(arg) == 'premesh') call bsa_enableOnlyPremesh()
      if (trim(arg) == 'modal') call bsa_exportModalData()
      if (arg(1:4) == 'sym=') then
         read(arg(5:), *) isym
         call bsa_setSpatialSymmetry(isym)
      end if
      if (arg(1:5) == 'podn=') then
         read(arg(6:), *) npod
         call bsa_setPODNOfModesKept(npod)
      end if
      if (arg(1:5) == 'podt=') then
         read(arg(6:), *) podt
         call bsa_setPODTruncationThreshold(podt)
      end if
      if (arg(1:5) == 'mtol=') then
         read(arg(6:), *) podt
         call bsa_setMesherTolerance(podt)
      end if
   end do
   call bsa_setAnalysisType(isuban)
   if (isuban /= 1) call bsa_setupMesher(isvd, 20, 2.0_real64, 2.0_real64, 2.0_real64, 1, 0)

   call bsa_setTotalNOfNodes(nn)
   call bsa_setNumOfNodalDOFs(nlibs)
   call bsa_setLoadedNodes(n_load)
   call bsa_setLoadedNodalDOFs(libs_load)
   call bsa_setModalInfo(nn*nlibs, nm, phi, natf)
   call bsa_setModalMatrices(nm, Mm, Km, Cm)
   ! modal damping ratios of the diagonal model: xsi = c / (2 sqrt(k m))
   allocate(xsi(nm))
   do i = 1, nm
      xsi(i) = Cm(i, i) / (2.0_real64 * sqrt(Km(i) * Mm(i)))
   end do
   call bsa_setTotDamping(xsi)
   if (grid_checks) call bsa_setTestMode(0)

   call bsa_setPSDType(psdid)
   call bsa_setWindDirections([1])
   call bsa_setWindTurbComps(tcs)
   call bsa_setWZMeanWindVel(umref)
   call bsa_setTurbWindScales(tscl)
   call bsa_setTurbWindSDT(tstd)
   call bsa_setNodalVel(unod)
   call bsa_setNodalWindZones(wz)
   call bsa_setSpatialNodalCorr(corr)
   call bsa_setWindFCoeffs(wfc)

   call bsa_Run(m2mf, m2mr, m2o2mr, m3mf_msh, m3mr_msh, m3mf, m3mr)
   used = bsa_getUsedModeShapes()
   if (slots) call check_slots()
   if (post) then
      call bsa_computeBRdecomp(m2mf, bkg, res)
      call bsa_modalRecombination(phi(:, used), m2mr, m3mr, m2o2mr, m2d, m3d, m22d, skd, &
                                  m2c, m3c, m22c, skc)
      call bsa_computePeakFactors(m2d, m22d, 600.0_real64, pgd, skd, ppd, pnd)
      call bsa_computePeakFactors(m2c, m22c, 600.0_real64, pgc, skc, ppc, pnc)
   end if

   open(newunit=uout, file=trim(fout), access='stream', form='unformatted', status='replace')
   write(uout) int(size(used), int32), used
   call put(m2mf)
   call put(m2mr)
   call put(m2o2mr)
   call put(m3mf)
   call put(m3mr)
   call put(m3mf_msh)
   call put(m3mr_msh)
   if (post) then
      call put(bkg)
      call put(res)
      call put(m2d)
      call put(m3d)
      call put(m22d)
      call put(skd)
      call put(m2c)
      call put(m3c)
      call put(m22c)
      call put(skc)
      call put(pgd)
      call put(ppd)
      call put(pnd)
      call put(pgc)
      call put(ppc)
      call put(pnc)
   end if
   if (slots) then
      call put_vals(reshape(sl_psdf, [size(sl_psdf)]))
      call put_vals(reshape(sl_psdr, [size(sl_psdr)]))
      call put_vals(reshape(sl_bf, [size(sl_bf)]))
      call put_vals(reshape(sl_br, [size(sl_br)]))
   end if
   close(uout)
   if (cb_zones > 0) print '(a, 1x, i0, 1x, i0, 1x, es25.17)', 'CALLBACK', cb_zones, cb_pts, cb_sum
   call bsa_Finalise()
   if (.not. bsa_isCleaned()) error stop 'bsa_Finalise did not clean up'
contains
   subroutine put(a)
      real(real64), allocatable, intent(in) :: a(:)
      if (allocated(a)) then
         write(uout) int(size(a), int32), a
      else
         write(uout) 0_int32
      end if
   end subroutine

   subroutine put_vals(a)
      real(real64), intent(in) :: a(:)
      write(uout) int(size(a), int32), a
   end subroutine

   subroutine check_slots()
      real(real64), allocatable :: f(:), suvw(:, :), spad(:), psd(:, :), bisp(:, :, :)
      real(real64), allocatable :: p1(:), b1(:), pr1(:), br1(:), fi(:), fj(:), bm(:, :), rm(:, :)
      real(real64) :: e(6), sf, sr, pf, pr
      integer(int32) :: nf, n3, n2, ii, ij, st, q, npt
      if (.not. (associated(getBFM_vect_cls) .and. associated(getBRM_vect_cls) .and. &
                 associated(getBFM_scalar_cls) .and. associated(getBRM_scalar_cls) .and. &
                 associated(getBFM_msh) .and. associated(getBRM_msh))) error stop 'slots not bound'
      nf = 2 * nfh - 1
      n2 = size(used)**2
      n3 = size(used)**3
      if (diagm) then
         n2 = size(used)
         n3 = size(used)
      end if
      allocate(f(nf), suvw(nf, 1), spad(1), p1(n2), b1(n3), pr1(n2), br1(n3))
      suvw = 0
      spad = 0
      do i = 1, nf
         f(i) = (i - nfh) * df
      end do
      call getBFM_vect_cls(f, suvw, psd, bisp)
      sl_psdf = psd
      sl_bf = bisp
      call getBRM_vect_cls(f, psd, bisp)
      sl_psdr = psd
      sl_br = bisp
      sf = maxval(abs(sl_bf)); sr = maxval(abs(sl_br))
      pf = maxval(abs(sl_psdf)); pr = maxval(abs(sl_psdr))
      e = 0
      st = max(1, nf / 9)
      npt = 0
      allocate(fi(nf * nf), fj(nf * nf))
      do ii = 1, nf, st
         do ij = 1, nf, st
            call getBFM_scalar_cls(ii, ij, f(ii), f(ij), suvw, spad, p1, b1)
            call getBRM_scalar_cls(ii, ij, f(ii), f(ij), p1, pr1, b1, br1)
            e(1) = max(e(1), maxval(abs(b1 - sl_bf(ii, ij, :))) / sf)
            e(2) = max(e(2), maxval(abs(br1 - sl_br(ii, ij, :))) / sr)
            if (ij == 1) then
               e(3) = max(e(3), maxval(abs(p1 - sl_psdf(ii, :))) / pf)
               e(4) = max(e(4), maxval(abs(pr1 - sl_psdr(ii, :))) / pr)
            else if (any(p1 /= 0) .or. any(pr1 /= 0)) then
               error stop 'scalar psd not zero off the first column'
            end if
            ! Mesher points: inside the grid S(fi + fj) is the same on both paths
            if (abs(f(ii) + f(ij)) <= f(nf)) then
               npt = npt + 1
               fi(npt) = f(ii)
               fj(npt) = f(ij)
            end if
         end do
      end do
      allocate(bm(n3, npt), rm(n3, npt))
      call getBFM_msh(bm, fi(1:npt), fj(1:npt))
      call getBRM_msh(rm, fi(1:npt), fj(1:npt), bm)
      do q = 1, npt
         ii = nint(fi(q) / df) + nfh
         ij = nint(fj(q) / df) + nfh
         e(5) = max(e(5), maxval(abs(bm(:, q) - sl_bf(ii, ij, :))) / sf)
         e(6) = max(e(6), maxval(abs(rm(:, q) - sl_br(ii, ij, :))) / sr)
      end do
      print '(a, 1x, a, 1x, es12.5)', 'SLOT', 'bfm_scalar', e(1)
      print '(a, 1x, a, 1x, es12.5)', 'SLOT', 'brm_scalar', e(2)
      print '(a, 1x, a, 1x, es12.5)', 'SLOT', 'psdf_scalar', e(3)
      print '(a, 1x, a, 1x, es12.5)', 'SLOT', 'psdr_scalar', e(4)
      print '(a, 1x, a, 1x, es12.5)', 'SLOT', 'bfm_msh', e(5)
      print '(a, 1x, a, 1x, es12.5)', 'SLOT', 'brm_msh', e(6)
      print '(a, 1x, a, 1x, i0)', 'SLOT', 'msh_points', npt
   end subroutine
end program bsalib_driver
