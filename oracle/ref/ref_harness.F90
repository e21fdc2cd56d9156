!! Test-only driver of the REFERENCE's own BFM/BRM operators (not product code).
!!
!! Built by oracle/ref/build_ref.sh from the reference's sources where they lie
!! (src/CONSTANTS, src/IO, src/settings, src/structure, src/wind, src/bsa/data,
!! src/utils, src/bsa/functions); the output goes to oracle/_ref/.  It fills the
!! reference's global singletons (settings, wd, struct_data in BsaLib_Data)
!! through the reference's own type-bound setters, then calls:
!!   wd%evalPSD                    (src/wind/WindPSDImpl.F90:123-144)
!!   getFM_full_tnm_vect_cls_      (src/bsa/functions/functionsImpl.F90:1312-1587)
!!   getRM_full_vect_cls_          (src/bsa/functions/functionsImpl.F90:1598-1738)
!!   getFM_full_tnm_scalar_msh_    (src/bsa/functions/functionsImpl.F90:247-465)
!!   getFM_diag_tnlm_vect_cls_     (src/bsa/functions/functionsImpl.F90:1747-1970)
!!   getRM_diag_vect_cls_          (src/bsa/functions/functionsImpl.F90:1975-2101)
!! and writes their outputs to a raw stream file for tests/golden.
!!
!! Usage:  ref_harness <input.bin> <output.bin> [time]
!! With "time" it only times getFM_full_tnm_vect_cls_ (wall clock) and prints
!! "REFTIME <seconds> <threads>".
program ref_harness
   use, intrinsic :: iso_fortran_env, only: int32, real64
   use BsaLib_CONSTANTS
   use BsaLib_Data
   use BsaLib_Functions, only: setBsaFunctionLocalVars, getFM_full_tnm_vect_cls_, getRM_full_vect_cls_, getFM_full_tnm_scalar_msh_, &
                               getFM_diag_tnlm_vect_cls_, getRM_diag_vect_cls_
   use omp_lib
   implicit none

   integer(int32) :: hdr(14)
   integer(int32) :: nn, nlibs, nnl, nlibsl, nm, nme, ntc, ndegw, nwz, psdid, nfh, imode, npts, ncorr
   integer(int32), allocatable, target :: n_load(:), libs_load(:), modes(:), tcs(:), wz(:)
   real(real64) :: df
   real(real64), allocatable, target :: phi(:, :), natf(:), Mm(:), Km(:), Cm(:, :), unod(:)
   real(real64), allocatable, target :: tscl(:, :, :), tstd(:, :), umref(:), corr(:, :)
   real(real64), allocatable, target :: wfc(:, :, :), psi(:, :, :), fi(:), fj(:)
   real(real64), allocatable :: f(:), suvw(:, :), psd(:, :), bisp(:, :, :), bfm(:, :)
   integer(int32) :: uin, uout, i, itc, idx, nf, ip
   character(len=512) :: fin, fout, opt
   real(real64) :: t0, t1

   call get_command_argument(1, fin)
   call get_command_argument(2, fout)
   opt = ''
   if (command_argument_count() >= 3) call get_command_argument(3, opt)

   open(newunit=uin, file=trim(fin), access='stream', form='unformatted', status='old')
   read(uin) hdr
   nn = hdr(1); nlibs = hdr(2); nnl = hdr(3); nlibsl = hdr(4); nm = hdr(5); nme = hdr(6)
   ntc = hdr(7); ndegw = hdr(8); nwz = hdr(9); psdid = hdr(10); nfh = hdr(11)
   imode = hdr(12); npts = hdr(13); ncorr = hdr(14)
   allocate(n_load(nnl), libs_load(nlibsl), modes(nme), tcs(ntc), wz(nn))
   read(uin) n_load, libs_load, modes, tcs, wz
   allocate(phi(nn*nlibs, nm), natf(nm), Mm(nm), Km(nm), Cm(nm, nm), unod(nn))
   allocate(tscl(3, 3, nwz), tstd(3, nwz), umref(nwz), corr(ncorr, 3))
   allocate(wfc(nlibsl, ndegw, nnl), psi(nme, nnl, ndegw), fi(max(npts, 1)), fj(max(npts, 1)))
   read(uin) df
   read(uin) phi, natf, Mm, Km, Cm, unod, tscl, tstd, umref, corr, wfc, psi
   if (npts > 0) read(uin) fi(1:npts), fj(1:npts)
   close(uin)

   allocate(settings, wd, struct_data)
   ! --- structure (src/structure/StructureImpl.F90) ---
   call struct_data%SetTotalNOfNodes(nn)
   call struct_data%SetNOfNodalDOFs(nlibs)
   call struct_data%SetLoadedNodes(nnl, n_load)
   call struct_data%SetLoadedNodalDOFs(nlibsl, libs_load)
   call struct_data%SetModalInfo(nn*nlibs, nm, phi, natf)
   call struct_data%SetModalMatrices(nm, Mm, Km, Cm)
   call struct_data%SetKeptModes(modes)
   ! --- wind (src/wind/WindSetImpl.F90, WindPSDImpl.F90) ---
   wd%nz_ = nwz
   call wd%SetPSDType(psdid)
   call wd%setTurbComps(tcs, ntc)
   call wd%setWindDirections([1])
   call wd%SetNodalVel(unod)
   call wd%SetNodalWindZones(wz)
   call wd%SetTurbWindScales(tscl)
   call wd%SetTurbWindSDT(tstd)
   call wd%SetWZMeanWindVel(umref)
   call wd%SetSpatialNodalCorr(corr)
   call wd%SetWindFCoeffs(wfc)
   call wd%SetPhitimesC(psi)
   ! --- settings + derived sizes (BsaLibImpl.F90:469-535 does the same) ---
   settings%i_def_scaling_ = 1
   settings%i_compute_psd_ = 1
   settings%i_compute_bisp_ = 1
   settings%nfreqs_ = 2*(nfh - 1) + 1
   settings%df_ = df
   dimM_psd_ = nme*nme
   dimM_bisp_ = nme*nme*nme
   call setBsaFunctionLocalVars()

   nf = settings%nfreqs_
   allocate(f(nf))
   do i = 1, nf
      f(i) = real(i - nfh, real64) * df
   enddo

   ! nodal PSD table as mainClassic_ builds it (BsaClassicImpl.F90:103-119)
   allocate(suvw(nf, nnl*ntc))
   idx = 1
   do itc = 1, ntc
      suvw(:, idx:idx + nnl - 1) = wd%evalPSD(nf, f, nnl, n_load, 1, tcs(itc))
      idx = idx + nnl
   enddo

   if (trim(opt) == 'time') then
      t0 = omp_get_wtime()
      call getFM_full_tnm_vect_cls_(f, suvw, psd, bisp)
      t1 = omp_get_wtime()
      print '(a, 1x, es16.8, 1x, i0)', 'REFTIME', t1 - t0, omp_get_max_threads()
      stop
   endif

   open(newunit=uout, file=trim(fout), access='stream', form='unformatted', status='replace')
   write(uout) suvw
   if (iand(imode, 1) /= 0) then
      call getFM_full_tnm_vect_cls_(f, suvw, psd, bisp)
      write(uout) psd, bisp
      call getRM_full_vect_cls_(f, psd, bisp)
      write(uout) psd, bisp
   endif
   if (iand(imode, 4) /= 0) then
      ! ONLY_DIAG model: NM modal diagonal spectra (dimensions as validateAll_ sets them)
      if (allocated(psd)) deallocate(psd)
      if (allocated(bisp)) deallocate(bisp)
      dimM_psd_ = nme
      dimM_bisp_ = nme
      call getFM_diag_tnlm_vect_cls_(f, suvw, psd, bisp)
      write(uout) psd, bisp
      call getRM_diag_vect_cls_(f, psd, bisp)
      write(uout) psd, bisp
      dimM_psd_ = nme*nme
      dimM_bisp_ = nme*nme*nme
   endif
   if (iand(imode, 2) /= 0 .and. npts > 0) then
      allocate(bfm(nme*nme*nme, 1))
      do ip = 1, npts
         call getFM_full_tnm_scalar_msh_(bfm, fi(ip:ip), fj(ip:ip))
         write(uout) bfm(:, 1)
      enddo
   endif
   close(uout)
end program ref_harness
