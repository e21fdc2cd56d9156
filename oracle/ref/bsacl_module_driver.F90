!! Test-only driver: the REFERENCE's own Fortran binding of BsaCL (module BsaCL,
!! BsaCL/BsaCL_Fortran/bsacl.F90, bsacl_impl.F90, bsacl_c.F90, compiled unmodified
!! from where they lie) linked against libbsahip.so instead of the reference's
!! BsaCL runtime.  It calls the module in the order bsa_Run does
!! (src/bsa/BsaLibImpl.F90:269-313) and mainClassic_ does (BsaClassicImpl.F90:74-89),
!! so a passing run shows that the reference's binding works unchanged.
!!
!! Usage:  bsacl_module_driver <input.bin> <output.bin>
!! Input: the layout tests write for the reference harness; the evaluation
!! points are the uniform Classic grid.  Output: NM_EFF^3 real64 (bsacl_Run's
!! result), preceded by the int32 return codes of Init, InitDeviceMemory,
!! SetKernelID and Run.
program bsacl_module_driver
   use, intrinsic :: iso_fortran_env, only: int32, real64
   use BsaCL
   implicit none

   integer(int32) :: hdr(14)
   integer(int32) :: nn, nlibs, nnl, nlibsl, nm, nme, ntc, ndegw, nwz, psdid, nfh, ncorr, i
   integer(int32), allocatable, target :: n_load(:), libs_load(:), modes(:), tcs(:), wz(:)
   real(real64) :: df
   real(real64), allocatable, target :: phi(:, :), natf(:), Mm(:), Km(:), Cm(:, :), unod(:)
   real(real64), allocatable, target :: tscl(:, :, :), tstd(:, :), umref(:), corr(:, :)
   real(real64), allocatable, target :: wfc(:, :, :), psi(:, :, :), f(:)
   real(real64), pointer :: res(:)
   integer(IK) :: rc(4)
   character(len=512) :: fin, fout
   integer :: uin, uout

   call get_command_argument(1, fin)
   call get_command_argument(2, fout)
   open(newunit=uin, file=trim(fin), access='stream', form='unformatted', status='old')
   read(uin) hdr
   nn = hdr(1); nlibs = hdr(2); nnl = hdr(3); nlibsl = hdr(4); nm = hdr(5); nme = hdr(6)
   ntc = hdr(7); ndegw = hdr(8); nwz = hdr(9); psdid = hdr(10); nfh = hdr(11); ncorr = hdr(14)
   allocate(n_load(nnl), libs_load(nlibsl), modes(nme), tcs(ntc), wz(nn))
   read(uin) n_load, libs_load, modes, tcs, wz
   allocate(phi(nn*nlibs, nm), natf(nm), Mm(nm), Km(nm), Cm(nm, nm), unod(nn))
   allocate(tscl(3, 3, nwz), tstd(3, nwz), umref(nwz), corr(ncorr, 3))
   allocate(wfc(nlibsl, ndegw, nnl), psi(nme, nnl, ndegw))
   read(uin) df
   read(uin) phi, natf, Mm, Km, Cm, unod, tscl, tstd, umref, corr, wfc, psi
   close(uin)
   allocate(f(2*nfh - 1))
   do i = 1, 2*nfh - 1
      f(i) = real(i - nfh, real64) * df
   end do

   call bsacl_AcquirePSDId(psdid)
   call bsacl_AcquireStructModMat(phi, natf)
   call bsacl_AcquireModalMatrices(Mm, Cm, Km)
   call bsacl_AcquireLoadedNodesList(n_load)
   call bsacl_AcquireTotalNOfNodes(nn)
   call bsacl_AcquireUsedModesList(modes)
   call bsacl_AcquireWindCoeffs(wfc)
   call bsacl_AcquirePhiTimesCMat(psi)
   call bsacl_AcquireTurbComponentsList(tcs)
   call bsacl_AcquireNodalCorrelation(corr)
   call bsacl_AcquireWindNodalVelocities(unod)
   call bsacl_AcquireWindNodalWindZones(wz)
   call bsacl_AcquireWindTurbScales(tscl, nwz)
   call bsacl_AcquireWindTurbStd(tstd, nwz)
   call bsacl_SetDeviceType(BSACL_DEVICE_TYPE_GPU)
   rc(1) = bsacl_Init(1)
   rc(2) = bsacl_InitDeviceMemory()
   call bsacl_AcquireComputationFreqs(0, size(f), f, size(f), f)
   rc(3) = bsacl_SetKernelID(2)
   allocate(res(nme**3))
   res = 0.0_real64
   rc(4) = bsacl_Run(0, res)
   call bsacl_Finalise()

   open(newunit=uout, file=trim(fout), access='stream', form='unformatted', status='replace')
   write(uout) int(rc, int32), res
   close(uout)
end program bsacl_module_driver
