!! Reference module procedures and constants that need no device: the text
!! exports of PSD and bispectrum series and the settings queries.  Writes
!! psd.txt and bisp.txt in the current directory and prints the queries.
program api_extras
   use BsaLib
   implicit none
   real(bsa_real_t) :: psd(3, 2), bisp(2, 2, 3), f(3)
   integer :: i
   psd = reshape([(0.5_RK * i, i = 1, 6)], [3, 2])
   bisp = reshape([(0.25_RK * i, i = 1, 12)], [2, 2, 3])
   f = [0.0_RK, 0.1_RK, 0.2_RK]
   call bsa_Init()
   call bsa_setExportInCurrDir()
   call bsa_exportPSDToFile('psd.txt', psd, f)
   call bsa_exportPSDToFile('psd_nof.txt', psd)
   call bsa_exportBispToFile('bisp.txt', bisp)
   call bsa_setPremeshMode(7)
   call bsa_setPolicyIDValidationValues(5, 2, 2, 4, 4)
   call bsa_setOutUnit(6)
   print '(a, 1x, l1)', 'FULL', bsa_isFullComp()
   call bsa_setSpectraExtension(1)
   print '(a, 1x, l1)', 'FULL', bsa_isFullComp()
   print '(a, 1x, es25.17, 1x, es25.17)', 'CST', CST_PIt2, CST_PIt3d2
   print '(a, 1x, a)', 'NAMES', BSA_EXPORT_M3MR_CLS_FNAME//' '//BRM_EXPORT_FNAME_MSH
   call bsa_Finalise()
end program api_extras
