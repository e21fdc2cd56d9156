﻿!mod$ v1 sum:a6a5cf63bc5f29be
!need$ f1de5abe9bfe2168 i iso_fortran_env
module driver_cb
use,intrinsic::iso_fortran_env,only:int64
use,intrinsic::iso_fortran_env,only:real64
integer(4)::cb_zones
integer(8)::cb_pts
real(8)::cb_sum
contains
subroutine cb_(f1,f2,brm,pdata)
real(8),intent(in)::f1(:)
real(8),intent(in)::f2(:)
real(8),intent(in)::brm(:,:)
class(*),intent(in),pointer::pdata
end
end
