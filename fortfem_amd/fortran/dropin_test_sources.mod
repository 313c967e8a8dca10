﻿!mod$ v1 sum:c27aadcb1cabcbcf
!need$ f425f7fa798ff9dd n fortfem_hip_types
module dropin_test_sources
use fortfem_hip_types,only:dp
contains
pure subroutine bubble_source(x,y,z,value)
real(8),intent(in)::x
real(8),intent(in)::y
real(8),intent(in)::z
real(8),intent(out)::value
end
pure subroutine zero_boundary(x,y,z,value)
real(8),intent(in)::x
real(8),intent(in)::y
real(8),intent(in)::z
real(8),intent(out)::value
end
pure subroutine sin_source(x,y,z,value)
real(8),intent(in)::x
real(8),intent(in)::y
real(8),intent(in)::z
real(8),intent(out)::value
end
end
