!> Derived types of the simulation with the reference's names and components
!! (src/common/m_derived_types.f90:25-37, 40-43, 107-117).
!!
!! A scalar_field is one variable of a field, sf(ix%beg:ix%end, iy%beg:iy%end)
!! as the reference allocates it (m_time_steppers.fpp:81-112).  Here sf points
!! at device memory (HBM) that only the HIP library reads and writes: host
!! code hands c_loc(sf(lbound(sf, 1), lbound(sf, 2))) across the C ABI and
!! never dereferences it.  The reference's fields are at most 2D; a 3D field
!! (the p > 0 extension) keeps the same rank-2 pointer with the y and z
!! extents folded into the second index, sf(-b:m+b, -b:(n+2b+1)(p+2b+1)-b-1).
module m_derived_types

    implicit none

    type scalar_field
        real(kind(0d0)), pointer, dimension(:, :) :: sf => null()
    end type scalar_field

    type vector_field
        type(scalar_field), allocatable, dimension(:) :: vf
    end type vector_field

    type int_bounds_info
        integer :: beg
        integer :: end
    end type int_bounds_info

    type physical_parameters
        real(kind(0d0)) :: gamma
        real(kind(0d0)) :: pi_inf
        real(kind(0d0)), dimension(2) :: Re
    end type physical_parameters

    type probe_parameters
        real(kind(0d0)) :: x
        real(kind(0d0)) :: y
    end type probe_parameters

end module m_derived_types
