!> Case parameters of the simulation host: the reference's &user_inputs
!! namelist variables (src/simulation/m_start_up.fpp:73-83, defaults as in
!! src/simulation/m_global_parameters.fpp:229-280) plus the 3D extension
!! (p, bc_z) and the choice of bulk arithmetic (bulk_precision), the derived
!! sizes, and the device allocation of the reference's field types.
module m_global_parameters

    use iso_c_binding
    use m_mfc_abi
    use m_derived_types

    implicit none

    integer, parameter :: path_len = 400, name_len = 50
    integer, parameter :: num_fluids_max = 10, num_probes_max = 10
    real(kind(0d0)), parameter :: dflt_real = -1d6
    integer, parameter :: dflt_int = -100
    real(kind(0d0)), parameter :: sgm_eps = 1d-16

    ! namelist variables
    character(len=path_len) :: case_dir
    logical :: run_time_info, weno_flat, riemann_flat, cu_mpi, probe_wrt, weno_Re_flux, parallel_io
    integer :: m, n, p
    real(kind(0d0)) :: dt
    integer :: t_step_start, t_step_stop, t_step_save, t_step_old
    integer :: num_fluids, time_stepper, weno_order, fd_order, num_probes, precision
    integer :: bulk_precision
    character(len=path_len) :: mlp_weights
    real(kind(0d0)) :: weno_eps
    type(int_bounds_info) :: bc_x, bc_y, bc_z
    type(physical_parameters), dimension(num_fluids_max) :: fluid_pp
    type(probe_parameters), dimension(num_probes_max) :: probe

    ! derived
    integer :: num_dims, sys_size, buff_size, bj, bk
    integer :: cont_end, mom_beg, E_idx, adv_beg
    real(kind(0d0)), allocatable, dimension(:) :: x_cb, y_cb, z_cb, dx, dy, dz
    real(kind(0d0)) :: mytime, finaltime

contains

    subroutine s_assign_default_values_to_user_inputs()
        integer :: i
        case_dir = '.'
        run_time_info = .false.
        weno_flat = .true.; riemann_flat = .true.; cu_mpi = .false.
        probe_wrt = .false.; weno_Re_flux = .false.; parallel_io = .false.
        m = dflt_int; n = 0; p = 0
        dt = dflt_real
        t_step_start = dflt_int; t_step_stop = dflt_int; t_step_save = dflt_int; t_step_old = dflt_int
        num_fluids = dflt_int; time_stepper = dflt_int; weno_order = dflt_int
        fd_order = dflt_int; num_probes = dflt_int; precision = 2
        bulk_precision = mfc_prec_fp64
        mlp_weights = ''
        weno_eps = dflt_real
        bc_x%beg = dflt_int; bc_x%end = dflt_int
        bc_y%beg = dflt_int; bc_y%end = dflt_int
        bc_z%beg = dflt_int; bc_z%end = dflt_int
        do i = 1, num_fluids_max
            fluid_pp(i)%gamma = dflt_real
            fluid_pp(i)%pi_inf = dflt_real
            fluid_pp(i)%Re(:) = dflt_real
        end do
        do i = 1, num_probes_max
            probe(i)%x = dflt_real
            probe(i)%y = dflt_real
        end do
    end subroutine s_assign_default_values_to_user_inputs

    !> Index layout and buffer size (m_global_parameters.fpp:285-381).
    subroutine s_initialize_global_parameters_module()
        num_dims = 1 + min(1, n) + min(1, p)
        cont_end = num_fluids
        mom_beg = num_fluids + 1
        E_idx = num_fluids + num_dims + 1
        adv_beg = E_idx + 1
        sys_size = E_idx + num_fluids
        ! m_global_parameters.fpp:355-359
        if (any(fluid_pp(1:num_fluids)%Re(1) > 0d0) .or. any(fluid_pp(1:num_fluids)%Re(2) > 0d0)) then
            buff_size = 2*((weno_order - 1)/2) + 2
        else
            buff_size = (weno_order - 1)/2 + 2
        end if
        bj = 0; bk = 0
        if (n > 0) bj = buff_size
        if (p > 0) bk = buff_size
        allocate (x_cb(-1 - buff_size:m + buff_size), dx(-buff_size:m + buff_size))
        if (n > 0) allocate (y_cb(-1 - buff_size:n + buff_size), dy(-buff_size:n + buff_size))
        if (p > 0) allocate (z_cb(-1 - buff_size:p + buff_size), dz(-buff_size:p + buff_size))
    end subroutine s_initialize_global_parameters_module

    !> The parameter block handed to mfc_rhs_init.
    function f_mfc_params() result(prm)
        type(mfc_params) :: prm
        integer :: i
        prm%m = m; prm%n = n; prm%p = p
        prm%num_fluids = num_fluids
        prm%weno_order = weno_order
        prm%time_stepper = time_stepper
        prm%t_step_stop = t_step_stop
        prm%bc_x_beg = bc_x%beg; prm%bc_x_end = bc_x%end
        prm%bc_y_beg = bc_y%beg; prm%bc_y_end = bc_y%end
        prm%bc_z_beg = bc_z%beg; prm%bc_z_end = bc_z%end
        prm%weno_eps = weno_eps
        prm%gammas = 0d0; prm%pi_infs = 0d0; prm%Re = dflt_real
        do i = 1, min(num_fluids, num_fluids_max)
            prm%gammas(i) = fluid_pp(i)%gamma
            prm%pi_infs(i) = fluid_pp(i)%pi_inf
            prm%Re(:, i) = fluid_pp(i)%Re(:)
        end do
        prm%precision = bulk_precision
        prm%weno_Re_flux = merge(1, 0, weno_Re_flux)
    end function f_mfc_params

    !> Number of doubles per variable of a ghosted (or interior) field.
    integer(c_size_t) function f_field_elems(ghosted)
        logical, intent(in) :: ghosted
        if (ghosted) then
            f_field_elems = int(m + 2*buff_size + 1, c_size_t)*int(n + 2*bj + 1, c_size_t)*int(p + 2*bk + 1, c_size_t)
        else
            f_field_elems = int(m + 1, c_size_t)*int(n + 1, c_size_t)*int(p + 1, c_size_t)
        end if
    end function f_field_elems

    !> Bounds of one variable of a ghosted (or interior) field as the
    !! reference allocates it (m_time_steppers.fpp:67-112): (-b:m+b, -b:n+b)
    !! and (0:m, 0:n); a 3D field folds its z extent into the second index.
    subroutine s_field_bounds(ghosted, lb, ub)
        logical, intent(in) :: ghosted
        integer, intent(out) :: lb(2), ub(2)
        if (ghosted) then
            lb = [-buff_size, -bj]
            ub = [m + buff_size, -bj + (n + 2*bj + 1)*(p + 2*bk + 1) - 1]
        else
            lb = [0, 0]
            ub = [m, (n + 1)*(p + 1) - 1]
        end if
    end subroutine s_field_bounds

    !> Allocates the size(f) variables of a field in one device block
    !! (returned in base) and points f(i)%sf at variable i, with the
    !! reference's bounds.  Only the library touches the memory.
    subroutine s_allocate_field(f, ghosted, base)
        type(scalar_field), dimension(:), intent(inout) :: f
        logical, intent(in) :: ghosted
        type(c_ptr), intent(out) :: base
        real(kind(0d0)), pointer :: tmp(:, :)
        integer :: lb(2), ub(2), i
        integer(c_intptr_t) :: addr
        integer(c_size_t) :: ne
        type(c_ptr) :: v
        call s_field_bounds(ghosted, lb, ub)
        ne = f_field_elems(ghosted)
        call s_mfc_check(mfc_device_alloc(8_c_size_t*ne*int(size(f), c_size_t), base), 'mfc_device_alloc')
        addr = transfer(base, addr)
        do i = 1, size(f)
            v = transfer(addr + 8_c_intptr_t*int(ne, c_intptr_t)*int(i - 1, c_intptr_t), v)
            call c_f_pointer(v, tmp, [ub(1) - lb(1) + 1, ub(2) - lb(2) + 1])
            f(i)%sf(lb(1):, lb(2):) => tmp
        end do
    end subroutine s_allocate_field

    subroutine s_free_field(f, base)
        type(scalar_field), dimension(:), intent(inout) :: f
        type(c_ptr), intent(inout) :: base
        integer :: i
        do i = 1, size(f)
            nullify (f(i)%sf)
        end do
        if (c_associated(base)) call s_mfc_check(mfc_device_free(base), 'mfc_device_free')
        base = c_null_ptr
    end subroutine s_free_field

    !> The device pointers of a field's variables, c_loc of each variable's
    !! first element: the layout the C ABI takes (include/mfc_rhs.h).
    subroutine s_field_pointers(f, ptrs)
        type(scalar_field), dimension(:), intent(in) :: f
        type(c_ptr), dimension(:), intent(out) :: ptrs
        integer :: i
        do i = 1, size(f)
            ptrs(i) = c_loc(f(i)%sf(lbound(f(i)%sf, 1), lbound(f(i)%sf, 2)))
        end do
    end subroutine s_field_pointers

    !> Host copy of one variable: all of sf, ghosts included, into the
    !! first size(v%sf) elements of host (x fastest).
    subroutine s_variable_to_host(v, host)
        type(scalar_field), intent(in) :: v
        real(kind(0d0)), intent(out), target :: host(*)
        call s_mfc_check(mfc_memcpy_d2h(c_loc(host(1)), c_loc(v%sf(lbound(v%sf, 1), lbound(v%sf, 2))), &
                                        8_c_size_t*int(size(v%sf), c_size_t)), 'mfc_memcpy_d2h')
    end subroutine s_variable_to_host

    !> Device copy of one variable from the first size(v%sf) elements of host.
    subroutine s_variable_to_device(host, v)
        real(kind(0d0)), intent(in), target :: host(*)
        type(scalar_field), intent(inout) :: v
        call s_mfc_check(mfc_memcpy_h2d(c_loc(v%sf(lbound(v%sf, 1), lbound(v%sf, 2))), c_loc(host(1)), &
                                        8_c_size_t*int(size(v%sf), c_size_t)), 'mfc_memcpy_h2d')
    end subroutine s_variable_to_device

end module m_global_parameters
