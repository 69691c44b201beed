!> Case parameters of the simulation host: the reference's &user_inputs
!! namelist variables (src/simulation/m_start_up.fpp:73-83, defaults as in
!! src/simulation/m_global_parameters.fpp:229-280) plus the 3D extension
!! (p, bc_z) and the choice of bulk arithmetic (bulk_precision), the derived
!! sizes and the device-resident field type.
module m_global_parameters

    use iso_c_binding
    use m_mfc_abi

    implicit none

    integer, parameter :: path_len = 400, name_len = 50
    integer, parameter :: num_fluids_max = 10, num_probes_max = 10
    real(kind(0d0)), parameter :: dflt_real = -1d6
    integer, parameter :: dflt_int = -100

    !> Derived types of the namelist (names and components as in
    !! src/common/m_derived_types.f90:40-43, 107-117).
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

    !> A ghosted field on the device: one allocation of sys_size variables,
    !! each (m+2b+1) x (n+2bj+1) x (p+2bk+1) doubles with x fastest, and the
    !! per-variable device pointers handed to the C ABI.
    type device_field
        type(c_ptr) :: base = c_null_ptr
        type(c_ptr), allocatable, dimension(:) :: v
        integer(c_size_t) :: nvar_elems = 0
    end type device_field

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
        buff_size = (weno_order - 1)/2 + 2
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

    subroutine s_allocate_device_field(f, ghosted)
        type(device_field), intent(inout) :: f
        logical, intent(in) :: ghosted
        integer :: i
        integer(c_intptr_t) :: addr
        f%nvar_elems = f_field_elems(ghosted)
        call s_mfc_check(mfc_device_alloc(8_c_size_t*f%nvar_elems*int(sys_size, c_size_t), f%base), 'mfc_device_alloc')
        allocate (f%v(sys_size))
        addr = transfer(f%base, addr)
        do i = 1, sys_size
            f%v(i) = transfer(addr + int(8, c_intptr_t)*int(f%nvar_elems, c_intptr_t)*int(i - 1, c_intptr_t), f%v(i))
        end do
    end subroutine s_allocate_device_field

    subroutine s_free_device_field(f)
        type(device_field), intent(inout) :: f
        if (c_associated(f%base)) call s_mfc_check(mfc_device_free(f%base), 'mfc_device_free')
        f%base = c_null_ptr
        if (allocated(f%v)) deallocate (f%v)
    end subroutine s_free_device_field

end module m_global_parameters
