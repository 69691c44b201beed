!> One call of the drop-in s_compute_rhs with the reference's argument types:
!! reads a case like p_main, declares q_cons_vf, q_prim_vf and rhs_vf as the
!! reference's type(scalar_field), dimension(sys_size) arrays (their sf on
!! the device), calls s_compute_rhs(q_cons_vf, q_prim_vf, rhs_vf, t_step_start)
!! and writes the interior of rhs_vf (and, with run_time_info, the ghosted
!! q_prim_vf) to rhs.dat / q_prim.dat as raw little-endian doubles, variable
!! after variable, x fastest.  It exists to check the drop-in interface.
program compute_rhs_check

    use iso_c_binding
    use m_mfc_abi
    use m_derived_types
    use m_global_parameters
    use m_start_up
    use m_rhs

    implicit none

    type(scalar_field), allocatable, dimension(:) :: q_cons_vf, q_prim_vf, rhs_vf
    type(c_ptr) :: cons_base, prim_base, rhs_base
    real(kind(0d0)), allocatable :: h(:, :, :), r(:, :, :)
    integer :: i

    call s_assign_default_values_to_user_inputs()
    call s_read_input_file()
    call s_check_input_file()
    call s_initialize_global_parameters_module()
    call s_initialize_rhs_module()
    allocate (q_cons_vf(1:sys_size), q_prim_vf(1:sys_size), rhs_vf(1:sys_size))
    call s_allocate_field(q_cons_vf, .true., cons_base)
    call s_allocate_field(q_prim_vf, .true., prim_base)
    call s_allocate_field(rhs_vf, .false., rhs_base)
    call s_read_data_files(q_cons_vf)
    call s_populate_grid_variables_buffers()
    call s_initialize_weno_module()

    call s_compute_rhs(q_cons_vf, q_prim_vf, rhs_vf, t_step_start)

    allocate (r(0:m, 0:n, 0:p), h(-buff_size:m + buff_size, -bj:n + bj, -bk:p + bk))
    open (3, FILE='rhs.dat', FORM='unformatted', ACCESS='stream', STATUS='replace')
    do i = 1, sys_size
        call s_variable_to_host(rhs_vf(i), r)
        write (3) r
    end do
    close (3)
    if (run_time_info) then
        open (3, FILE='q_prim.dat', FORM='unformatted', ACCESS='stream', STATUS='replace')
        do i = 1, sys_size
            call s_variable_to_host(q_prim_vf(i), h)
            write (3) h
        end do
        close (3)
    end if
    call s_free_field(q_cons_vf, cons_base)
    call s_free_field(q_prim_vf, prim_base)
    call s_free_field(rhs_vf, rhs_base)
    call s_finalize_rhs_module()

end program compute_rhs_check
