!> Simulation driver with the structure of src/simulation/p_main.fpp:15-347:
!! read and check the case, read the initial condition and grid, set the
!! library up, then step (dt fixed, end-of-run adjustment as :287), save
!! with a NaN guard every t_step_save steps and record the mean step time.
program p_main

    use iso_c_binding
    use m_mfc_abi
    use m_global_parameters
    use m_start_up
    use m_rhs
    use m_time_steppers
    use m_data_output

    implicit none

    real(kind(0d0)), allocatable, target :: q_host(:, :, :, :)
    integer :: t_step
    real(kind(0d0)) :: time_avg, icfl
    integer(c_size_t) :: bytes

    call s_assign_default_values_to_user_inputs()
    call s_read_input_file()
    call s_check_input_file()
    call s_initialize_global_parameters_module()

    allocate (q_host(-buff_size:m + buff_size, -bj:n + bj, -bk:p + bk, sys_size))
    call s_read_serial_data_files(q_host)
    call s_populate_grid_variables_buffers()
    call s_initialize_rhs_module()
    call s_initialize_time_steppers_module()

    bytes = 8_c_size_t*f_field_elems(.true.)*int(sys_size, c_size_t)
    call s_mfc_check(mfc_memcpy_h2d(q_cons_ts(1)%base, c_loc(q_host), bytes), 'mfc_memcpy_h2d')
    if (run_time_info) call s_open_run_time_information_file()

    t_step = t_step_start
    if (t_step == 0) then
        mytime = 0d0
    else
        mytime = t_step*dt
    end if
    finaltime = t_step_stop*dt
    time_avg = 0d0

    do
        print '(" [",I3,"%]  Time step ",I8," of ",I0," @ t_step = ",I0)', &
            int(100*(real(t_step + 1)/(t_step_stop - t_step_start + 1))), &
            t_step - t_step_start + 1, t_step_stop - t_step_start + 1, t_step
        mytime = mytime + dt
        call s_tvd_rk(t_step, time_avg, icfl)
        if (run_time_info) call s_write_run_time_information(t_step, icfl)
        if (t_step == t_step_stop) then
            print *, "Final Time", time_avg
            call s_write_time_data(time_avg)
            exit
        else
            if ((mytime + dt) >= finaltime) dt = finaltime - mytime
            t_step = t_step + 1
        end if
        if (mod(t_step - t_step_start, t_step_save) == 0 .or. t_step == t_step_stop) then
            call s_mfc_check(mfc_memcpy_d2h(c_loc(q_host), q_cons_ts(1)%base, bytes), 'mfc_memcpy_d2h')
            call s_write_serial_data_files(q_host, t_step)
        end if
    end do

    if (run_time_info) call s_close_run_time_information_file()
    call s_finalize_time_steppers_module()
    call s_finalize_rhs_module()
    deallocate (q_host)

end program p_main
