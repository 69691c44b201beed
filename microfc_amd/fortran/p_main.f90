!> Simulation driver with the structure of src/simulation/p_main.fpp:15-347:
!! read and check the case, read the initial condition and grid into the
!! device-resident q_cons_ts(1)%vf, set the library up, then step with the
!! configured TVD-RK scheme (dt fixed, end-of-run adjustment as :287), check
!! for NaN and save every t_step_save steps (:289-305), and record the mean
!! step time in time_data.dat (:239-270).
program p_main

    use iso_c_binding
    use m_mfc_abi
    use m_derived_types
    use m_global_parameters
    use m_start_up
    use m_rhs
    use m_data_output
    use m_time_steppers

    implicit none

    integer :: t_step
    real(kind(0d0)) :: time_avg

    call s_assign_default_values_to_user_inputs()
    call s_read_input_file()
    call s_check_input_file()
    call s_initialize_global_parameters_module()
    call s_initialize_data_output_module()

    call s_initialize_rhs_module()
    call s_initialize_time_steppers_module()
    call s_read_data_files(q_cons_ts(1)%vf)
    call s_populate_grid_variables_buffers()
    call s_initialize_weno_module()

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
        if (time_stepper == 1) then
            call s_1st_order_tvd_rk(t_step, time_avg)
        elseif (time_stepper == 2) then
            call s_2nd_order_tvd_rk(t_step, time_avg)
        elseif (time_stepper == 3) then
            call s_3rd_order_tvd_rk(t_step, time_avg)
        end if
        if (t_step == t_step_stop) then
            print *, "Final Time", time_avg
            call s_write_time_data(time_avg)
            exit
        else
            if ((mytime + dt) >= finaltime) dt = finaltime - mytime
            t_step = t_step + 1
        end if
        if (mod(t_step - t_step_start, t_step_save) == 0 .or. t_step == t_step_stop) then
            call s_check_for_nan(t_step)
            call s_write_data_files(q_cons_ts(1)%vf, t_step)
        end if
    end do

    call s_finalize_time_steppers_module()
    call s_finalize_rhs_module()

contains

    !> NaN guard of p_main.fpp:291-302: every interior value of every
    !! conservative variable, from a host copy of the device state; the
    !! first NaN prints its indices and stops with "Error".
    subroutine s_check_for_nan(t_step)
        integer, intent(in) :: t_step
        real(kind(0d0)), allocatable :: h(:, :, :)
        integer :: i, j, k, l
        allocate (h(-buff_size:m + buff_size, -bj:n + bj, -bk:p + bk))
        do i = 1, sys_size
            call s_variable_to_host(q_cons_ts(1)%vf(i), h)
            do l = 0, p
                do k = 0, n
                    do j = 0, m
                        if (h(j, k, l) /= h(j, k, l)) then
                            print *, j, k, 0, t_step, m, n
                            stop "Error"
                        end if
                    end do
                end do
            end do
        end do
        deallocate (h)
    end subroutine s_check_for_nan

end program p_main
