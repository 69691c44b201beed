!> Run-time information and serial output (src/simulation/m_data_output.fpp):
!! run_time.inf with the reference's header, one line of stability criteria
!! per step (ICFL, and VCFL and Rc in viscous runs) and the closing summary;
!! the abort when ICFL is NaN or above 1 (:296-305); the serial writer of
!! <case_dir>/p_all/p0/<t_step>/ (:316-381), the parallel_io writer of
!! restart_data/lustre_<t_step>.dat (:477-540) and time_data.dat.
module m_data_output

    use iso_c_binding
    use m_mfc_abi
    use m_derived_types
    use m_global_parameters
    use m_start_up, only: s_abort, f_t_step_dir

    implicit none

    private; public :: s_initialize_data_output_module, s_open_run_time_information_file, &
        s_write_run_time_information, s_write_serial_data_files, s_write_parallel_data_files, &
        s_write_data_files, s_close_run_time_information_file, s_write_time_data, f_viscous

    !> Stability criteria extrema over all the time steps (:74-76).
    real(kind(0d0)) :: icfl_max, vcfl_max, Rc_min

    integer, parameter :: run_time_unit = 11

contains

    !> any(Re_size > 0): some fluid has a positive Reynolds number.
    logical function f_viscous()
        f_viscous = any(fluid_pp(1:num_fluids)%Re(1) > 0d0) .or. any(fluid_pp(1:num_fluids)%Re(2) > 0d0)
    end function f_viscous

    !> s_initialize_data_output_module (:715-740): the extrema's starting values.
    subroutine s_initialize_data_output_module()
        icfl_max = 0d0
        vcfl_max = 0d0
        Rc_min = 1d3
    end subroutine s_initialize_data_output_module

    !> s_open_run_time_information_file (:90-146): appends to
    !! <case_dir>/run_time.inf, writing the description and date header when
    !! the file is new, then the table header of the criteria this run has.
    subroutine s_open_run_time_information_file()
        character(len=path_len + name_len) :: file_path
        character(len=8) :: file_date
        logical :: file_exist
        file_path = trim(case_dir)//'/run_time.inf'
        inquire (FILE=trim(file_path), EXIST=file_exist)
        open (run_time_unit, FILE=trim(file_path), FORM='formatted', POSITION='append', STATUS='unknown')
        if (.not. file_exist) then
            write (run_time_unit, '(A)') 'Description: Stability information at '// &
                'each time-step of the simulation. This'
            write (run_time_unit, '(13X,A)') 'data is composed of the inviscid '// &
                'Courant–Friedrichs–Lewy (ICFL)'
            write (run_time_unit, '(13X,A)') 'number and the cell Reynolds (Rc) '// &
                'number. Please note that only'
            write (run_time_unit, '(13X,A)') 'those stability conditions pertinent '// &
                'to the physics included in'
            write (run_time_unit, '(13X,A)') 'the current computation are displayed.'
            call date_and_time(DATE=file_date)
            write (run_time_unit, '(A)') 'Date: '//file_date(5:6)//'/'//file_date(7:8)//'/'//file_date(3:4)
        end if
        write (run_time_unit, '(A)') ''; write (run_time_unit, '(A)') ''
        if (f_viscous()) then
            write (run_time_unit, '(A)') '==== Time-steps ====== Time ======= ICFL '// &
                'Max ==== VCFL Max ====== Rc Min ======='
        else
            write (run_time_unit, '(A)') '=========== Time-steps ============== Time '// &
                '============== ICFL Max ============='
        end if
    end subroutine s_open_run_time_information_file

    !> s_write_run_time_information(q_prim_vf, t_step) (:197-310): the
    !! stability criteria of the primitive field over the interior (computed
    !! on the device), their extrema over the run, one table line, and the
    !! abort when ICFL is NaN or above 1.
    subroutine s_write_run_time_information(q_prim_vf, t_step)
        type(scalar_field), dimension(sys_size), intent(IN) :: q_prim_vf
        integer, intent(IN) :: t_step
        type(c_ptr), dimension(sys_size), target :: pp
        real(c_double) :: crit(3)
        real(kind(0d0)) :: icfl_max_glb, vcfl_max_glb, Rc_min_glb
        call s_field_pointers(q_prim_vf, pp)
        call s_mfc_check(mfc_stability_criteria(c_loc(pp), dt, crit), 'mfc_stability_criteria')
        icfl_max_glb = crit(1)
        vcfl_max_glb = crit(2)
        Rc_min_glb = crit(3)
        if (icfl_max_glb > icfl_max) icfl_max = icfl_max_glb
        if (f_viscous()) then
            if (vcfl_max_glb > vcfl_max) vcfl_max = vcfl_max_glb
            if (Rc_min_glb < Rc_min) Rc_min = Rc_min_glb
            write (run_time_unit, '(6X,I8,6X,F10.6,6X,F9.6,6X,F9.6,6X,F10.6)') &
                t_step, t_step*dt, icfl_max_glb, vcfl_max_glb, Rc_min_glb
        else
            write (run_time_unit, '(13X,I8,14X,F10.6,13X,F9.6)') t_step, t_step*dt, icfl_max_glb
        end if
        if (icfl_max_glb /= icfl_max_glb) then
            print '(A)', 'ICFL is NaN. Exiting ...'
            stop 1
        elseif (icfl_max_glb > 1d0) then
            print '(A)', 'ICFL is greater than 1.0. Exiting ...'
            print *, 'icfl', icfl_max_glb
            stop 1
        end if
    end subroutine s_write_run_time_information

    !> s_close_run_time_information_file (:686-705): the extrema over all
    !! the steps and the run time, then closes the file.
    subroutine s_close_run_time_information_file()
        real(kind(0d0)) :: run_time
        write (run_time_unit, '(A)') '----------------------------------------'// &
            '----------------------------------------'
        write (run_time_unit, '(A)') ''
        write (run_time_unit, '(A,F9.6)') 'ICFL Max: ', icfl_max
        if (f_viscous()) write (run_time_unit, '(A,F9.6)') 'VCFL Max: ', vcfl_max
        if (f_viscous()) write (run_time_unit, '(A,F10.6)') 'Rc Min: ', Rc_min
        call cpu_time(run_time)
        write (run_time_unit, '(A)') ''
        write (run_time_unit, '(A,I0,A)') 'Run-time: ', int(anint(run_time)), 's'
        write (run_time_unit, '(A)') '========================================'// &
            '========================================'
        close (run_time_unit)
    end subroutine s_close_run_time_information_file

    !> s_write_serial_data_files(q_cons_vf, t_step) (:316-381): the cell
    !! boundaries and the interior of every conservative variable, copied
    !! from the device, as Fortran unformatted records.
    subroutine s_write_serial_data_files(q_cons_vf, t_step)
        type(scalar_field), dimension(sys_size), intent(IN) :: q_cons_vf
        integer, intent(IN) :: t_step
        character(len=path_len + 3*name_len) :: fp
        real(kind(0d0)), allocatable :: h(:, :, :)
        integer :: i
        call execute_command_line('mkdir -p '//trim(f_t_step_dir(t_step)))
        open (2, FILE=trim(f_t_step_dir(t_step))//'/x_cb.dat', FORM='unformatted', STATUS='replace')
        write (2) x_cb(-1:m); close (2)
        if (n > 0) then
            open (2, FILE=trim(f_t_step_dir(t_step))//'/y_cb.dat', FORM='unformatted', STATUS='replace')
            write (2) y_cb(-1:n); close (2)
        end if
        if (p > 0) then
            open (2, FILE=trim(f_t_step_dir(t_step))//'/z_cb.dat', FORM='unformatted', STATUS='replace')
            write (2) z_cb(-1:p); close (2)
        end if
        allocate (h(-buff_size:m + buff_size, -bj:n + bj, -bk:p + bk))
        do i = 1, sys_size
            call s_variable_to_host(q_cons_vf(i), h)
            write (fp, '(A,I0,A)') trim(f_t_step_dir(t_step))//'/q_cons_vf', i, '.dat'
            open (2, FILE=trim(fp), FORM='unformatted', STATUS='replace')
            write (2) h(0:m, 0:n, 0:p); close (2)
        end do
        deallocate (h)
    end subroutine s_write_serial_data_files

    !> s_write_parallel_data_files(q_cons_vf, t_step) (:477-540) on one rank:
    !! restart_data/lustre_<t_step>.dat with the interior of every variable as
    !! raw doubles, variable i at byte offset (m+1)(n+1)(p+1)*8*(i-1).
    subroutine s_write_parallel_data_files(q_cons_vf, t_step)
        type(scalar_field), dimension(sys_size), intent(IN) :: q_cons_vf
        integer, intent(IN) :: t_step
        character(len=path_len + 2*name_len) :: fp
        real(kind(0d0)), allocatable :: h(:, :, :)
        integer :: i
        call execute_command_line('mkdir -p '//trim(case_dir)//'/restart_data')
        write (fp, '(A,I0,A)') trim(case_dir)//'/restart_data/lustre_', t_step, '.dat'
        open (2, FILE=trim(fp), FORM='unformatted', ACCESS='stream', STATUS='replace')
        allocate (h(-buff_size:m + buff_size, -bj:n + bj, -bk:p + bk))
        do i = 1, sys_size
            call s_variable_to_host(q_cons_vf(i), h)
            write (2) h(0:m, 0:n, 0:p)
        end do
        close (2)
        deallocate (h)
    end subroutine s_write_parallel_data_files

    !> The writer parallel_io selects (the reference's s_write_data_files
    !! procedure pointer).
    subroutine s_write_data_files(q_cons_vf, t_step)
        type(scalar_field), dimension(sys_size), intent(IN) :: q_cons_vf
        integer, intent(IN) :: t_step
        if (parallel_io) then
            call s_write_parallel_data_files(q_cons_vf, t_step)
        else
            call s_write_serial_data_files(q_cons_vf, t_step)
        end if
    end subroutine s_write_data_files

    !> time_data.dat as p_main.fpp:255-270 appends it.
    subroutine s_write_time_data(time_final)
        real(kind(0d0)), intent(in) :: time_final
        logical :: file_exists
        inquire (FILE='time_data.dat', EXIST=file_exists)
        if (file_exists) then
            open (1, file='time_data.dat', position='append', status='old')
        else
            open (1, file='time_data.dat', status='new')
        end if
        write (1, *) 1, time_final
        close (1)
    end subroutine s_write_time_data

end module m_data_output
