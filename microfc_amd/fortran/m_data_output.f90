!> Run-time information and serial output (src/simulation/m_data_output.fpp):
!! run_time.inf with the ICFL of each step (abort when it is NaN or above 1,
!! :296-305), the NaN guard before each save (p_main.fpp:291-305), and the
!! serial writer of <case_dir>/p_all/p0/<t_step>/ (:316-381).
module m_data_output

    use iso_c_binding
    use m_mfc_abi
    use m_global_parameters
    use m_start_up, only: s_abort, f_t_step_dir

    implicit none

contains

    subroutine s_open_run_time_information_file()
        open (11, FILE=trim(case_dir)//'/run_time.inf', FORM='formatted', STATUS='replace')
        write (11, '(A)') '    Time-step        Time            ICFL Max'
    end subroutine s_open_run_time_information_file

    subroutine s_write_run_time_information(t_step, icfl)
        integer, intent(in) :: t_step
        real(kind(0d0)), intent(in) :: icfl
        write (11, '(13X,I8,14X,F10.6,13X,F9.6)') t_step, t_step*dt, icfl
        if (icfl /= icfl) then
            print '(A)', 'ICFL is NaN. Exiting ...'
            stop 1
        elseif (icfl > 1d0) then
            print '(A)', 'ICFL is greater than 1.0. Exiting ...'
            print *, 'icfl', icfl
            stop 1
        end if
    end subroutine s_write_run_time_information

    subroutine s_close_run_time_information_file()
        close (11)
    end subroutine s_close_run_time_information_file

    !> NaN guard and serial writer; q is the host copy of q_cons_ts(1).
    subroutine s_write_serial_data_files(q, t_step)
        real(kind(0d0)), intent(in) :: q(-buff_size:, -bj:, -bk:, :)
        integer, intent(in) :: t_step
        character(len=path_len + 3*name_len) :: fp
        integer :: i, j, k, l
        do i = 1, sys_size
            do l = 0, p
                do k = 0, n
                    do j = 0, m
                        if (q(j, k, l, i) /= q(j, k, l, i)) then
                            print *, j, k, l, t_step, m, n, p
                            stop "Error"
                        end if
                    end do
                end do
            end do
        end do
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
        do i = 1, sys_size
            write (fp, '(A,I0,A)') trim(f_t_step_dir(t_step))//'/q_cons_vf', i, '.dat'
            open (2, FILE=trim(fp), FORM='unformatted', STATUS='replace')
            write (2) q(0:m, 0:n, 0:p, i); close (2)
        end do
    end subroutine s_write_serial_data_files

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
