!> Input file, initial condition and grid (src/simulation/m_start_up.fpp):
!! the &user_inputs namelist, the serial data files
!! <case_dir>/p_all/p0/<t_step_start>/{x_cb,y_cb,z_cb,q_cons_vf<i>}.dat
!! (Fortran unformatted sequential, as the reference's pre_process writes
!! them) and the ghost buffers of the grid.
module m_start_up

    use iso_c_binding
    use m_mfc_abi
    use m_derived_types
    use m_global_parameters

    implicit none

contains

    !> s_read_input_file (m_start_up.fpp:64-113): the reference's namelist
    !! variables plus p, bc_z and bulk_precision.
    subroutine s_read_input_file()
        character(len=name_len) :: file_path = './simulation.inp'
        logical :: file_exist
        integer :: ios
        namelist /user_inputs/ case_dir, run_time_info, m, n, p, dt, &
            t_step_start, t_step_stop, t_step_save, &
            num_fluids, time_stepper, &
            weno_eps, weno_flat, riemann_flat, cu_mpi, &
            bc_x, bc_y, bc_z, fluid_pp, probe_wrt, &
            fd_order, probe, num_probes, t_step_old, &
            weno_Re_flux, weno_order, precision, parallel_io, bulk_precision, mlp_weights
        inquire (FILE=trim(file_path), EXIST=file_exist)
        if (.not. file_exist) then
            print '(A)', trim(file_path)//' is missing. Exiting ...'
            stop 1
        end if
        open (1, FILE=trim(file_path), FORM='formatted', ACTION='read', STATUS='old')
        read (1, NML=user_inputs, iostat=ios); close (1)
        if (ios /= 0) then
            print '(A,I0)', 'Error reading the &user_inputs namelist, iostat = ', ios
            stop 1
        end if
    end subroutine s_read_input_file

    !> The checks of s_check_input_file (m_start_up.fpp:117-290) that the
    !! library does not repeat; mfc_rhs_init checks the rest.
    subroutine s_check_input_file()
        integer :: i, j
        if (dt <= 0d0) call s_abort('Unsupported value of dt. Exiting ...')
        if (t_step_start < 0) call s_abort('Unsupported value of t_step_start. Exiting ...')
        if (t_step_stop <= t_step_start) &
            call s_abort('Unsupported combination of values of t_step_start and t_step_stop. Exiting ...')
        if (t_step_save > t_step_stop - t_step_start) &
            call s_abort('Unsupported combination of values of t_step_start, t_step_stop and t_step_save. Exiting ...')
        if (num_fluids < 1 .or. num_fluids > num_fluids_max) call s_abort('Unsupported value of num_fluids. Exiting ...')
        if (probe_wrt) call s_abort('probe_wrt is not supported by this build. Exiting ...')
        if (cu_mpi) call s_abort('Unsupported value of cu_mpi. Exiting ...')
        if (n == 0 .and. bc_y%beg == dflt_int) then
            bc_y%beg = -3; bc_y%end = -3
        end if
        if (p == 0 .and. bc_z%beg == dflt_int) then
            bc_z%beg = -3; bc_z%end = -3
        end if
    end subroutine s_check_input_file

    subroutine s_abort(msg)
        character(len=*), intent(in) :: msg
        print '(A)', msg
        stop 1
    end subroutine s_abort

    function f_t_step_dir(t_step) result(dir)
        integer, intent(in) :: t_step
        character(len=path_len + 2*name_len) :: dir
        write (dir, '(A,I0)') trim(case_dir)//'/p_all/p0/', t_step
    end function f_t_step_dir

    subroutine s_read_record(path, a, count)
        character(len=*), intent(in) :: path
        integer, intent(in) :: count
        real(kind(0d0)), intent(out) :: a(count)
        logical :: file_exist
        inquire (FILE=trim(path), EXIST=file_exist)
        if (.not. file_exist) call s_abort(trim(path)//' is missing. Exiting ...')
        open (2, FILE=trim(path), FORM='unformatted', ACTION='read', STATUS='old')
        read (2) a; close (2)
    end subroutine s_read_record

    !> s_read_serial_data_files(q_cons_vf) (m_start_up.fpp:298-390): the cell
    !! boundaries (cell widths from them) and the interior of every
    !! conservative variable, copied to the device (ghost cells zero until the
    !! first ghost fill).
    subroutine s_read_serial_data_files(q_cons_vf)
        type(scalar_field), dimension(sys_size), intent(INOUT) :: q_cons_vf
        character(len=path_len + 3*name_len) :: fp
        real(kind(0d0)), allocatable :: tmp(:, :, :), h(:, :, :)
        integer :: i
        call s_read_record(trim(f_t_step_dir(t_step_start))//'/x_cb.dat', x_cb(-1:m), m + 2)
        dx(0:m) = x_cb(0:m) - x_cb(-1:m - 1)
        if (n > 0) then
            call s_read_record(trim(f_t_step_dir(t_step_start))//'/y_cb.dat', y_cb(-1:n), n + 2)
            dy(0:n) = y_cb(0:n) - y_cb(-1:n - 1)
        end if
        if (p > 0) then
            call s_read_record(trim(f_t_step_dir(t_step_start))//'/z_cb.dat', z_cb(-1:p), p + 2)
            dz(0:p) = z_cb(0:p) - z_cb(-1:p - 1)
        end if
        allocate (tmp(0:m, 0:n, 0:p), h(-buff_size:m + buff_size, -bj:n + bj, -bk:p + bk))
        h = 0d0
        do i = 1, sys_size
            write (fp, '(A,I0,A)') trim(f_t_step_dir(t_step_start))//'/q_cons_vf', i, '.dat'
            call s_read_record(fp, tmp, (m + 1)*(n + 1)*(p + 1))
            h(0:m, 0:n, 0:p) = tmp
            call s_variable_to_device(h, q_cons_vf(i))
        end do
        deallocate (tmp, h)
    end subroutine s_read_serial_data_files

    !> s_read_parallel_data_files(q_cons_vf) (m_start_up.fpp:393-511) on one
    !! rank: restart_data/lustre_{x,y,z}_cb.dat hold the cell boundaries and
    !! lustre_<t_step_start>.dat the interior of every variable, raw doubles,
    !! variable i at byte offset (m+1)(n+1)(p+1)*8*(i-1) (what MPI-IO's file
    !! views lay down for a single rank; the reference's 2D has p+1 = 1).
    subroutine s_read_parallel_data_files(q_cons_vf)
        type(scalar_field), dimension(sys_size), intent(INOUT) :: q_cons_vf
        real(kind(0d0)), allocatable :: tmp(:, :, :), h(:, :, :)
        character(len=path_len + 2*name_len) :: fp
        integer(kind=8) :: nint
        integer :: i
        call s_read_raw(trim(case_dir)//'/restart_data/lustre_x_cb.dat', x_cb(-1:m), m + 2, 1_8)
        dx(0:m) = x_cb(0:m) - x_cb(-1:m - 1)
        if (n > 0) then
            call s_read_raw(trim(case_dir)//'/restart_data/lustre_y_cb.dat', y_cb(-1:n), n + 2, 1_8)
            dy(0:n) = y_cb(0:n) - y_cb(-1:n - 1)
        end if
        if (p > 0) then
            call s_read_raw(trim(case_dir)//'/restart_data/lustre_z_cb.dat', z_cb(-1:p), p + 2, 1_8)
            dz(0:p) = z_cb(0:p) - z_cb(-1:p - 1)
        end if
        nint = int(m + 1, 8)*int(n + 1, 8)*int(p + 1, 8)
        write (fp, '(A,I0,A)') trim(case_dir)//'/restart_data/lustre_', t_step_start, '.dat'
        allocate (tmp(0:m, 0:n, 0:p), h(-buff_size:m + buff_size, -bj:n + bj, -bk:p + bk))
        h = 0d0
        do i = 1, sys_size
            call s_read_raw(fp, tmp, int(nint), 1_8 + 8_8*nint*int(i - 1, 8))
            h(0:m, 0:n, 0:p) = tmp
            call s_variable_to_device(h, q_cons_vf(i))
        end do
        deallocate (tmp, h)
    end subroutine s_read_parallel_data_files

    !> count raw doubles from byte position pos (1-based) of a stream file.
    subroutine s_read_raw(path, a, count, pos)
        character(len=*), intent(in) :: path
        integer, intent(in) :: count
        real(kind(0d0)), intent(out) :: a(count)
        integer(kind=8), intent(in) :: pos
        logical :: file_exist
        integer :: ios
        inquire (FILE=trim(path), EXIST=file_exist)
        if (.not. file_exist) call s_abort('File '//trim(path)//' is missing. Exiting...')
        open (2, FILE=trim(path), FORM='unformatted', ACCESS='stream', ACTION='read', STATUS='old')
        read (2, POS=pos, iostat=ios) a
        close (2)
        if (ios /= 0) call s_abort('File '//trim(path)//' is too short. Exiting...')
    end subroutine s_read_raw

    !> The reader parallel_io selects (the reference's s_read_data_files
    !! procedure pointer, m_start_up.fpp:1010-1016).
    subroutine s_read_data_files(q_cons_vf)
        type(scalar_field), dimension(sys_size), intent(INOUT) :: q_cons_vf
        if (parallel_io) then
            call s_read_parallel_data_files(q_cons_vf)
        else
            call s_read_serial_data_files(q_cons_vf)
        end if
    end subroutine s_read_data_files

    !> s_populate_grid_variables_buffers (m_start_up.fpp:517-655) for one
    !! direction with s cells, boundary codes bcb/bce.
    subroutine s_populate_grid_buffers(s, bcb, bce, cb, ds)
        integer, intent(in) :: s, bcb, bce
        real(kind(0d0)), intent(inout) :: cb(-1 - buff_size:), ds(-buff_size:)
        integer :: i
        do i = 1, buff_size
            if (bcb <= -3) then
                ds(-i) = ds(0)
            elseif (bcb == -2) then
                ds(-i) = ds(i - 1)
            else
                ds(-i) = ds(s - (i - 1))
            end if
        end do
        do i = 1, buff_size
            cb(-1 - i) = cb(-i) - ds(-i)
        end do
        do i = 1, buff_size
            if (bce <= -3) then
                ds(s + i) = ds(s)
            elseif (bce == -2) then
                ds(s + i) = ds(s - (i - 1))
            else
                ds(s + i) = ds(i - 1)
            end if
        end do
        do i = 1, buff_size
            cb(s + i) = cb(s + (i - 1)) + ds(s + i)
        end do
    end subroutine s_populate_grid_buffers

    subroutine s_populate_grid_variables_buffers()
        call s_populate_grid_buffers(m, bc_x%beg, bc_x%end, x_cb, dx)
        if (n > 0) call s_populate_grid_buffers(n, bc_y%beg, bc_y%end, y_cb, dy)
        if (p > 0) call s_populate_grid_buffers(p, bc_z%beg, bc_z%end, z_cb, dz)
    end subroutine s_populate_grid_variables_buffers

end module m_start_up
