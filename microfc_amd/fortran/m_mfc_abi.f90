!> ISO_C_BINDING interfaces to the HIP RHS library (include/mfc_rhs.h).
!! Every function returns 0 on success, a negative code on failure
!! (mfc_last_error then holds the message) and 1 for the t_step_stop early
!! return.  Field arguments are c_loc() of an array of sys_size device
!! pointers, one per variable.
module m_mfc_abi

    use iso_c_binding

    implicit none

    integer, parameter :: mfc_num_fluids_max = 10
    integer, parameter :: mfc_prec_fp64 = 0, mfc_prec_fp32_bulk = 1, mfc_prec_fp32_learned = 2, &
                          mfc_prec_fp64_fast = 3
    integer, parameter :: mfc_mlp_nparam = 2128

    type, bind(C) :: mfc_params
        integer(c_int) :: m, n, p
        integer(c_int) :: num_fluids
        integer(c_int) :: weno_order
        integer(c_int) :: time_stepper
        integer(c_int) :: t_step_stop
        integer(c_int) :: bc_x_beg, bc_x_end, bc_y_beg, bc_y_end, bc_z_beg, bc_z_end
        real(c_double) :: weno_eps
        real(c_double) :: gammas(mfc_num_fluids_max)
        real(c_double) :: pi_infs(mfc_num_fluids_max)
        integer(c_int) :: precision
        real(c_double) :: Re(2, mfc_num_fluids_max)
        integer(c_int) :: weno_Re_flux
    end type mfc_params

    interface

        integer(c_int) function mfc_rhs_init(prm) bind(C, name='mfc_rhs_init')
            import :: c_int, mfc_params
            type(mfc_params), intent(in) :: prm
        end function mfc_rhs_init

        integer(c_int) function mfc_rhs_set_grid(dir, s_cb, ds) bind(C, name='mfc_rhs_set_grid')
            import :: c_int, c_double
            integer(c_int), value :: dir
            real(c_double), intent(in) :: s_cb(*), ds(*)
        end function mfc_rhs_set_grid

        integer(c_int) function mfc_compute_rhs(q_cons, q_prim, rhs, t_step) bind(C, name='mfc_compute_rhs')
            import :: c_int, c_ptr
            type(c_ptr), value :: q_cons, q_prim, rhs
            integer(c_int), value :: t_step
        end function mfc_compute_rhs

        integer(c_int) function mfc_tvd_rk_step(t_step, dt, q1, q2, q_prim, rhs) bind(C, name='mfc_tvd_rk_step')
            import :: c_int, c_double, c_ptr
            integer(c_int), value :: t_step
            real(c_double), value :: dt
            type(c_ptr), value :: q1, q2, q_prim, rhs
        end function mfc_tvd_rk_step

        integer(c_int) function mfc_cons_to_prim(q_cons, q_prim) bind(C, name='mfc_cons_to_prim')
            import :: c_int, c_ptr
            type(c_ptr), value :: q_cons, q_prim
        end function mfc_cons_to_prim

        integer(c_int) function mfc_icfl_max(q_prim, dt, icfl) bind(C, name='mfc_icfl_max')
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: q_prim
            real(c_double), value :: dt
            real(c_double), intent(out) :: icfl
        end function mfc_icfl_max

        integer(c_int) function mfc_stability_criteria(q_prim, dt, crit) bind(C, name='mfc_stability_criteria')
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: q_prim
            real(c_double), value :: dt
            real(c_double), intent(out) :: crit(3)
        end function mfc_stability_criteria

        integer(c_int) function mfc_cfl_dt(q_cons, cfl, dt) bind(C, name='mfc_cfl_dt')
            import :: c_int, c_double, c_ptr
            type(c_ptr), value :: q_cons
            real(c_double), value :: cfl
            real(c_double), intent(out) :: dt
        end function mfc_cfl_dt

        integer(c_int) function mfc_device_alloc(bytes, ptr) bind(C, name='mfc_device_alloc')
            import :: c_int, c_size_t, c_ptr
            integer(c_size_t), value :: bytes
            type(c_ptr), intent(out) :: ptr
        end function mfc_device_alloc

        integer(c_int) function mfc_device_free(ptr) bind(C, name='mfc_device_free')
            import :: c_int, c_ptr
            type(c_ptr), value :: ptr
        end function mfc_device_free

        integer(c_int) function mfc_memcpy_h2d(dst, src, bytes) bind(C, name='mfc_memcpy_h2d')
            import :: c_int, c_ptr, c_size_t
            type(c_ptr), value :: dst, src
            integer(c_size_t), value :: bytes
        end function mfc_memcpy_h2d

        integer(c_int) function mfc_memcpy_d2h(dst, src, bytes) bind(C, name='mfc_memcpy_d2h')
            import :: c_int, c_ptr, c_size_t
            type(c_ptr), value :: dst, src
            integer(c_size_t), value :: bytes
        end function mfc_memcpy_d2h

        integer(c_int) function mfc_sync() bind(C, name='mfc_sync')
            import :: c_int
        end function mfc_sync

        integer(c_int) function mfc_rhs_finalize() bind(C, name='mfc_rhs_finalize')
            import :: c_int
        end function mfc_rhs_finalize

        type(c_ptr) function mfc_last_error() bind(C, name='mfc_last_error')
            import :: c_ptr
        end function mfc_last_error

        integer(c_int) function mfc_mlp_init(params) bind(C, name='mfc_mlp_init')
            import :: c_int, c_float
            real(c_float), intent(in) :: params(*)
        end function mfc_mlp_init

        type(c_ptr) function mfc_mlp_last_error() bind(C, name='mfc_mlp_last_error')
            import :: c_ptr
        end function mfc_mlp_last_error

    end interface

contains

    !> Stops with the library's message when rc < 0 (the reference's
    !! s_mpi_abort convention: print and stop).
    subroutine s_mfc_check(rc, what)
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: what
        if (rc >= 0) return
        call s_mfc_stop(rc, what, mfc_last_error())
    end subroutine s_mfc_check

    !> The same for the learned-weight entry points (mfc_mlp_*).
    subroutine s_mfc_mlp_check(rc, what)
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: what
        if (rc >= 0) return
        call s_mfc_stop(rc, what, mfc_mlp_last_error())
    end subroutine s_mfc_mlp_check

    subroutine s_mfc_stop(rc, what, cmsg)
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: what
        type(c_ptr), intent(in) :: cmsg
        character(kind=c_char), pointer :: msg(:)
        integer :: i
        character(len=512) :: text
        call c_f_pointer(cmsg, msg, [512])
        text = ''
        do i = 1, 512
            if (msg(i) == c_null_char) exit
            text(i:i) = msg(i)
        end do
        print '(A,A,A,I0,A,A)', 'Error: ', what, ' failed (', rc, '): ', trim(text)
        stop 1
    end subroutine s_mfc_stop

end module m_mfc_abi
