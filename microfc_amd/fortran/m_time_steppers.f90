!> Drop-in for src/simulation/m_time_steppers.fpp with the reference's
!! declarations (:30-39): q_cons_ts(1:num_ts)%vf(1:sys_size), q_prim_vf and
!! rhs_vf are the reference's field types, their sf pointing at device
!! memory.  One library call advances a whole TVD-RK step, the stage updates
!! fused into the last directional sweep.
module m_time_steppers

    use iso_c_binding
    use m_mfc_abi
    use m_derived_types
    use m_global_parameters
    use m_data_output

    implicit none

    type(vector_field), allocatable, dimension(:) :: q_cons_ts
    type(scalar_field), allocatable, dimension(:) :: q_prim_vf
    type(scalar_field), allocatable, dimension(:) :: rhs_vf

    integer, private :: num_ts
    type(c_ptr), private, allocatable, dimension(:) :: cons_base
    type(c_ptr), private :: prim_base = c_null_ptr, rhs_base = c_null_ptr

contains

    !> s_initialize_time_steppers_module (m_time_steppers.fpp:53-125): the
    !! time-stage states, the primitive field and rhs on the device, and the
    !! run-time information file when run_time_info is set.
    subroutine s_initialize_time_steppers_module()
        integer :: i
        num_ts = 2
        if (time_stepper == 1) num_ts = 1
        allocate (q_cons_ts(1:num_ts), cons_base(1:num_ts))
        do i = 1, num_ts
            allocate (q_cons_ts(i)%vf(1:sys_size))
            call s_allocate_field(q_cons_ts(i)%vf, .true., cons_base(i))
        end do
        allocate (q_prim_vf(1:sys_size), rhs_vf(1:sys_size))
        call s_allocate_field(q_prim_vf, .true., prim_base)
        call s_allocate_field(rhs_vf, .false., rhs_base)
        if (run_time_info) call s_open_run_time_information_file()
    end subroutine s_initialize_time_steppers_module

    !> s_1st_order_tvd_rk(t_step, time_avg) (m_time_steppers.fpp:129-193).
    subroutine s_1st_order_tvd_rk(t_step, time_avg)
        integer, intent(IN) :: t_step
        real(kind(0d0)), intent(INOUT) :: time_avg
        call s_tvd_rk(t_step, time_avg)
    end subroutine s_1st_order_tvd_rk

    !> s_2nd_order_tvd_rk(t_step, time_avg) (m_time_steppers.fpp:197-267).
    subroutine s_2nd_order_tvd_rk(t_step, time_avg)
        integer, intent(IN) :: t_step
        real(kind(0d0)), intent(INOUT) :: time_avg
        call s_tvd_rk(t_step, time_avg)
    end subroutine s_2nd_order_tvd_rk

    !> s_3rd_order_tvd_rk(t_step, time_avg) (m_time_steppers.fpp:271-362).
    subroutine s_3rd_order_tvd_rk(t_step, time_avg)
        integer, intent(IN) :: t_step
        real(kind(0d0)), intent(INOUT) :: time_avg
        call s_tvd_rk(t_step, time_avg)
    end subroutine s_3rd_order_tvd_rk

    !> One step of the configured scheme: the run-time information of the
    !! first stage's primitive field (the reference writes it right after the
    !! first s_compute_rhs), nothing more when t_step == t_step_stop (:296),
    !! and the mean wall time per step from step 4 on (:281, :352-358).
    subroutine s_tvd_rk(t_step, time_avg)
        integer, intent(IN) :: t_step
        real(kind(0d0)), intent(INOUT) :: time_avg
        type(c_ptr), dimension(sys_size), target :: p1, p2, pp, pr
        type(c_ptr) :: prim, q2
        integer(kind=8) :: c0, c1, rate
        real(kind(0d0)) :: sec
        call system_clock(c0, rate)
        call s_field_pointers(q_cons_ts(1)%vf, p1)
        call s_field_pointers(rhs_vf, pr)
        prim = c_null_ptr
        if (run_time_info) then
            call s_field_pointers(q_prim_vf, pp)
            prim = c_loc(pp)
        end if
        q2 = c_null_ptr
        if (num_ts > 1) then
            call s_field_pointers(q_cons_ts(2)%vf, p2)
            q2 = c_loc(p2)
        end if
        call s_mfc_check(mfc_tvd_rk_step(int(t_step, c_int), dt, c_loc(p1), q2, prim, c_loc(pr)), 'mfc_tvd_rk_step')
        if (run_time_info) call s_write_run_time_information(q_prim_vf, t_step)
        call s_mfc_check(mfc_sync(), 'mfc_sync')
        call system_clock(c1)
        sec = real(c1 - c0, kind(0d0))/real(rate, kind(0d0))
        if (t_step >= 4) then
            time_avg = (abs(sec) + (t_step - 4)*time_avg)/(t_step - 3)
        else
            time_avg = 0d0
        end if
    end subroutine s_tvd_rk

    !> s_finalize_time_steppers_module (m_time_steppers.fpp:430-455).
    subroutine s_finalize_time_steppers_module()
        integer :: i
        if (run_time_info) call s_close_run_time_information_file()
        do i = 1, num_ts
            call s_free_field(q_cons_ts(i)%vf, cons_base(i))
            deallocate (q_cons_ts(i)%vf)
        end do
        call s_free_field(q_prim_vf, prim_base)
        call s_free_field(rhs_vf, rhs_base)
        deallocate (q_cons_ts, cons_base, q_prim_vf, rhs_vf)
    end subroutine s_finalize_time_steppers_module

end module m_time_steppers
