!> Drop-in for src/simulation/m_time_steppers.fpp: the state q_cons_ts(1:2),
!! the primitive field and rhs_vf live on the device; one call of the
!! library advances a whole TVD-RK step with the stage updates fused into the
!! last directional sweep.
module m_time_steppers

    use iso_c_binding
    use m_mfc_abi
    use m_global_parameters

    implicit none

    type(device_field), dimension(2), target :: q_cons_ts
    type(device_field), target :: q_prim_vf
    type(device_field), target :: rhs_vf

contains

    !> s_initialize_time_steppers_module (m_time_steppers.fpp:53-125).
    subroutine s_initialize_time_steppers_module()
        call s_allocate_device_field(q_cons_ts(1), .true.)
        if (time_stepper > 1) call s_allocate_device_field(q_cons_ts(2), .true.)
        call s_allocate_device_field(q_prim_vf, .true.)
        call s_allocate_device_field(rhs_vf, .false.)
    end subroutine s_initialize_time_steppers_module

    !> s_{1st,2nd,3rd}_order_tvd_rk(t_step, time_avg)
    !! (m_time_steppers.fpp:129-362), including the run-time information call
    !! on the first stage's primitive field and the mean wall time per step
    !! from step 4 on (:281, :352-358).  Returns without advancing when
    !! t_step == t_step_stop.
    subroutine s_tvd_rk(t_step, time_avg, icfl)
        integer, intent(in) :: t_step
        real(kind(0d0)), intent(inout) :: time_avg
        real(kind(0d0)), intent(out) :: icfl
        type(c_ptr) :: prim, q2
        integer(c_int) :: rc
        integer(kind=8) :: c0, c1, rate
        real(kind(0d0)) :: sec
        call system_clock(c0, rate)
        prim = c_null_ptr
        if (run_time_info) prim = c_loc(q_prim_vf%v)
        q2 = c_null_ptr
        if (time_stepper > 1) q2 = c_loc(q_cons_ts(2)%v)
        rc = mfc_tvd_rk_step(int(t_step, c_int), dt, c_loc(q_cons_ts(1)%v), q2, prim, c_loc(rhs_vf%v))
        call s_mfc_check(rc, 'mfc_tvd_rk_step')
        icfl = -1d0
        if (run_time_info) call s_mfc_check(mfc_icfl_max(c_loc(q_prim_vf%v), dt, icfl), 'mfc_icfl_max')
        call s_mfc_check(mfc_sync(), 'mfc_sync')
        call system_clock(c1)
        sec = real(c1 - c0, kind(0d0))/real(rate, kind(0d0))
        if (t_step >= 4) then
            time_avg = (abs(sec) + (t_step - 4)*time_avg)/(t_step - 3)
        else
            time_avg = 0d0
        end if
    end subroutine s_tvd_rk

    subroutine s_finalize_time_steppers_module()
        call s_free_device_field(q_cons_ts(1))
        call s_free_device_field(q_cons_ts(2))
        call s_free_device_field(q_prim_vf)
        call s_free_device_field(rhs_vf)
    end subroutine s_finalize_time_steppers_module

end module m_time_steppers
