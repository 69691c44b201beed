!> Drop-in for src/simulation/m_rhs.fpp: the same module procedures, with the
!! bodies replaced by calls into the HIP library through ISO_C_BINDING.
module m_rhs

    use iso_c_binding
    use m_mfc_abi
    use m_global_parameters

    implicit none

    private; public :: s_initialize_rhs_module, s_compute_rhs, s_finalize_rhs_module

contains

    !> s_initialize_rhs_module (m_rhs.fpp:123): sets the library up for the
    !! case (the module chain of p_main.fpp:125-183 and the WENO
    !! coefficients of m_weno.fpp:103-363, which need the ghosted grid).
    subroutine s_initialize_rhs_module()
        call s_mfc_check(mfc_rhs_init(f_mfc_params()), 'mfc_rhs_init')
        call s_mfc_check(mfc_rhs_set_grid(0_c_int, x_cb, dx), 'mfc_rhs_set_grid(x)')
        if (n > 0) call s_mfc_check(mfc_rhs_set_grid(1_c_int, y_cb, dy), 'mfc_rhs_set_grid(y)')
        if (p > 0) call s_mfc_check(mfc_rhs_set_grid(2_c_int, z_cb, dz), 'mfc_rhs_set_grid(z)')
    end subroutine s_initialize_rhs_module

    !> s_compute_rhs(q_cons_vf, q_prim_vf, rhs_vf, t_step) (m_rhs.fpp:405-679).
    !! The fields are device-resident; q_prim_vf is written when
    !! run_time_info is set, as in the reference (:659-675).
    subroutine s_compute_rhs(q_cons_vf, q_prim_vf, rhs_vf, t_step)
        type(device_field), intent(inout), target :: q_cons_vf, q_prim_vf, rhs_vf
        integer, intent(in) :: t_step
        type(c_ptr) :: prim
        prim = c_null_ptr
        if (run_time_info) prim = c_loc(q_prim_vf%v)
        call s_mfc_check(mfc_compute_rhs(c_loc(q_cons_vf%v), prim, c_loc(rhs_vf%v), int(t_step, c_int)), &
                         'mfc_compute_rhs')
    end subroutine s_compute_rhs

    subroutine s_finalize_rhs_module()
        call s_mfc_check(mfc_rhs_finalize(), 'mfc_rhs_finalize')
    end subroutine s_finalize_rhs_module

end module m_rhs
