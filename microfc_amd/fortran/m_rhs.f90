!> Drop-in for src/simulation/m_rhs.fpp: the same module procedures, with the
!! bodies replaced by calls into the HIP library through ISO_C_BINDING.
module m_rhs

    use iso_c_binding
    use m_mfc_abi
    use m_derived_types
    use m_global_parameters

    implicit none

    private; public :: s_initialize_rhs_module, s_initialize_weno_module, s_compute_rhs, s_finalize_rhs_module

contains

    !> s_initialize_rhs_module (m_rhs.fpp:123): sets the library up for the
    !! case (the module chain of p_main.fpp:125-183, whose input checks run
    !! before any device call) and loads the learned weights when asked.
    subroutine s_initialize_rhs_module()
        call s_mfc_check(mfc_rhs_init(f_mfc_params()), 'mfc_rhs_init')
        if (bulk_precision == mfc_prec_fp32_learned) call s_load_mlp_weights()
    end subroutine s_initialize_rhs_module

    !> s_initialize_weno_module (m_weno.fpp:103-159): the WENO coefficients
    !! of each direction (m_weno.fpp:168-363) from the ghosted grid, which
    !! p_main reads before this call, as the reference does.
    subroutine s_initialize_weno_module()
        call s_mfc_check(mfc_rhs_set_grid(0_c_int, x_cb, dx), 'mfc_rhs_set_grid(x)')
        if (n > 0) call s_mfc_check(mfc_rhs_set_grid(1_c_int, y_cb, dy), 'mfc_rhs_set_grid(y)')
        if (p > 0) call s_mfc_check(mfc_rhs_set_grid(2_c_int, z_cb, dz), 'mfc_rhs_set_grid(z)')
    end subroutine s_initialize_weno_module

    !> Learned WENO5 weights (no reference counterpart): loads the network
    !! named by mlp_weights, mfc_mlp_nparam little-endian FP32 values in the
    !! parameter layout of include/mfc_rhs.h, into the library.
    subroutine s_load_mlp_weights()
        real(c_float) :: params(mfc_mlp_nparam)
        logical :: file_exist
        integer :: ios
        inquire (FILE=trim(mlp_weights), EXIST=file_exist)
        if (len_trim(mlp_weights) == 0 .or. .not. file_exist) then
            print '(A)', 'Error: bulk_precision = 2 needs mlp_weights, the file of trained MLP parameters'
            stop 1
        end if
        open (3, FILE=trim(mlp_weights), FORM='unformatted', ACCESS='stream', ACTION='read', STATUS='old')
        read (3, iostat=ios) params; close (3)
        if (ios /= 0) then
            print '(A,I0,A)', 'Error: mlp_weights must hold ', mfc_mlp_nparam, ' FP32 values'
            stop 1
        end if
        call s_mfc_mlp_check(mfc_mlp_init(params), 'mfc_mlp_init')
    end subroutine s_load_mlp_weights

    !> s_compute_rhs(q_cons_vf, q_prim_vf, rhs_vf, t_step) (m_rhs.fpp:405-679),
    !! with the reference's declarations (:405-410): the fields are the
    !! reference's scalar_field arrays, their sf pointing at device memory.
    !! q_prim_vf is written when run_time_info is set, as in the reference
    !! (:659-675); the ghost cells of q_cons_vf are filled in place.  The work
    !! is queued on the library's stream, which every later library call
    !! (including the host copies) follows.
    subroutine s_compute_rhs(q_cons_vf, q_prim_vf, rhs_vf, t_step)
        type(scalar_field), dimension(sys_size), intent(INOUT) :: q_cons_vf, q_prim_vf, rhs_vf
        integer, intent(IN) :: t_step
        type(c_ptr), dimension(sys_size), target :: pc, pp, pr
        type(c_ptr) :: prim
        call s_field_pointers(q_cons_vf, pc)
        call s_field_pointers(rhs_vf, pr)
        prim = c_null_ptr
        if (run_time_info) then
            call s_field_pointers(q_prim_vf, pp)
            prim = c_loc(pp)
        end if
        call s_mfc_check(mfc_compute_rhs(c_loc(pc), prim, c_loc(pr), int(t_step, c_int)), 'mfc_compute_rhs')
    end subroutine s_compute_rhs

    subroutine s_finalize_rhs_module()
        call s_mfc_check(mfc_rhs_finalize(), 'mfc_rhs_finalize')
    end subroutine s_finalize_rhs_module

end module m_rhs
