!----------------------------------------------------------------------------
! ISO_C_BINDING interface to the MI355X force engine (include/mdforce.h).
!
! The derived types mirror mdf_topology / mdf_params / mdf_outputs field by field;
! the interfaces bind the C entry points.  A Fortran caller keeps its arrays and
! passes c_loc() of their first elements (site arrays 1..ntota map to C 0..ntota-1).
!----------------------------------------------------------------------------
module mdforce_c
    use, intrinsic :: iso_c_binding
    implicit none

    integer(c_int), parameter :: MDF_OK = 0, MDF_ERR_ARG = 1, MDF_ERR_HIP = 2, &
        MDF_ERR_LIST_OVERFLOW = 3, MDF_ERR_VDW_TYPE = 4, MDF_ERR_KSET = 5, &
        MDF_ERR_NONFINITE = 6, MDF_ERR_NO_DEVICE = 7
    integer(c_int), parameter :: MDF_FP64 = 0, MDF_FP32_ACC64 = 1
    integer(c_int), parameter :: MDF_CONST_REFERENCE = 0, MDF_CONST_EXACT = 1
    integer(c_int), parameter :: MDF_ERFC_AS7126 = 0, MDF_ERFC_ACCURATE = 1
    integer(c_int), parameter :: MDF_ALPHA_REFERENCE = 0, MDF_ALPHA_GIVEN = 1
    integer(c_int), parameter :: MDF_Z_OPEN = 0, MDF_Z_PERIODIC = 1, MDF_Z_PERIODIC_OPEN_LIST = 2

    type, bind(C) :: mdf_topology
        integer(c_int) :: n_total, n_move, n_rigid_sites, n_liquid_sites, n_rig, n_types, n_ltype
        type(c_ptr) :: site_type, charge, nsi, ijtype, vdw_kind
        type(c_ptr) :: vdw_eps4, vdw_sig6, vdw_a, vdw_rhoi, vdw_c
        integer(c_int) :: n_constraints
        type(c_ptr) :: con_a, con_b, con_len
        integer(c_int) :: n_angles
        type(c_ptr) :: ang_i, ang_j, ang_k, ang_kt, ang_theta
        real(c_double) :: wall_lw
        real(c_double) :: wall_cell(3)
        integer(c_int) :: wall_nss(2)
    end type mdf_topology

    type, bind(C) :: mdf_params
        real(c_double) :: box(3)
        real(c_double) :: rcut, skin, alpha
        integer(c_int) :: alpha_mode, kmax, z_mode, ew3dc, wall_lr, precision, const_mode, &
                          erfc_mode, max_neighbors, device, virial
    end type mdf_params

    integer(c_int), parameter :: MDF_MTS_HELD = 0, MDF_MTS_NETWORK = 1

    type, bind(C) :: mdf_mts_params
        integer(c_int) :: interval, model, train_frames, train_iters
        real(c_double) :: tol
        real(c_float) :: lr
        integer(c_int32_t) :: seed
    end type mdf_mts_params

    type, bind(C) :: mdf_chempot_params
        integer(c_int) :: n_insert, n_bins, bin_mid, ew3dc
        real(c_double) :: wall_min, wall_max, temperature, seed
        integer(c_int) :: max_partners, pad
    end type mdf_chempot_params

    type, bind(C) :: mdf_mts_stats
        integer(c_int) :: state, last_exact
        integer(c_int64_t) :: n_exact, n_surrogate, n_fallback, n_judged, adam_steps
        real(c_double) :: err_last, err_max, err_mean, loss_last
    end type mdf_mts_stats

    type, bind(C) :: mdf_outputs
        type(c_ptr) :: f_site, f_com, torque, e_vdw, e_coul_real
        real(c_double) :: e_recip, e_recip_ll, e_recip_wl, e_recip_ww, e_self, e_slab, &
                          e_wall_lr, e_conf, e_avex, press_wall
        real(c_double) :: f_wall(9), virial(9)
        integer(c_int) :: list_rebuilt, n_pairs_listed, n_wall_pairs_listed, n_kvec, n_list_builds
    end type mdf_outputs

    interface
        integer(c_int) function mdf_create(top, par, x, y, z, cx, cy, cz, sys) bind(C, name='mdf_create')
            import :: c_int, c_ptr, mdf_topology, mdf_params
            type(mdf_topology), intent(in) :: top
            type(mdf_params), intent(in) :: par
            type(c_ptr), value :: x, y, z, cx, cy, cz
            type(c_ptr), intent(out) :: sys
        end function mdf_create

        integer(c_int) function mdf_compute(sys, x, y, z, cx, cy, cz, out) bind(C, name='mdf_compute')
            import :: c_int, c_ptr, mdf_outputs
            type(c_ptr), value :: sys, x, y, z, cx, cy, cz
            type(mdf_outputs), intent(inout) :: out
        end function mdf_compute

        integer(c_int) function mdf_zero(sys) bind(C, name='mdf_zero')
            import :: c_int, c_ptr
            type(c_ptr), value :: sys
        end function mdf_zero

        integer(c_int) function mdf_autotune(sys, reps, chosen) bind(C, name='mdf_autotune')
            import :: c_int, c_ptr
            type(c_ptr), value :: sys
            integer(c_int), value :: reps
            integer(c_int), intent(out) :: chosen(4)
        end function mdf_autotune

        integer(c_int) function mdf_mts_enable(sys, p) bind(C, name='mdf_mts_enable')
            import :: c_int, c_ptr, mdf_mts_params
            type(c_ptr), value :: sys
            type(mdf_mts_params), intent(in) :: p
        end function mdf_mts_enable

        integer(c_int) function mdf_mts_get_stats(sys, st) bind(C, name='mdf_mts_get_stats')
            import :: c_int, c_ptr, mdf_mts_stats
            type(c_ptr), value :: sys
            type(mdf_mts_stats), intent(out) :: st
        end function mdf_mts_get_stats

        integer(c_int) function mdf_observables_enable(sys, rdel) bind(C, name='mdf_observables_enable')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: sys
            real(c_double), value :: rdel
        end function mdf_observables_enable

        integer(c_int) function mdf_get_observables(sys, ngc, efipos, nbins, reset) &
                bind(C, name='mdf_get_observables')
            import :: c_int, c_ptr, c_int32_t
            type(c_ptr), value :: sys, ngc, efipos
            integer(c_int32_t), intent(out) :: nbins
            integer(c_int32_t), value :: reset
        end function mdf_get_observables

        integer(c_int) function mdf_angle_histogram_enable(sys, ang_type, n_atypes) &
                bind(C, name='mdf_angle_histogram_enable')
            import :: c_int, c_ptr, c_int32_t
            type(c_ptr), value :: sys, ang_type
            integer(c_int32_t), value :: n_atypes
        end function mdf_angle_histogram_enable

        integer(c_int) function mdf_get_angle_histogram(sys, ngang, reset) bind(C, name='mdf_get_angle_histogram')
            import :: c_int, c_ptr, c_int32_t
            type(c_ptr), value :: sys, ngang
            integer(c_int32_t), value :: reset
        end function mdf_get_angle_histogram

        integer(c_int) function mdf_chempot_enable(sys, p) bind(C, name='mdf_chempot_enable')
            import :: c_int, c_ptr, mdf_chempot_params
            type(c_ptr), value :: sys
            type(mdf_chempot_params), intent(in) :: p
        end function mdf_chempot_enable

        integer(c_int) function mdf_get_chempot(sys, pave, inave, seed, reset) bind(C, name='mdf_get_chempot')
            import :: c_int, c_ptr, c_double, c_int32_t
            type(c_ptr), value :: sys, pave, inave
            real(c_double), intent(out) :: seed
            integer(c_int32_t), value :: reset
        end function mdf_get_chempot

        subroutine mdf_destroy(sys) bind(C, name='mdf_destroy')
            import :: c_ptr
            type(c_ptr), value :: sys
        end subroutine mdf_destroy

        type(c_ptr) function mdf_last_error() bind(C, name='mdf_last_error')
            import :: c_ptr
        end function mdf_last_error
    end interface

contains

    ! message of the last failing call as a Fortran string
    function mdf_error_message() result(msg)
        character(len=512) :: msg
        type(c_ptr) :: p
        character(kind=c_char), pointer :: s(:)
        integer :: i
        msg = ''
        p = mdf_last_error()
        if (.not. c_associated(p)) return
        call c_f_pointer(p, s, [512])
        do i = 1, 512
            if (s(i) == c_null_char) exit
            msg(i:i) = s(i)
        end do
    end function mdf_error_message

end module mdforce_c
