!----------------------------------------------------------------------------
! Module FORCE on the MI355X force engine.
!
! A drop-in for md_sim's own force module (src/forcem.f90, or src/forcem_back.f90 for the
! 3-D periodic form): the same module name, the same entry points force_init /
! force_zero / force_get, and the same module variables that the driver reads
! (src/move.f90:237-330, src/main.f90:37-62,132-134,326-328).  The bodies call the C ABI
! (include/mdforce.h) through ISO_C_BINDING instead of the CPU loops; everything else in
! the driver (Gear predictor-corrector, quaternions, SHAKE, measurement) is unchanged.
!
! Build: compile this file in place of forcem.f90, after the reference's nnlist and
! measure modules, with md_sim_amd/fortran/mdforce_c.f90, and link -lmdforce.
!
! Run-time choices (environment):
!   MDF_FORCE_VARIANT  slab | bulk   form of src/forcem.f90 (open z, slab correction, wall
!                                    tail) or src/forcem_back.f90 (3-D periodic pair loops);
!                                    default: bulk when the crystal has no sites, else slab
!   MDF_PRECISION      fp64 | fp32   FP32 pair/structure-factor arithmetic with FP64
!                                    accumulation; default fp64
!   MDF_DEVICE         HIP device index; default 0
!   MDF_MTS            k-space force by multiple time steps: the full sum every MDF_MTS-th
!                      step, the surrogate plus the last correction in between (no
!                      counterpart in the reference); unset or 1: full sum every step
!   MDF_MTS_MODEL      network | held; default network
!   MDF_MTS_TRAIN      full-sum training steps before the scheme starts; default 200
!   MDF_MTS_TOL        judged relative force RMS error that triggers the fallback; default 1e-3
!   MDF_MTS_LOG        file for one line of the scheme's counters per step
!   MDF_CHEMPOT_EW3DC  1 | 0: the Widom insertions' slab term with the step's coefficient, or
!                      none (the reference's chempot reads an unassigned local there);
!                      default 1
!
! The hot-loop observables ngc and efipos (src/forcem.f90:347-354,388-395,493-495,930-934)
! and, when tot.inp sets lchempot, the Widom insertions (src/forcem.f90:126,1226-1507) run
! on the device too; each step adds their counts into measure's arrays, and the insertions'
! generator state is mirrored into rndmod's seed.
!----------------------------------------------------------------------------
module force
    use nnlist
    use measure
    use rndmod, only: seed
    use mdforce_c, only: mdf_topology, mdf_params, mdf_outputs, mdf_create, mdf_zero, mdf_destroy, &
                         mdf_error_message, MDF_OK, MDF_FP64, MDF_FP32_ACC64, MDF_CONST_REFERENCE, &
                         MDF_ERFC_AS7126, MDF_ALPHA_REFERENCE, MDF_Z_OPEN, MDF_Z_PERIODIC_OPEN_LIST, &
                         mdf_mts_params, mdf_mts_stats, mdf_mts_enable, mdf_mts_get_stats, MDF_MTS_HELD, &
                         MDF_MTS_NETWORK, mdf_observables_enable, mdf_get_observables, mdf_chempot_params, &
                         mdf_chempot_enable, mdf_get_chempot, mdf_angle_histogram_enable, &
                         mdf_get_angle_histogram
    use, intrinsic :: iso_c_binding
    implicit none

    ! the driver-visible state of the reference's module (src/forcem.f90:16-49)
    real*8, dimension(:), allocatable :: fx, fy, fz
    real*8, dimension(:), allocatable :: ftx, fty, ftz, fcx, fcy, fcz, tx, ty, tz
    real*8, dimension(-1:1) :: fwsx, fwsy, fwsz
    real*8, dimension(:), allocatable :: elj, eqq
    real*8 :: alfa, spii, facq, press, uconf, uewa, urf, uself, ewqWW, ewqLL, ewqWL, ulr, vir
    real*8 :: totpres
    integer*8 :: kmax
    real*8 :: uavex, uavextot

    ! engine handle and the arrays the C side points into
    type(c_ptr), private :: sys = c_null_ptr
    integer, private :: mts_log = 0                ! unit of the scheme's log (0: none)
    type(mdf_topology), private :: top
    type(mdf_params), private :: par
    type(mdf_outputs), private :: out
    integer(c_int), allocatable, target, private :: c_itype(:), c_nsi(:), c_ijtype(:), c_vkind(:)
    integer(c_int), allocatable, target, private :: c_cona(:), c_conb(:), c_angi(:), c_angj(:), c_angk(:), c_angt(:)
    real(c_double), allocatable, target, private :: c_q(:), c_eps4(:), c_sig6(:), c_a(:), c_rhoi(:), c_c(:)
    real(c_double), allocatable, target, private :: c_conl(:), c_angkt(:), c_angth(:)
    real(c_double), allocatable, target, private :: o_fsite(:), o_fcom(:), o_tq(:), o_evdw(:), o_ecoul(:)
    ! one step's observables and insertions, fetched from the device
    integer(c_int64_t), allocatable, target, private :: o_ngc(:), o_inave(:), o_ngang(:)
    real(c_double), allocatable, target, private :: o_efipos(:), o_pave(:)
    integer(c_int32_t), private :: o_nbins = 0

    ! mdf_compute with the positions passed as arrays (no TARGET needed on the driver's arrays)
    interface
        integer(c_int) function mdf_compute_arr(s, x, y, z, cx, cy, cz, o) bind(C, name='mdf_compute')
            import :: c_int, c_ptr, c_double, mdf_outputs
            type(c_ptr), value :: s
            real(c_double), dimension(*), intent(in) :: x, y, z, cx, cy, cz
            type(mdf_outputs), intent(inout) :: o
        end function mdf_compute_arr
    end interface

    private :: env_value, build_topology

    contains

    ! value of an environment variable ('' when unset)
    function env_value(name) result(v)
        character(len=*), intent(in) :: name
        character(len=512) :: v
        integer :: st
        v = ''
        call get_environment_variable(name, v, status=st)
        if (st /= 0) v = ''
    end function env_value

    ! the reference's topology tables as the C ABI wants them (0-based sites, int32)
    subroutine build_topology
        integer*8 :: it, jt, ix, iy, is, n, k
        top%n_total = int(ntota, c_int)
        top%n_move = int(nmove, c_int)
        top%n_rigid_sites = int(nriga, c_int)
        top%n_liquid_sites = int(nliqa, c_int)
        top%n_rig = int(nrig, c_int)
        top%n_types = int(nit, c_int)
        top%n_ltype = int(nltype, c_int)
        allocate(c_itype(ntota), c_q(ntota), c_nsi(nit), c_ijtype(nit*nit), c_vkind(nijt))
        allocate(c_eps4(nijt), c_sig6(nijt), c_a(nijt), c_rhoi(nijt), c_c(nijt))
        c_itype = int(itype(1:ntota), c_int)
        c_q = qs(1:ntota)
        c_nsi = int(nsi(1:nit), c_int)
        do it = 1, nit
            do jt = 1, nit
                c_ijtype((it - 1)*nit + jt) = int(ijtype(it, jt), c_int)
            end do
        end do
        ! parameters after sysdef_init's transforms: 4 eps, sigma^6, 1/rho (src/sysdef.f90:372-374)
        c_vkind = int(vdwnum(1:nijt), c_int)
        c_eps4 = eps(1:nijt)
        c_sig6 = sig(1:nijt)
        c_a = aa(1:nijt)
        c_rhoi = rhoi(1:nijt)
        c_c = cc(1:nijt)
        top%site_type = c_loc(c_itype); top%charge = c_loc(c_q); top%nsi = c_loc(c_nsi)
        top%ijtype = c_loc(c_ijtype); top%vdw_kind = c_loc(c_vkind)
        top%vdw_eps4 = c_loc(c_eps4); top%vdw_sig6 = c_loc(c_sig6); top%vdw_a = c_loc(c_a)
        top%vdw_rhoi = c_loc(c_rhoi); top%vdw_c = c_loc(c_c)

        ! constraints, molecule by molecule as nnlist_constr walks them (src/nnlist.f90:189-203)
        n = 0
        do it = 1, ntype
            if (nconst(it) > 0) n = n + nt(it)*nconst(it)
        end do
        allocate(c_cona(max(n, 1_8)), c_conb(max(n, 1_8)), c_conl(max(n, 1_8)))
        k = 0
        ix = 0
        do it = 1, ntype
            if (nconst(it) <= 0) cycle
            do iy = 1, nt(it)
                ix = ix + 1
                do is = 1, nconst(it)
                    k = k + 1
                    c_cona(k) = int(constra(is, ix) - 1, c_int)
                    c_conb(k) = int(constrb(is, ix) - 1, c_int)
                    c_conl(k) = length(it, is)
                end do
            end do
        end do
        top%n_constraints = int(n, c_int)
        top%con_a = c_loc(c_cona); top%con_b = c_loc(c_conb); top%con_len = c_loc(c_conl)
        if (n == 0) then
            top%con_a = c_null_ptr; top%con_b = c_null_ptr; top%con_len = c_null_ptr
        end if

        ! angle terms, angle group by angle group as force_intra walks them (src/forcem.f90:571-576)
        n = 0
        do ix = 1, nang
            n = n + nangt(iatype(ix))
        end do
        allocate(c_angi(max(n, 1_8)), c_angj(max(n, 1_8)), c_angk(max(n, 1_8)))
        allocate(c_angkt(max(n, 1_8)), c_angth(max(n, 1_8)), c_angt(max(n, 1_8)))
        k = 0
        do ix = 1, nang
            it = iatype(ix)
            do is = 1, nangt(it)
                k = k + 1
                c_angi(k) = int(ang1(is, ix) - 1, c_int)
                c_angj(k) = int(ang2(is, ix) - 1, c_int)
                c_angk(k) = int(ang3(is, ix) - 1, c_int)
                c_angkt(k) = kt(it, is)
                c_angth(k) = theta(it, is)
                c_angt(k) = int(it, c_int)
            end do
        end do
        top%n_angles = int(n, c_int)
        top%ang_i = c_loc(c_angi); top%ang_j = c_loc(c_angj); top%ang_k = c_loc(c_angk)
        top%ang_kt = c_loc(c_angkt); top%ang_theta = c_loc(c_angth)
        if (n == 0) then
            top%ang_i = c_null_ptr; top%ang_j = c_null_ptr; top%ang_k = c_null_ptr
            top%ang_kt = c_null_ptr; top%ang_theta = c_null_ptr
        end if

        ! frozen crystal: pore width, cell, sites per crystal site type (src/forcem.f90:170-179)
        top%wall_lw = lw
        top%wall_cell = (/ cx, cy, cz /)
        top%wall_nss = int(nss(ntype, 1:2), c_int)
    end subroutine build_topology

    subroutine force_init
        integer(c_int) :: rc
        character(len=64) :: v
        logical :: bulk
        integer :: ios

        call build_topology
        v = env_value('MDF_FORCE_VARIANT')
        if (trim(v) == 'bulk') then
            bulk = .true.
        else if (trim(v) == 'slab') then
            bulk = .false.
        else
            bulk = (nwall == 0)
        end if
        par%box = (/ lx, ly, lz /)
        par%rcut = rcut
        par%skin = skin
        par%alpha = 0.0d0
        par%alpha_mode = MDF_ALPHA_REFERENCE
        par%kmax = int(kmax, c_int)
        if (bulk) then
            par%z_mode = MDF_Z_PERIODIC_OPEN_LIST
            par%ew3dc = 0
            par%wall_lr = 0
        else
            par%z_mode = MDF_Z_OPEN
            par%ew3dc = 1
            par%wall_lr = 1
        end if
        par%precision = MDF_FP64
        if (trim(env_value('MDF_PRECISION')) == 'fp32') par%precision = MDF_FP32_ACC64
        par%const_mode = MDF_CONST_REFERENCE
        par%erfc_mode = MDF_ERFC_AS7126
        par%max_neighbors = 0
        par%device = 0
        v = env_value('MDF_DEVICE')
        if (len_trim(v) > 0) read(v, *, iostat=ios) par%device
        par%virial = 0

        allocate(fx(nmove), fy(nmove), fz(nmove))
        allocate(ftx(nriga+1:nmove), fty(nriga+1:nmove), ftz(nriga+1:nmove))
        allocate(fcx(nrig), fcy(nrig), fcz(nrig), tx(nrig), ty(nrig), tz(nrig))
        allocate(elj(nijt), eqq(nijt))
        allocate(o_fsite(3*nmove), o_fcom(3*max(nrig, 1_8)), o_tq(3*max(nrig, 1_8)), o_evdw(nijt), o_ecoul(nijt))
        out%f_site = c_loc(o_fsite); out%f_com = c_loc(o_fcom); out%torque = c_loc(o_tq)
        out%e_vdw = c_loc(o_evdw); out%e_coul_real = c_loc(o_ecoul)

        rc = mdf_create(top, par, c_loc_or_null(x0), c_loc_or_null(y0), c_loc_or_null(z0), &
                        c_loc_or_null(cx0), c_loc_or_null(cy0), c_loc_or_null(cz0), sys)
        if (rc /= MDF_OK) then
            print *, 'force_init(): ', trim(mdf_error_message())
            stop 'force_init(): force engine could not be created'
        end if
        call mts_init
        call observables_init
        totpres = 0.0
        call force_zero
    end subroutine force_init

    ! the hot-loop observables with the driver's bin width, and the Widom insertions when
    ! tot.inp asks for them (its values as read: measure_init converts lwmin/lwmax later)
    subroutine observables_init
        type(mdf_chempot_params) :: cp
        integer(c_int) :: rc
        integer :: ios, ew
        character(len=512) :: v
        rc = mdf_observables_enable(sys, real(rdel, c_double))
        if (rc /= MDF_OK) then
            print *, 'force_init(): ', trim(mdf_error_message())
            stop 'force_init(): observables could not be set up'
        end if
        rc = mdf_get_observables(sys, c_null_ptr, c_null_ptr, o_nbins, 0_c_int32_t)
        allocate(o_ngc(max(1, o_nbins*int(nijt))), o_efipos(max(1, 2*o_nbins)))
        ! force_intra's angle histogram, sized as measure_init allocates ngang(nit-2, 0:201)
        ! (src/measure.f90:96-99)
        if (top%n_angles > 0) then
            rc = mdf_angle_histogram_enable(sys, c_loc(c_angt), int(nit - 2, c_int32_t))
            if (rc /= MDF_OK) then
                print *, 'force_init(): ', trim(mdf_error_message())
                stop 'force_init(): the angle histogram could not be set up'
            end if
            allocate(o_ngang(202*max(1, int(nit - 2))))
        end if
        if (.not. lchempot) return
        ew = 1
        v = env_value('MDF_CHEMPOT_EW3DC')
        if (len_trim(v) > 0) read(v, *, iostat=ios) ew
        cp%n_insert = int(ichmax, c_int)
        cp%n_bins = int(lchmax, c_int)
        cp%bin_mid = int(ilmid, c_int)
        cp%ew3dc = int(ew, c_int)
        cp%wall_min = lwmin
        cp%wall_max = lwmax
        cp%temperature = temper
        cp%seed = 0.0d0
        cp%max_partners = 0
        cp%pad = 0
        rc = mdf_chempot_enable(sys, cp)
        if (rc /= MDF_OK) then
            print *, 'force_init(): ', trim(mdf_error_message())
            stop 'force_init(): Widom insertions could not be set up'
        end if
        allocate(o_pave(max(1_8, lchmax)), o_inave(max(1_8, lchmax)))
    end subroutine observables_init

    ! add the step's ngc/efipos counts, and its insertions, into measure's arrays (the
    ! reference adds them inside its loops and chempot, src/forcem.f90:349-354,495,930-934,
    ! 1277,1495-1506)
    subroutine observables_step
        integer(c_int) :: rc
        integer*8 :: nb
        real(c_double) :: sd
        rc = mdf_get_observables(sys, c_loc(o_ngc), c_loc(o_efipos), o_nbins, 1_c_int32_t)
        if (rc /= MDF_OK) then
            print *, 'force_get(): ', trim(mdf_error_message())
            stop 'force_get(): observables could not be read'
        end if
        nb = min(int(o_nbins, 8), int(size(ngc, 2), 8))
        ngc(:, 0:nb-1) = ngc(:, 0:nb-1) + reshape(o_ngc(1:nijt*nb), (/ int(nijt), int(nb) /))
        nb = min(int(o_nbins, 8), int(size(efipos, 2), 8))
        efipos(:, 0:nb-1) = efipos(:, 0:nb-1) + reshape(o_efipos(1:2*nb), (/ 2, int(nb) /))
        ! the step's angle counts into ngang (src/forcem.f90:617-625)
        if (allocated(o_ngang)) then
            rc = mdf_get_angle_histogram(sys, c_loc(o_ngang), 1_c_int32_t)
            if (rc /= MDF_OK) then
                print *, 'force_get(): ', trim(mdf_error_message())
                stop 'force_get(): the angle histogram could not be read'
            end if
            if (allocated(ngang)) ngang(:, 0:201) = ngang(:, 0:201) + reshape(o_ngang, (/ int(nit - 2), 202 /))
        end if
        if (.not. lchempot) return
        rc = mdf_get_chempot(sys, c_loc(o_pave), c_loc(o_inave), sd, 1_c_int32_t)
        if (rc /= MDF_OK) then
            print *, 'force_get(): ', trim(mdf_error_message())
            stop 'force_get(): Widom insertions failed'
        end if
        pave(1:lchmax) = pave(1:lchmax) + o_pave(1:lchmax)
        inave(1:lchmax) = inave(1:lchmax) + o_inave(1:lchmax)
        pavesum(1:lchmax) = pave(1:lchmax)
        inavesum(1:lchmax) = inave(1:lchmax)
        seed = sd
    end subroutine observables_step

    ! multiple time steps for the k-space force, when MDF_MTS asks for them
    subroutine mts_init
        type(mdf_mts_params) :: mp
        character(len=512) :: v
        integer :: interval, ios
        integer(c_int) :: rc
        v = env_value('MDF_MTS')
        interval = 1
        if (len_trim(v) > 0) read(v, *, iostat=ios) interval
        if (interval <= 1) return
        mp%interval = int(interval, c_int)
        mp%model = MDF_MTS_NETWORK
        if (trim(env_value('MDF_MTS_MODEL')) == 'held') mp%model = MDF_MTS_HELD
        mp%train_frames = 200
        v = env_value('MDF_MTS_TRAIN')
        if (len_trim(v) > 0) read(v, *, iostat=ios) mp%train_frames
        mp%train_iters = 2
        mp%tol = 1.0d-3
        v = env_value('MDF_MTS_TOL')
        if (len_trim(v) > 0) read(v, *, iostat=ios) mp%tol
        mp%lr = 1.0e-3
        mp%seed = 1
        rc = mdf_mts_enable(sys, mp)
        if (rc /= MDF_OK) then
            print *, 'force_init(): ', trim(mdf_error_message())
            stop 'force_init(): multiple time steps could not be set up'
        end if
        v = env_value('MDF_MTS_LOG')
        if (len_trim(v) > 0) then
            mts_log = 97
            open(mts_log, file=trim(v), status='replace')
        end if
    end subroutine mts_init

    ! address of a real*8 module array for mdf_create, which copies what it needs
    function c_loc_or_null(a) result(p)
        real*8, dimension(:), intent(in), target :: a
        type(c_ptr) :: p
        p = c_null_ptr
        if (size(a) > 0) p = c_loc(a(1))
    end function c_loc_or_null

    subroutine force_zero
        uavextot = 0.0
    end subroutine force_zero

    subroutine force_get
        integer(c_int) :: rc
        integer*8 :: i, it, l, m

        rc = mdf_compute_arr(sys, x0, y0, z0, cx0, cy0, cz0, out)
        if (rc /= MDF_OK) then
            print *, 'force_get(): ', trim(mdf_error_message())
            stop 'force_get(): force evaluation failed'
        end if
        m = nmove
        fx = o_fsite(1:m)
        fy = o_fsite(m+1:2*m)
        fz = o_fsite(2*m+1:3*m)
        ! non-rigid site forces, rigid-body COM forces and torques (src/forcem.f90:993-1024);
        ! like the reference, fx(1:nrig) ends up holding the COM sums
        ftx(nriga+1:nmove) = fx(nriga+1:nmove)
        fty(nriga+1:nmove) = fy(nriga+1:nmove)
        ftz(nriga+1:nmove) = fz(nriga+1:nmove)
        if (nrig > 0) then
            fcx = o_fcom(1:nrig); fcy = o_fcom(nrig+1:2*nrig); fcz = o_fcom(2*nrig+1:3*nrig)
            tx = o_tq(1:nrig); ty = o_tq(nrig+1:2*nrig); tz = o_tq(2*nrig+1:3*nrig)
            fx(1:nrig) = fcx; fy(1:nrig) = fcy; fz(1:nrig) = fcz
        end if
        elj = o_evdw
        eqq = o_ecoul
        uewa = out%e_recip; ewqLL = out%e_recip_ll; ewqWL = out%e_recip_wl; ewqWW = out%e_recip_ww
        uself = out%e_self; urf = out%e_slab; ulr = out%e_wall_lr
        uconf = out%e_conf
        press = out%press_wall
        uavex = out%e_avex
        uavextot = uavextot + uavex
        totpres = totpres + press
        ! wall forces per side (-1, 0 = total, +1) as force_sum leaves them for measure_do
        fwsx = out%f_wall(1:3); fwsy = out%f_wall(4:6); fwsz = out%f_wall(7:9)
        fwxmeas = fwsx; fwymeas = fwsy; fwzmeas = fwsz
        ! z force profile of the mobile non-rigid sites (src/forcem.f90:1094-1101)
        if (me == 0 .and. nliqa > nriga) then
            do i = nriga + 1, nliqa
                it = itype(i)
                l = int((z0(i) + lwh)*rdeli + 0.5d0)
                nefipos(it - 2, l) = nefipos(it - 2, l) + 1
                tefipos(it - 2, l) = tefipos(it - 2, l) + ftz(i)
            end do
        end if
        call observables_step
        if (mts_log > 0) call mts_log_line
    end subroutine force_get

    ! one line per step: state, last step's kind, full sums, steps without, fallbacks,
    ! judgements, last and largest judged error
    subroutine mts_log_line
        type(mdf_mts_stats) :: st
        integer(c_int) :: rc
        rc = mdf_mts_get_stats(sys, st)
        if (rc /= MDF_OK) return
        write(mts_log, '(2i3,4i9,2es14.5)') st%state, st%last_exact, st%n_exact, st%n_surrogate, st%n_fallback, &
            st%n_judged, st%err_last, st%err_max
    end subroutine mts_log_line

end module force
