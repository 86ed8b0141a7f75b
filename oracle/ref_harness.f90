!----------------------------------------------------------------------------
! Golden-vector harness for the reference force path (test infrastructure).
!
! Replaces the reference's program unit (src/main.f90:19-108): it performs the
! same set-up as `input`/`init` (src/main.f90:113-315), sets the
! site offsets the way move_predict does (src/move.f90:157-161), runs the pieces
! of one force_get (src/forcem.f90:93-129) and dumps inputs and outputs to a
! binary stream file.  Site forces are dumped before force_sum, which overwrites
! fx(1:nrig) with COM sums (src/forcem.f90:1019-1024).  On several MPI ranks it only times
! repeated force evaluations (the reference's own decomposition of the force path).
!
! Dump record: name (16 chars), kind (int64: 1 real*8, 2 int64), count (int64),
! data.  Built only in the development container, from the reference sources.
!----------------------------------------------------------------------------
program ref_harness
    use move
    use measure
    use comm
    use rndmod
    implicit none
    include 'mpif.h'
    integer :: ierr, u
    integer*8 :: i, ix, n, kl, km, kn, kcnt, kmn, knn, nsel, nrep, irep, c0, c1, crate
    character*32 :: arg2
    character*40 :: base
    character*256 :: outname
    real*8, allocatable :: f1x(:), f1y(:), f1z(:), f2x(:), f2y(:), f2z(:)
    real*8, allocatable :: kak(:), kre(:), kim(:), kvec(:)
    integer*8, allocatable :: ibuf(:)
    real*8, allocatable :: cbuf(:)
    character*256 :: chemenv
    integer*8 :: nchcall, ic
    logical :: chzero
    real*8, allocatable :: chp(:), chs(:), chpar(:)
    integer*8, allocatable :: chi(:)

    nproc = 1
    me = 0
    call mpi_init(ierr)
    call mpi_comm_size(mpi_comm_world, nproc, ierr)
    call mpi_comm_rank(mpi_comm_world, me, ierr)

    call get_command_argument(1, outname)
    ! optional second argument: number of extra timed force evaluations
    nrep = 0
    call get_command_argument(2, arg2)
    if (len_trim(arg2) > 0) read(arg2, *) nrep
    ! several ranks: the reference's own decomposition of the force path (leads and
    ! k-vectors over ranks, force_sum's all-reduce), for timing only - the dump is one rank's
    if (nproc /= 1 .and. nrep <= 0) stop 'ref_harness: several ranks only time (give nrep)'
    open(7, file='harness.inp', status='old')
    read(7, *) base
    read(7, *) dt, rcut, skin, rdel, kmax
    close(7)

    ! the tot.inp switches the force path reads; all optional features off
    temper = -1.0
    timeq = 0.0
    timrun = 0.0
    timprn = dt
    l2d = .false.
    lhb = .false.
    ldiff = .false.
    lcouett = .false.
    lfext = .false.
    lchempot = .false.
    ! optional Widom insertions after the step (src/forcem.f90:126): REF_CHEMPOT holds the
    ! tot.inp values "ichmax lwmin lwmax lchmax ilmid" (src/main.f90:148-152) and the number
    ! of chempot calls to make on the same configuration
    nchcall = 0
    call get_environment_variable('REF_CHEMPOT_ZERO', chemenv)
    chzero = len_trim(chemenv) > 0
    call get_environment_variable('REF_CHEMPOT', chemenv)
    if (len_trim(chemenv) > 0) then
        read(chemenv, *) ichmax, lwmin, lwmax, lchmax, ilmid, nchcall
        lchempot = .true.
    end if
    dshx = 0.0 ; dshy = 0.0 ; dshz = 0.0
    n = len_trim(base)
    filfld = base(1:n)//'.fld'
    filtop = base(1:n)//'.top'
    filcfi = base(1:n)//'.cfg'
    filcfo = base(1:n)//'_n.cfg'
    filmeas = base(1:n)//'.crl'

    call sysdef_zero
    call sysdef_input
    call sysdef_alloc
    call config_zero
    ! as src/main.f90:249-300: rank 0 reads the configuration, the others receive it
    if (me == 0) call config_input
    if (nproc > 1) then
        call bcast_dbl(lx) ; call bcast_dbl(ly) ; call bcast_dbl(lz)
        call bcast_int1d(itype)
        call bcast_dbl1d(x0) ; call bcast_dbl1d(y0) ; call bcast_dbl1d(z0)
        call bcast_dbl1d(cx0) ; call bcast_dbl1d(cy0) ; call bcast_dbl1d(cz0)
        call bcast_dbl1d(wx0) ; call bcast_dbl1d(wy0) ; call bcast_dbl1d(wz0)
        call bcast_dbl1d(q10) ; call bcast_dbl1d(q20) ; call bcast_dbl1d(q30) ; call bcast_dbl1d(q40)
        call bcast_dbl1d(a11) ; call bcast_dbl1d(a12) ; call bcast_dbl1d(a13)
        call bcast_dbl1d(a21) ; call bcast_dbl1d(a22) ; call bcast_dbl1d(a23)
        call bcast_dbl1d(a31) ; call bcast_dbl1d(a32) ; call bcast_dbl1d(a33)
        call bcast_dbl1d(rsx) ; call bcast_dbl1d(rsy) ; call bcast_dbl1d(rsz)
    end if
    call sysdef_init
    call config_init
    call force_init
    call move_init
    call measure_init

    ! site offsets from the lead site, as the predictor leaves them
    do i = 1, nrig
        ix = 3*(i-1)
        rsx(ix+2:ix+3) = x0(ix+2:ix+3) - x0(ix+1)
        rsy(ix+2:ix+3) = y0(ix+2:ix+3) - y0(ix+1)
        rsz(ix+2:ix+3) = z0(ix+2:ix+3) - z0(ix+1)
    end do

    ! accumulators as force_get zeroes them; the hot-loop observables as measure_init does
    ngc = 0
    efipos = 0.0
    fx = 0.0 ; fy = 0.0 ; fz = 0.0
    fwsx = 0.0 ; fwsy = 0.0 ; fwsz = 0.0
    vir = 0.0 ; elj = 0.0 ; eqq = 0.0 ; uewa = 0.0

    allocate(f1x(nmove), f1y(nmove), f1z(nmove), f2x(nmove), f2y(nmove), f2z(nmove))
    call force_real
    f1x = fx ; f1y = fy ; f1z = fz
    call force_fouri
    ! the angle histogram as measure_init leaves it (src/measure.f90:96-99,198)
    if (allocated(ngang)) ngang = 0
    call force_intra
    f2x = fx ; f2y = fy ; f2z = fz

    ! selected k-vectors in the reference's enumeration order (src/forcem.f90:870-915)
    nsel = 0
    do ix = 1, 2
    if (ix == 2) allocate(kak(nsel), kre(nsel), kim(nsel), kvec(3*nsel))
    kcnt = 0
    nsel = 0
    kmn = 0
    knn = 1
    do kl = 0, klmax
        do km = kmn, kmmax
            do kn = knn, knmax
                kcnt = kcnt + 1
                if (myk(kcnt) >= 0) then
                    nsel = nsel + 1
                    if (ix == 1) cycle
                    kak(nsel) = Ak(kcnt)
                    kre(nsel) = dble(Qsum(kcnt))
                    kim(nsel) = dimag(Qsum(kcnt))
                    kvec(3*nsel-2) = dble(kl)
                    kvec(3*nsel-1) = dble(km)
                    kvec(3*nsel) = dble(kn)
                end if
            end do
            knn = -knmax
        end do
        kmn = -kmmax
    end do
    end do

    call force_sum

    ! Widom insertions on the step's configuration: cumulative pave/inave and the generator
    ! state after each call
    if (lchempot .and. nchcall > 0) then
        ! chempot's k loop runs over every kcount, also the ones force_fouri_init leaves
        ! unassigned (myk < 0, src/forcem.f90:757-770): define them as zero, as a fresh
        ! allocation would read
        if (chzero) then
            where (myk(1:kcnt) < 0) Ak(1:kcnt) = 0.0
            where (myk(1:kcnt) < 0) Qsum(1:kcnt) = (0.0, 0.0)
        end if
        allocate(chp(lchmax*nchcall), chi(lchmax*nchcall), chs(nchcall), chpar(13))
        do ic = 1, nchcall
            call chempot
            chp((ic-1)*lchmax+1:ic*lchmax) = pave(1:lchmax)
            chi((ic-1)*lchmax+1:ic*lchmax) = inave(1:lchmax)
            chs(ic) = seed
        end do
        chpar = (/ lwmin, lwmax, facbin, chtemper, lwh, ulr_w(1), ulr_l(1), flr_lja(1), flr_ljc(1), &
                   flr_bua(1), flr_buc(1), flr_bur(1), rcut /)
    end if

    if (nproc == 1) then
    u = 11
    open(u, file=trim(outname), access='stream', form='unformatted', status='replace')
    allocate(ibuf(max(ntota, int(ipack,8), int(wpack,8), nit*nit, nijt, 10_8)))
    ibuf(1:10) = (/ ntota, nmove, nriga, nliqa, nrig, nwall, nijt, int(ipack,8), int(wpack,8), nsel /)
    call put_i(u, 'counts', ibuf, 10_8)
    call put_r(u, 'x0', x0, ntota)
    call put_r(u, 'y0', y0, ntota)
    call put_r(u, 'z0', z0, ntota)
    call put_r(u, 'qs', qs, ntota)
    ibuf(1:ntota) = itype(1:ntota)
    call put_i(u, 'itype', ibuf, ntota)
    call put_r(u, 'cx0', cx0, nrig)
    call put_r(u, 'cy0', cy0, nrig)
    call put_r(u, 'cz0', cz0, nrig)
    ibuf(1:ipack) = list1(1:ipack)
    call put_i(u, 'list1', ibuf, int(ipack,8))
    ibuf(1:ipack) = list2(1:ipack)
    call put_i(u, 'list2', ibuf, int(ipack,8))
    ibuf(1:wpack) = listw1(1:wpack)
    call put_i(u, 'listw1', ibuf, int(wpack,8))
    ibuf(1:wpack) = listw2(1:wpack)
    call put_i(u, 'listw2', ibuf, int(wpack,8))
    call put_r(u, 'f_real_x', f1x, nmove)
    call put_r(u, 'f_real_y', f1y, nmove)
    call put_r(u, 'f_real_z', f1z, nmove)
    call put_r(u, 'f_site_x', f2x, nmove)
    call put_r(u, 'f_site_y', f2y, nmove)
    call put_r(u, 'f_site_z', f2z, nmove)
    call put_r(u, 'fcx', fcx, nrig)
    call put_r(u, 'fcy', fcy, nrig)
    call put_r(u, 'fcz', fcz, nrig)
    call put_r(u, 'tx', tx, nrig)
    call put_r(u, 'ty', ty, nrig)
    call put_r(u, 'tz', tz, nrig)
    call put_r(u, 'elj', elj, nijt)
    call put_r(u, 'eqq', eqq, nijt)
    ! observables of the hot loops (src/forcem.f90:347-354,388-395,486-495,930-934):
    ! ngc(nijt, 0:lmax) and efipos(2, 0:lmax), column-major
    call put_i(u, 'ngc', ngc, int(size(ngc), 8))
    call put_r(u, 'efipos', efipos, int(size(efipos), 8))
    if (lchempot .and. nchcall > 0) then
        call put_r(u, 'chem_pave', chp, lchmax*nchcall)
        call put_i(u, 'chem_inave', chi, lchmax*nchcall)
        call put_r(u, 'chem_seed', chs, nchcall)
        call put_r(u, 'chem_par', chpar, 13_8)
        ! the wall tail tables of every type (src/forcem.f90:170-228), for the force path's ulr
        call put_r(u, 'wall_ulr_w', ulr_w, int(nit, 8))
        call put_r(u, 'wall_ulr_l', ulr_l, int(nit, 8))
        call put_r(u, 'wall_lja', flr_lja, int(nit, 8))
        call put_r(u, 'wall_ljc', flr_ljc, int(nit, 8))
        call put_r(u, 'wall_bua', flr_bua, int(nit, 8))
        call put_r(u, 'wall_buc', flr_buc, int(nit, 8))
        call put_r(u, 'wall_bur', flr_bur, int(nit, 8))
    end if
    call put_r(u, 'k_ak', kak, nsel)
    call put_r(u, 'k_sre', kre, nsel)
    call put_r(u, 'k_sim', kim, nsel)
    call put_r(u, 'k_vec', kvec, 3*nsel)
    call put_s(u, 'uewa', uewa)
    call put_s(u, 'uself', uself)
    call put_s(u, 'urf', urf)
    call put_s(u, 'ulr', ulr)
    call put_s(u, 'uconf', uconf)
    call put_s(u, 'press', press)
    call put_s(u, 'ewqWW', ewqWW)
    call put_s(u, 'ewqLL', ewqLL)
    call put_s(u, 'ewqWL', ewqWL)
    call put_s(u, 'uavex', uavex)
    call put_s(u, 'alfa', alfa)
    call put_s(u, 'facq', facq)
    call put_s(u, 'spii', spii)
    call put_s(u, 'rskinsq', rskinsq)
    call put_r(u, 'fwsx', fwsx, 3_8)
    call put_r(u, 'fwsy', fwsy, 3_8)
    call put_r(u, 'fwsz', fwsz, 3_8)
    ibuf(1:3) = (/ int(klmax,8), int(kmmax,8), int(knmax,8) /)
    call put_i(u, 'kmaxes', ibuf, 3_8)

    ! the derived topology the force path reads (after sysdef_init's transforms), so a
    ! system can be rebuilt from the dump alone
    ibuf(1:4) = (/ nit, nltype, ntype, nwtype /)
    call put_i(u, 'types', ibuf, 4_8)
    ibuf(1:nit) = nsi(1:nit)
    call put_i(u, 'nsi', ibuf, nit)
    do i = 1, nit
        ibuf((i-1)*nit+1:i*nit) = ijtype(i, 1:nit)
    end do
    call put_i(u, 'ijtype', ibuf, nit*nit)
    ibuf(1:nijt) = vdwnum(1:nijt)
    call put_i(u, 'vdwnum', ibuf, nijt)
    call put_r(u, 'eps', eps, nijt)
    call put_r(u, 'sig', sig, nijt)
    call put_r(u, 'aa', aa, nijt)
    call put_r(u, 'rhoi', rhoi, nijt)
    call put_r(u, 'cc', cc, nijt)
    ! constraints, molecule by molecule (src/sysdef.f90:193-204)
    n = 0
    do i = 1, ncons
        n = n + nconst(ictype(i))
    end do
    allocate(cbuf(3*max(n, 1_8)))
    kcnt = 0
    do i = 1, ncons
        do ix = 1, nconst(ictype(i))
            kcnt = kcnt + 1
            cbuf(kcnt) = dble(constra(ix, i))
            cbuf(n + kcnt) = dble(constrb(ix, i))
            cbuf(2*n + kcnt) = length(ictype(i), ix)
        end do
    end do
    call put_r(u, 'constraints', cbuf, 3*n)
    deallocate(cbuf)
    ! angles, group by group (src/forcem.f90:571-576); theta in radians
    n = 0
    do i = 1, nang
        n = n + nangt(iatype(i))
    end do
    allocate(cbuf(5*max(n, 1_8)))
    kcnt = 0
    do i = 1, nang
        do ix = 1, nangt(iatype(i))
            kcnt = kcnt + 1
            cbuf(kcnt) = dble(ang1(ix, i))
            cbuf(n + kcnt) = dble(ang2(ix, i))
            cbuf(2*n + kcnt) = dble(ang3(ix, i))
            cbuf(3*n + kcnt) = kt(iatype(i), ix)
            cbuf(4*n + kcnt) = theta(iatype(i), ix)
        end do
    end do
    call put_r(u, 'angles', cbuf, 5*n)
    ! each angle term's group type (iatype), the first index of ngang
    kcnt = 0
    do i = 1, nang
        do ix = 1, nangt(iatype(i))
            kcnt = kcnt + 1
            cbuf(kcnt) = dble(iatype(i))
        end do
    end do
    call put_r(u, 'angle_type', cbuf, n)
    deallocate(cbuf)
    ! force_intra's angle histogram ngang(nit-2, 0:201), column-major (src/forcem.f90:617-625)
    if (allocated(ngang)) then
        ibuf(1:2) = (/ int(size(ngang, 1), 8), int(size(ngang, 2), 8) /)
        call put_i(u, 'ngang_shape', ibuf, 2_8)
        call put_i(u, 'ngang', ngang, int(size(ngang), 8))
    end if
    allocate(cbuf(12))
    cbuf(1:12) = (/ lx, ly, lz, lw, cx, cy, cz, dble(nss(ntype, 1)), dble(nss(ntype, 2)), rcut, skin, dble(kmax) /)
    call put_r(u, 'geometry', cbuf, 12_8)
    deallocate(cbuf)
    close(u)
    end if

    ! timing: repeated force evaluations on the same configuration (no list rebuild)
    if (nrep > 0) then
        call system_clock(c0, crate)
        do irep = 1, nrep
            fx = 0.0 ; fy = 0.0 ; fz = 0.0
            fwsx = 0.0 ; fwsy = 0.0 ; fwsz = 0.0
            vir = 0.0 ; elj = 0.0 ; eqq = 0.0 ; uewa = 0.0
            call force_real
            call force_fouri
            call force_intra
            call force_sum
        end do
        call system_clock(c1)
        if (me == 0) write(*, '(a,i8,a,es16.8)') 'TIMING nrep=', nrep, ' seconds=', dble(c1 - c0)/dble(crate)
    end if

    call mpi_finalize(ierr)

contains

    subroutine put_hdr(u, name, kind, cnt)
        integer :: u
        character(len=*) :: name
        integer*8 :: kind, cnt
        character(len=16) :: nm
        nm = name
        write(u) nm, kind, cnt
    end subroutine put_hdr

    subroutine put_r(u, name, a, cnt)
        integer :: u
        character(len=*) :: name
        integer*8 :: cnt
        real*8 :: a(*)
        call put_hdr(u, name, 1_8, cnt)
        if (cnt > 0) write(u) a(1:cnt)
    end subroutine put_r

    subroutine put_s(u, name, a)
        integer :: u
        character(len=*) :: name
        real*8 :: a
        real*8 :: b(1)
        b(1) = a
        call put_r(u, name, b, 1_8)
    end subroutine put_s

    subroutine put_i(u, name, a, cnt)
        integer :: u
        character(len=*) :: name
        integer*8 :: cnt
        integer*8 :: a(*)
        call put_hdr(u, name, 2_8, cnt)
        if (cnt > 0) write(u) a(1:cnt)
    end subroutine put_i

end program ref_harness
