!----------------------------------------------------------------------------
! NVE trajectory driver for the drift oracle (test infrastructure).
!
! The reference's own time loop (src/main.f90:38-47) without measurement and output:
! move_predict (Gear predictor for COMs, quaternions and angular momenta,
! src/move.f90:48-218), force_get, move_correct (src/move.f90:221-380), with the
! thermostat off (temper = -1, src/main.f90:124-128).  Set-up as in ref_harness.f90.
! Linked either with the reference's force module (CPU) or with the engine's drop-in
! module (md_sim_amd/fortran/force_mdf.f90), so the two runs differ only in the force
! evaluation.  Writes one line per step: step, htot, uconf, ke, tempnow, ulr (the wall
! tail energy, which the reference keeps out of uconf and htot).
! Optional: REF_CHEMPOT="ichmax lwmin lwmax lchmax ilmid" switches the Widom insertions on
! (tot.inp's chempot lines, src/main.f90:148-152); MD_NVE_OBS=<file> writes the measurement
! arrays the force path fills (ngc, efipos; pave, inave and the generator state with
! insertions) after the last step.
! Usage: md_nve <output> <nsteps>   (harness.inp and the system files in the cwd)
!----------------------------------------------------------------------------
program md_nve
    use move
    use measure
    use rndmod
    implicit none
    include 'mpif.h'
    integer :: ierr, u
    integer*8 :: n, nsteps, i, ix
    character*40 :: base
    character*256 :: outname, arg2, chemenv, obsname

    nproc = 1
    me = 0
    call mpi_init(ierr)
    call mpi_comm_size(mpi_comm_world, nproc, ierr)
    call mpi_comm_rank(mpi_comm_world, me, ierr)
    if (nproc /= 1) stop 'md_nve: run on one rank'
    call get_command_argument(1, outname)
    call get_command_argument(2, arg2)
    read(arg2, *) nsteps
    open(7, file='harness.inp', status='old')
    read(7, *) base
    read(7, *) dt, rcut, skin, rdel, kmax
    close(7)

    temper = -1.0
    timeq = 0.0
    timrun = dble(nsteps)*dt
    timprn = dt
    l2d = .false.
    lhb = .false.
    ldiff = .false.
    lcouett = .false.
    lfext = .false.
    lchempot = .false.
    call get_environment_variable('REF_CHEMPOT', chemenv)
    if (len_trim(chemenv) > 0) then
        read(chemenv, *) ichmax, lwmin, lwmax, lchmax, ilmid
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
    call config_input
    call sysdef_init
    call config_init
    call force_init
    call move_init
    call measure_init

    u = 11
    open(u, file=trim(outname), status='replace')
    do itime = 1, nsteps
        call move_predict
        call force_get
        call move_correct
        write(u, '(i8,5(1x,es24.16))') itime, htot, uconf, ke, tempnow, ulr
    end do
    close(u)
    call get_environment_variable('MD_NVE_OBS', obsname)
    if (len_trim(obsname) > 0) then
        open(12, file=trim(obsname), status='replace')
        write(12, '(a,i12)') 'ngc ', size(ngc, kind=8)
        write(12, '(i12)') ngc
        write(12, '(a,i12)') 'efipos ', size(efipos, kind=8)
        write(12, '(es26.17)') efipos
        if (lchempot) then
            write(12, '(a,i12)') 'pave ', lchmax
            write(12, '(es26.17)') pave(1:lchmax)
            write(12, '(a,i12)') 'inave ', lchmax
            write(12, '(i12)') inave(1:lchmax)
            write(12, '(a,i12)') 'seed ', 1_8
            write(12, '(es26.17)') seed
        end if
        close(12)
    end if
    call mpi_finalize(ierr)
end program md_nve
