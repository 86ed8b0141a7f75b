!----------------------------------------------------------------------------
! One force evaluation from Fortran through the C ABI: reads a system file
! (sysfile.py), creates the engine (force_init), evaluates the forces
! (force_get), writes site forces, COM forces, torques, energies.
! Usage: mdf_force_once system.bin forces.bin
!----------------------------------------------------------------------------
program mdf_force_once
    use, intrinsic :: iso_c_binding
    use mdforce_c
    implicit none
    type(mdf_topology) :: top
    type(mdf_params) :: par
    type(mdf_outputs) :: out
    type(c_ptr) :: sys
    integer(c_int) :: hdr(8), ipar(11), nss(2), nsteps, rc
    integer(c_int), allocatable, target :: itype(:), nsi(:), ijtype(:), vkind(:)
    real(c_double), allocatable, target :: q(:), eps4(:), sig6(:), aa(:), rhoi(:), cc(:)
    real(c_double), allocatable, target :: x(:), y(:), z(:), cx(:), cy(:), cz(:)
    real(c_double), allocatable, target :: fsite(:), fcom(:), tq(:), evdw(:), ecoul(:)
    real(c_double) :: wl(4), dpar(6), dt
    integer :: nt, nm, nr, nit, nijt
    character(len=256) :: fin, fout

    call get_command_argument(1, fin)
    call get_command_argument(2, fout)
    open(10, file=trim(fin), access='stream', form='unformatted', status='old')
    read(10) hdr
    nt = hdr(1); nm = hdr(2); nr = hdr(5); nit = hdr(6); nijt = hdr(8)
    allocate(itype(nt), q(nt), nsi(nit), ijtype(nit*nit), vkind(nijt))
    allocate(eps4(nijt), sig6(nijt), aa(nijt), rhoi(nijt), cc(nijt))
    read(10) itype, q, nsi, ijtype, vkind, eps4, sig6, aa, rhoi, cc
    read(10) wl, nss, dpar, ipar
    allocate(x(nt), y(nt), z(nt), cx(max(nr,1)), cy(max(nr,1)), cz(max(nr,1)))
    read(10) x, y, z
    read(10) cx(1:nr), cy(1:nr), cz(1:nr)
    read(10) nsteps, dt
    close(10)

    top%n_total = hdr(1); top%n_move = hdr(2); top%n_rigid_sites = hdr(3)
    top%n_liquid_sites = hdr(4); top%n_rig = hdr(5); top%n_types = hdr(6); top%n_ltype = hdr(7)
    top%site_type = c_loc(itype); top%charge = c_loc(q); top%nsi = c_loc(nsi)
    top%ijtype = c_loc(ijtype); top%vdw_kind = c_loc(vkind)
    top%vdw_eps4 = c_loc(eps4); top%vdw_sig6 = c_loc(sig6); top%vdw_a = c_loc(aa)
    top%vdw_rhoi = c_loc(rhoi); top%vdw_c = c_loc(cc)
    top%n_constraints = 0; top%con_a = c_null_ptr; top%con_b = c_null_ptr; top%con_len = c_null_ptr
    top%n_angles = 0; top%ang_i = c_null_ptr; top%ang_j = c_null_ptr; top%ang_k = c_null_ptr
    top%ang_kt = c_null_ptr; top%ang_theta = c_null_ptr
    top%wall_lw = wl(1); top%wall_cell = wl(2:4); top%wall_nss = nss
    par%box = dpar(1:3); par%rcut = dpar(4); par%skin = dpar(5); par%alpha = dpar(6)
    par%alpha_mode = ipar(1); par%kmax = ipar(2); par%z_mode = ipar(3); par%ew3dc = ipar(4)
    par%wall_lr = ipar(5); par%precision = ipar(6); par%const_mode = ipar(7)
    par%erfc_mode = ipar(8); par%max_neighbors = ipar(9); par%device = ipar(10); par%virial = ipar(11)

    rc = mdf_create(top, par, c_loc(x), c_loc(y), c_loc(z), c_loc(cx), c_loc(cy), c_loc(cz), sys)
    if (rc /= MDF_OK) then
        print *, 'mdf_create failed: ', trim(mdf_error_message())
        stop 1
    end if
    allocate(fsite(3*nm), fcom(3*max(nr,1)), tq(3*max(nr,1)), evdw(nijt), ecoul(nijt))
    out%f_site = c_loc(fsite); out%f_com = c_loc(fcom); out%torque = c_loc(tq)
    out%e_vdw = c_loc(evdw); out%e_coul_real = c_loc(ecoul)
    rc = mdf_compute(sys, c_loc(x), c_loc(y), c_loc(z), c_loc(cx), c_loc(cy), c_loc(cz), out)
    if (rc /= MDF_OK) then
        print *, 'mdf_compute failed: ', trim(mdf_error_message())
        stop 1
    end if
    open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
    write(11) fsite, fcom(1:3*nr), tq(1:3*nr), evdw, ecoul
    write(11) out%e_recip, out%e_recip_ll, out%e_recip_wl, out%e_recip_ww, out%e_self, &
              out%e_slab, out%e_wall_lr, out%e_conf, out%e_avex, out%press_wall
    close(11)
    call mdf_destroy(sys)
end program mdf_force_once
