"""Binary system file read by the Fortran programs in this directory (stream access,
little-endian): counts, topology tables, run parameters, positions."""
import numpy as np

from .. import abi


def write_sysfile(system, path, positions=None, nsteps=0, dt=None, **abi_kw):
    sa = abi.build_abi(system, **abi_kw)
    t, p, k = sa.top, sa.par, sa.keep
    x, y, z, cx, cy, cz = positions if positions is not None else abi.positions(system)
    with open(path, "wb") as f:
        w = lambda a, dt_: f.write(np.ascontiguousarray(a, dt_).tobytes())
        w([t.n_total, t.n_move, t.n_rigid_sites, t.n_liquid_sites, t.n_rig, t.n_types, t.n_ltype, system.nijt], np.int32)
        w(k["site_type"], np.int32)
        w(k["charge"], np.float64)
        w(k["nsi"], np.int32)
        w(k["ijtype"], np.int32)
        w(k["vdw_kind"], np.int32)
        for n in ("vdw_eps4", "vdw_sig6", "vdw_a", "vdw_rhoi", "vdw_c"):
            w(k[n], np.float64)
        w([t.wall_lw, *t.wall_cell], np.float64)
        w(list(t.wall_nss), np.int32)
        w([*p.box, p.rcut, p.skin, p.alpha], np.float64)
        w([p.alpha_mode, p.kmax, p.z_mode, p.ew3dc, p.wall_lr, p.precision, p.const_mode, p.erfc_mode,
           p.max_neighbors, p.device, p.virial], np.int32)
        for a in (x, y, z, cx, cy, cz):
            w(a, np.float64)
        w([nsteps], np.int32)
        w([dt if dt is not None else system.dt], np.float64)


def read_forces(path, n_move, n_rig, nijt):
    raw = np.fromfile(path, dtype=np.float64)
    o = 0
    out = {}
    out["f_site"] = raw[o:o + 3 * n_move].reshape(3, n_move); o += 3 * n_move
    out["f_com"] = raw[o:o + 3 * n_rig].reshape(3, n_rig); o += 3 * n_rig
    out["torque"] = raw[o:o + 3 * n_rig].reshape(3, n_rig); o += 3 * n_rig
    out["e_vdw"] = raw[o:o + nijt]; o += nijt
    out["e_coul_real"] = raw[o:o + nijt]; o += nijt
    for kname in abi.SCALAR_OUTPUTS:
        out[kname] = raw[o]; o += 1
    return out
