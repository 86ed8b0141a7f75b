"""ctypes mirror of include/mdforce.h and the conversion System -> ABI arguments.

The host side of the engine mirrors the reference's force module (src/forcem.f90:54-129):
``create`` ~ force_init, ``compute`` ~ force_get, ``zero`` ~ force_zero, and a non-zero
status ~ the reference's ``stop`` (raised as MDForceError with the library's message).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import Dict, Optional

import numpy as np

MDF_OK = 0
MDF_ERR_LIST_OVERFLOW = 3
STATUS_NAMES = {0: "OK", 1: "ERR_ARG", 2: "ERR_HIP", 3: "ERR_LIST_OVERFLOW", 4: "ERR_VDW_TYPE",
                5: "ERR_KSET", 6: "ERR_NONFINITE", 7: "ERR_NO_DEVICE"}
FP64, FP32_ACC64 = 0, 1
CONST_REFERENCE, CONST_EXACT = 0, 1
ERFC_AS7126, ERFC_ACCURATE = 0, 1
ALPHA_REFERENCE, ALPHA_GIVEN = 0, 1
Z_OPEN, Z_PERIODIC, Z_PERIODIC_OPEN_LIST = 0, 1, 2

_pi32 = C.POINTER(C.c_int32)
_pd = C.POINTER(C.c_double)


class Topology(C.Structure):
    _fields_ = [("n_total", C.c_int32), ("n_move", C.c_int32), ("n_rigid_sites", C.c_int32),
                ("n_liquid_sites", C.c_int32), ("n_rig", C.c_int32), ("n_types", C.c_int32),
                ("n_ltype", C.c_int32),
                ("site_type", _pi32), ("charge", _pd), ("nsi", _pi32), ("ijtype", _pi32),
                ("vdw_kind", _pi32), ("vdw_eps4", _pd), ("vdw_sig6", _pd), ("vdw_a", _pd),
                ("vdw_rhoi", _pd), ("vdw_c", _pd),
                ("n_constraints", C.c_int32), ("con_a", _pi32), ("con_b", _pi32), ("con_len", _pd),
                ("n_angles", C.c_int32), ("ang_i", _pi32), ("ang_j", _pi32), ("ang_k", _pi32),
                ("ang_kt", _pd), ("ang_theta", _pd),
                ("wall_lw", C.c_double), ("wall_cell", C.c_double * 3), ("wall_nss", C.c_int32 * 2)]


class Params(C.Structure):
    _fields_ = [("box", C.c_double * 3), ("rcut", C.c_double), ("skin", C.c_double),
                ("alpha", C.c_double), ("alpha_mode", C.c_int32), ("kmax", C.c_int32),
                ("z_mode", C.c_int32), ("ew3dc", C.c_int32), ("wall_lr", C.c_int32),
                ("precision", C.c_int32), ("const_mode", C.c_int32), ("erfc_mode", C.c_int32),
                ("max_neighbors", C.c_int32), ("device", C.c_int32), ("virial", C.c_int32)]


class Outputs(C.Structure):
    _fields_ = [("f_site", _pd), ("f_com", _pd), ("torque", _pd), ("e_vdw", _pd), ("e_coul_real", _pd),
                ("e_recip", C.c_double), ("e_recip_ll", C.c_double), ("e_recip_wl", C.c_double),
                ("e_recip_ww", C.c_double), ("e_self", C.c_double), ("e_slab", C.c_double),
                ("e_wall_lr", C.c_double), ("e_conf", C.c_double), ("e_avex", C.c_double),
                ("press_wall", C.c_double), ("f_wall", C.c_double * 9), ("virial", C.c_double * 9),
                ("list_rebuilt", C.c_int32), ("n_pairs_listed", C.c_int32),
                ("n_wall_pairs_listed", C.c_int32), ("n_kvec", C.c_int32), ("n_list_builds", C.c_int32)]


# reciprocal-force surrogate with multiple time steps (mdf_mts_params / mdf_mts_stats)
MTS_MODELS = {"held": 0, "network": 1}
PH_R = float(np.float32(8.31441))   # the reference's gas constant, a default-real literal (src/global.f90:21)
MTS_WEIGHTS = 8448


class MtsParams(C.Structure):
    _fields_ = [("interval", C.c_int32), ("model", C.c_int32), ("train_frames", C.c_int32),
                ("train_iters", C.c_int32), ("tol", C.c_double), ("lr", C.c_float), ("seed", C.c_uint32)]


class ChempotParams(C.Structure):
    _fields_ = [("n_insert", C.c_int32), ("n_bins", C.c_int32), ("bin_mid", C.c_int32), ("ew3dc", C.c_int32),
                ("wall_min", C.c_double), ("wall_max", C.c_double), ("temperature", C.c_double),
                ("seed", C.c_double), ("max_partners", C.c_int32), ("pad", C.c_int32)]


class MtsStats(C.Structure):
    _fields_ = [("state", C.c_int32), ("last_exact", C.c_int32), ("n_exact", C.c_int64),
                ("n_surrogate", C.c_int64), ("n_fallback", C.c_int64), ("n_judged", C.c_int64),
                ("adam_steps", C.c_int64), ("err_last", C.c_double), ("err_max", C.c_double),
                ("err_mean", C.c_double), ("loss_last", C.c_double)]


SCALAR_OUTPUTS = ("e_recip", "e_recip_ll", "e_recip_wl", "e_recip_ww", "e_self", "e_slab",
                  "e_wall_lr", "e_conf", "e_avex", "press_wall")


def _p(a: np.ndarray, t):
    return a.ctypes.data_as(C.POINTER(t)) if a is not None and a.size else None


@dataclass
class SystemArrays:
    """Contiguous arrays a Topology/Params pair points into (kept alive with them)."""
    top: Topology
    par: Params
    keep: Dict[str, np.ndarray]
    n_pairs: int


def build_abi(system, precision: int = FP64, const_mode: int = CONST_REFERENCE,
              erfc_mode: int = ERFC_AS7126, alpha: Optional[float] = None,
              z_mode: Optional[int] = None, ew3dc: Optional[bool] = None,
              wall_lr: Optional[bool] = None, max_neighbors: int = 0, device: int = 0,
              virial: bool = True, constraints=None, angles=None,
              n_liquid_sites: Optional[int] = None) -> SystemArrays:
    """Topology and parameters of a ``systems.System``.  Defaults follow the reference
    variant the system corresponds to: slab systems use src/forcem.f90 (open z, slab
    correction and wall tail on); bulk systems use src/forcem_back.f90 (periodic z in the
    pair loops over the reference's open-z list, no slab or wall terms).

    ``constraints`` = (site_a, site_b, length) and ``angles`` = (i, j, k, kt, theta) add the
    reference's surface-group terms (0-based sites): constrained surface pairs are left out
    of the pair list and enter the self energy (src/nnlist.f90:41-43,68-70;
    src/forcem.f90:808-819); angles add k(theta - theta0) bending forces on j and k
    (src/forcem.f90:565-614).  ``n_liquid_sites`` moves the liquid/surface boundary
    (``nliqa``): moving sites from that index on are surface sites."""
    from .topology import TopologyData
    s = system
    derived = isinstance(s, TopologyData)
    if derived:
        keep = dict(site_type=np.ascontiguousarray(s.site_type, np.int32),
                    charge=np.ascontiguousarray(s.charge, np.float64),
                    nsi=np.ascontiguousarray(s.nsi, np.int32),
                    ijtype=np.ascontiguousarray(s.ijtype.ravel(), np.int32),
                    vdw_kind=np.ascontiguousarray(s.vdw_kind, np.int32))
        for k in ("vdw_eps4", "vdw_sig6", "vdw_a", "vdw_rhoi", "vdw_c"):
            keep[k] = np.ascontiguousarray(getattr(s, k), np.float64)
        if constraints is None:
            constraints = s.constraints
        if angles is None:
            angles = s.angles
    else:
        ijtype, kind, eps4, sig6, A, rhoi, Cc = s.vdw_tables()
        keep = dict(site_type=np.ascontiguousarray(s.site_types(), np.int32),
                    charge=np.ascontiguousarray(s.charges(), np.float64),
                    nsi=np.ascontiguousarray(s.nsi(), np.int32),
                    ijtype=np.ascontiguousarray(ijtype.ravel(), np.int32),
                    vdw_kind=kind, vdw_eps4=eps4, vdw_sig6=sig6, vdw_a=A, vdw_rhoi=rhoi, vdw_c=Cc)
    top = Topology()
    top.n_total, top.n_move, top.n_rigid_sites = s.ntota, s.nmove, s.nriga
    top.n_liquid_sites, top.n_rig, top.n_types = s.nliqa, s.nrig, s.nit
    top.n_ltype = s.n_ltype if derived else s.nltype
    top.site_type = _p(keep["site_type"], C.c_int32)
    top.charge = _p(keep["charge"], C.c_double)
    top.nsi = _p(keep["nsi"], C.c_int32)
    top.ijtype = _p(keep["ijtype"], C.c_int32)
    top.vdw_kind = _p(keep["vdw_kind"], C.c_int32)
    for k in ("vdw_eps4", "vdw_sig6", "vdw_a", "vdw_rhoi", "vdw_c"):
        setattr(top, k, _p(keep[k], C.c_double))
    if n_liquid_sites is not None:
        top.n_liquid_sites = int(n_liquid_sites)
    top.n_constraints = 0
    if constraints is not None:
        ca, cb, cl = constraints
        keep["con_a"] = np.ascontiguousarray(ca, np.int32)
        keep["con_b"] = np.ascontiguousarray(cb, np.int32)
        keep["con_len"] = np.ascontiguousarray(cl, np.float64)
        top.n_constraints = len(keep["con_a"])
        top.con_a, top.con_b = _p(keep["con_a"], C.c_int32), _p(keep["con_b"], C.c_int32)
        top.con_len = _p(keep["con_len"], C.c_double)
    top.n_angles = 0
    if angles is not None:
        ai, aj, ak, kt, th = angles
        for k, v, t in (("ang_i", ai, np.int32), ("ang_j", aj, np.int32), ("ang_k", ak, np.int32),
                        ("ang_kt", kt, np.float64), ("ang_theta", th, np.float64)):
            keep[k] = np.ascontiguousarray(v, t)
            setattr(top, k, _p(keep[k], C.c_int32 if t is np.int32 else C.c_double))
        top.n_angles = len(keep["ang_i"])
    if derived:
        top.wall_lw = s.wall_lw
        top.wall_cell[:] = list(s.wall_cell)
        top.wall_nss[:] = list(s.wall_nss)
    else:
        cr = s.crystal
        top.wall_lw = cr.lw
        top.wall_cell[:] = list(cr.cell)
        top.wall_nss[:] = [cr.sites[0][1], cr.sites[1][1]]
    par = Params()
    par.box[:] = list(s.box)
    par.rcut, par.skin = s.rcut, s.skin
    par.alpha = alpha if alpha is not None else 0.0
    par.alpha_mode = ALPHA_GIVEN if alpha is not None else ALPHA_REFERENCE
    par.kmax = s.kmax
    bulk = bool(s.z_periodic)
    par.z_mode = z_mode if z_mode is not None else (Z_PERIODIC_OPEN_LIST if bulk else Z_OPEN)
    par.ew3dc = int(ew3dc if ew3dc is not None else not bulk)
    par.wall_lr = int(wall_lr if wall_lr is not None else not bulk)
    par.precision, par.const_mode, par.erfc_mode = precision, const_mode, erfc_mode
    par.max_neighbors = max_neighbors
    par.device = device
    par.virial = int(bool(virial))
    return SystemArrays(top, par, keep, s.nijt)


def positions(system):
    if hasattr(system, "positions"):
        return system.positions()
    pos = system.site_positions()
    x = np.ascontiguousarray(pos[:, 0])
    y = np.ascontiguousarray(pos[:, 1])
    z = np.ascontiguousarray(pos[:, 2])
    com = system.com
    return x, y, z, np.ascontiguousarray(com[:, 0]), np.ascontiguousarray(com[:, 1]), np.ascontiguousarray(com[:, 2])


class ForceResult(dict):
    """Outputs of one force evaluation as numpy arrays / floats."""

    @property
    def f_site(self):
        return self["f_site"]


def make_outputs(n_move: int, n_rig: int, n_pairs: int):
    arrs = dict(f_site=np.zeros(3 * n_move), f_com=np.zeros(3 * max(n_rig, 0)),
                torque=np.zeros(3 * max(n_rig, 0)), e_vdw=np.zeros(n_pairs), e_coul_real=np.zeros(n_pairs))
    o = Outputs()
    for k, a in arrs.items():
        setattr(o, k, _p(a, C.c_double))
    return o, arrs


def collect(o: Outputs, arrs, n_move: int, n_rig: int) -> ForceResult:
    r = ForceResult()
    r["f_site"] = arrs["f_site"].reshape(3, n_move).copy()
    r["f_com"] = arrs["f_com"].reshape(3, n_rig).copy()
    r["torque"] = arrs["torque"].reshape(3, n_rig).copy()
    r["e_vdw"] = arrs["e_vdw"].copy()
    r["e_coul_real"] = arrs["e_coul_real"].copy()
    for k in SCALAR_OUTPUTS:
        r[k] = float(getattr(o, k))
    r["f_wall"] = np.array(o.f_wall[:]).reshape(3, 3)
    r["virial"] = np.array(o.virial[:]).reshape(3, 3)
    r["list_rebuilt"] = int(o.list_rebuilt)
    r["n_pairs_listed"] = int(o.n_pairs_listed)
    r["n_wall_pairs_listed"] = int(o.n_wall_pairs_listed)
    r["n_kvec"] = int(o.n_kvec)
    r["n_list_builds"] = int(o.n_list_builds)
    return r
