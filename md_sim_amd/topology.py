"""A system as the reference's force path sees it after set-up: derived per-site tables,
type-pair tables in internal form, constraints, angles, frozen-crystal geometry and
positions.

``TopologyData`` is what ``sysdef_init`` / ``config_init`` leave in the reference's modules
(src/sysdef.f90:356-457, src/config.f90:55-95,303-329): site types numbered 1..nit, the
canonical pair numbering ``ijtype``, vdW parameters as 4 eps / sigma^6 / 1/rho, angles in
radians, unwrapped site positions and molecule centres of mass.  It is built from the
reference's input files (``refio.load_reference_inputs``) or from a dump of the reference's
own modules, and it goes to the engine through ``abi.build_abi`` like a generated
``systems.System``.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, Optional, Tuple

import numpy as np


@dataclass
class TopologyData:
    name: str
    box: Tuple[float, float, float]
    rcut: float
    skin: float
    kmax: int
    site_type: np.ndarray          # [ntota] 1-based site types
    charge: np.ndarray             # [ntota]
    nsi: np.ndarray                # [nit] sites carried by a lead of each type (0: not a lead)
    ijtype: np.ndarray             # [nit, nit] 1-based canonical pair numbers
    vdw_kind: np.ndarray           # [nijt] 0 zero, 1 Lennard-Jones, 2 Buckingham
    vdw_eps4: np.ndarray
    vdw_sig6: np.ndarray
    vdw_a: np.ndarray
    vdw_rhoi: np.ndarray
    vdw_c: np.ndarray
    n_move: int
    n_rigid_sites: int
    n_liquid_sites: int
    n_rig: int
    n_ltype: int
    wall_lw: float
    wall_cell: Tuple[float, float, float]
    wall_nss: Tuple[int, int]
    pos: np.ndarray                # [ntota, 3]
    com: np.ndarray                # [nrig, 3]
    constraints: Optional[Tuple[np.ndarray, np.ndarray, np.ndarray]] = None   # 0-based a, b, length
    angles: Optional[Tuple[np.ndarray, np.ndarray, np.ndarray, np.ndarray, np.ndarray]] = None
    angle_type: Optional[np.ndarray] = None   # 1-based group type of each angle term (iatype)
    z_periodic: bool = False
    dt: float = 1.0
    meta: Dict[str, object] = field(default_factory=dict)

    # the counts the host code reads from a System
    @property
    def ntota(self) -> int:
        return len(self.site_type)

    @property
    def nmove(self) -> int:
        return self.n_move

    @property
    def nrig(self) -> int:
        return self.n_rig

    @property
    def nriga(self) -> int:
        return self.n_rigid_sites

    @property
    def nliqa(self) -> int:
        return self.n_liquid_sites

    @property
    def nwall(self) -> int:
        return self.ntota - self.n_move

    @property
    def nit(self) -> int:
        return len(self.nsi)

    @property
    def nijt(self) -> int:
        return self.nit * (self.nit + 1) // 2

    def positions(self):
        p = self.pos
        c = self.com
        return (np.ascontiguousarray(p[:, 0]), np.ascontiguousarray(p[:, 1]), np.ascontiguousarray(p[:, 2]),
                np.ascontiguousarray(c[:, 0]), np.ascontiguousarray(c[:, 1]), np.ascontiguousarray(c[:, 2]))

    # ---- construction from a dump of the reference's modules ----
    @classmethod
    def from_dump(cls, d: Dict[str, np.ndarray], name: str = "dump") -> "TopologyData":
        """From the arrays the golden-vector harness writes (oracle/ref_harness.f90): counts,
        per-site tables, derived type tables, constraints, angles and geometry; site indices
        there are 1-based."""
        cnt = np.asarray(d["counts"]).astype(np.int64)
        ntota, nmove, nriga, nliqa, nrig = (int(v) for v in cnt[:5])
        nit, nltype = (int(v) for v in np.asarray(d["types"])[:2])
        geo = np.asarray(d["geometry"], np.float64)
        c = np.asarray(d["constraints"], np.float64)
        nc = len(c) // 3
        cons = None
        if nc:
            c = c.reshape(3, nc)
            cons = (c[0].astype(np.int32) - 1, c[1].astype(np.int32) - 1, c[2].copy())
        a = np.asarray(d["angles"], np.float64)
        na = len(a) // 5
        ang = None
        if na:
            a = a.reshape(5, na)
            ang = (a[0].astype(np.int32) - 1, a[1].astype(np.int32) - 1, a[2].astype(np.int32) - 1, a[3].copy(),
                   a[4].copy())
        atype = np.asarray(d["angle_type"]).astype(np.int32) if "angle_type" in d else None
        pos = np.stack([d["x0"], d["y0"], d["z0"]], 1).astype(np.float64)
        com = np.stack([d["cx0"], d["cy0"], d["cz0"]], 1).astype(np.float64) if nrig else np.zeros((0, 3))
        return cls(name=name, box=(float(geo[0]), float(geo[1]), float(geo[2])), rcut=float(geo[9]),
                   skin=float(geo[10]), kmax=int(geo[11]),
                   site_type=np.asarray(d["itype"]).astype(np.int32), charge=np.asarray(d["qs"], np.float64),
                   nsi=np.asarray(d["nsi"]).astype(np.int32),
                   ijtype=np.asarray(d["ijtype"]).astype(np.int32).reshape(nit, nit),
                   vdw_kind=np.asarray(d["vdwnum"]).astype(np.int32), vdw_eps4=np.asarray(d["eps"], np.float64),
                   vdw_sig6=np.asarray(d["sig"], np.float64), vdw_a=np.asarray(d["aa"], np.float64),
                   vdw_rhoi=np.asarray(d["rhoi"], np.float64), vdw_c=np.asarray(d["cc"], np.float64),
                   n_move=nmove, n_rigid_sites=nriga, n_liquid_sites=nliqa, n_rig=nrig, n_ltype=nltype,
                   wall_lw=float(geo[3]), wall_cell=(float(geo[4]), float(geo[5]), float(geo[6])),
                   wall_nss=(int(geo[7]), int(geo[8])), pos=pos, com=com, constraints=cons, angles=ang,
                   angle_type=atype)
