"""Seeded synthetic systems for the force engine: bulk SPC/E water and SPC/E water on a
rock-salt slab, described the way the reference describes a system.

The reference reads a force field (``.fld``), a topology (``.top``) and a configuration
(``.cfg``) (src/sysdef.f90:83-340, src/config.f90:174-242) and derives per-site tables from
them in ``sysdef_init`` (src/sysdef.f90:356-457).  ``System`` holds the same information:

* molecule types in the reference's order: one rigid water type first, optional mobile ion
  types (``atmol``), and the frozen crystal last (src/sysdef.f90:226-249, 308);
* site types numbered 1..nit in that order (src/sysdef.f90:381-404), crystal site types
  last (src/sysdef.f90:406-419);
* the van der Waals table over canonical type pairs it<=jt (src/sysdef.f90:362-374);
* positions: rigid molecules as centre of mass + quaternion, rebuilt into sites with the
  reference's rotation convention (src/config.f90:303-329), then ions, then crystal sites.

Energies are kJ/mol, lengths Angstrom, as in the reference (src/forcem.f90:60).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

import numpy as np

KCAL = 4.184  # kJ per kcal

# SPC/E water: O-H 1 A, H-O-H 109.47 deg; charges -0.8476 / +0.4238 e; O-O LJ
SPCE_Q_O = -0.8476
SPCE_Q_H = 0.4238
SPCE_SIGMA = 3.166
SPCE_EPS = 0.1553 * KCAL
MASS_O, MASS_H = 15.9994e-3, 1.008e-3  # kg/mol, the reference's mass unit (src/sim/y.fld)
MASS_CAT, MASS_AN = 22.98977e-3, 35.453e-3


@dataclass
class MolType:
    name: str
    kind: str                      # 'rigmol' | 'atmol' | 'crystal'
    count: int                     # nt(it): molecules (1 for the crystal)
    sites: List[Tuple[str, int, float, float]]  # (name, nss, mass, charge) per site type
    geometry: Optional[np.ndarray] = None       # rigmol: (ns, 3) body coordinates
    nsites: int = 0                # crystal: ns(it) = number of crystal sites
    cell: Tuple[float, float, float] = (1.0, 1.0, 1.0)  # crystal: cx, cy, cz
    lw: float = 0.0                # crystal: pore width (walls at z = +-lw/2)

    @property
    def ns(self) -> int:
        if self.kind == "crystal":
            return self.nsites
        return sum(s[1] for s in self.sites)


@dataclass
class System:
    """One configuration plus its force field.  Arrays follow the reference's site order."""
    name: str
    box: Tuple[float, float, float]
    moltypes: List[MolType]
    vdw: Dict[Tuple[int, int], Tuple[str, Tuple[float, ...]]]  # (it, jt) -> (kind, params)
    com: np.ndarray                # (nrig, 3)
    quat: np.ndarray               # (nrig, 4)
    ion_pos: np.ndarray            # (nion_sites, 3)
    wall_pos: np.ndarray           # (nwall, 3)
    wall_kind: np.ndarray          # (nwall,) 1 or 2: crystal site type
    # run parameters (the reference's tot.inp, src/main.f90:122-155)
    rcut: float = 9.0
    kmax: int = 6
    skin: float = 1.0
    dt: float = 2.0
    rdel: float = 0.1
    temperature: float = 298.0
    z_periodic: bool = False       # True: bulk 3-D form (src/forcem_back.f90)
    com_vel: Optional[np.ndarray] = None   # (nrig, 3) A/fs
    ang_vel: Optional[np.ndarray] = None   # (nrig, 3) body-frame rad/fs
    ion_vel: Optional[np.ndarray] = None
    meta: Dict[str, object] = field(default_factory=dict)

    # ---- counts (src/sysdef.f90:298-319) ----
    @property
    def water(self) -> MolType:
        return self.moltypes[0]

    @property
    def crystal(self) -> MolType:
        return self.moltypes[-1]

    @property
    def nrig(self) -> int:
        return self.water.count

    @property
    def nriga(self) -> int:
        return self.water.count * self.water.ns

    @property
    def nliqa(self) -> int:
        return sum(m.count * m.ns for m in self.moltypes if m.kind in ("rigmol", "atmol"))

    @property
    def nmove(self) -> int:
        return self.nliqa

    @property
    def nwall(self) -> int:
        return self.crystal.nsites

    @property
    def ntota(self) -> int:
        return self.nmove + self.nwall

    @property
    def nit(self) -> int:
        return sum(len(m.sites) for m in self.moltypes)

    @property
    def nijt(self) -> int:
        return self.nit * (self.nit + 1) // 2

    @property
    def nltype(self) -> int:
        """Index (1-based) of the last liquid molecule type (src/sysdef.f90:226-249)."""
        last = 0
        for it, m in enumerate(self.moltypes, start=1):
            if m.kind in ("rigmol", "atmol"):
                last = it
        return last

    # ---- derived per-site tables (src/sysdef.f90:356-419) ----
    def body_coords(self) -> np.ndarray:
        """Rigid body coordinates relative to the centre of mass, as sysdef_init shifts
        them (src/sysdef.f90:430-441)."""
        w = self.water
        masses = np.array([s[2] for s in w.sites for _ in range(s[1])])
        g = np.array(w.geometry, dtype=np.float64)
        return g - (g * masses[:, None]).sum(0) / masses.sum()

    def site_types(self) -> np.ndarray:
        """1-based site type per site: moving sites by molecule site order, crystal sites
        remapped to the last two types (src/sysdef.f90:381-419)."""
        out = np.zeros(self.ntota, dtype=np.int32)
        i = 0
        base = 0
        for m in self.moltypes:
            if m.kind == "crystal":
                continue
            for _ in range(m.count):
                for ist, s in enumerate(m.sites):
                    out[i:i + s[1]] = base + ist + 1
                    i += s[1]
            base += len(m.sites)
        out[self.nmove:] = base + self.wall_kind
        return out

    def charges(self) -> np.ndarray:
        q = np.zeros(self.ntota)
        i = 0
        for m in self.moltypes:
            if m.kind == "crystal":
                continue
            for _ in range(m.count):
                for s in m.sites:
                    q[i:i + s[1]] = s[3]
                    i += s[1]
        cq = np.array([s[3] for s in self.crystal.sites])
        q[self.nmove:] = cq[self.wall_kind - 1]
        return q

    def nsi(self) -> np.ndarray:
        """Sites carried by a lead site of each type; 0 marks non-lead sites
        (src/sysdef.f90:387-392).  Crystal types keep 0."""
        out = np.zeros(self.nit, dtype=np.int32)
        base = 0
        for m in self.moltypes:
            if m.kind == "crystal":
                break
            for ist in range(len(m.sites)):
                if m.kind == "rigmol":
                    out[base + ist] = m.ns if ist == 0 else 0
                else:
                    out[base + ist] = 1
            base += len(m.sites)
        return out

    def vdw_tables(self):
        """Canonical pair numbering ijtype(it,jt) (src/sysdef.f90:362-369) and parameters in
        the reference's internal form: eps <- 4 eps, sig <- sig^6, rhoi <- 1/rho
        (src/sysdef.f90:372-374).  Returns (ijtype[nit,nit] 1-based, kind, eps4, sig6, A,
        rhoi, C) with kind 0 zero / 1 Lennard-Jones / 2 Buckingham."""
        nit = self.nit
        ijtype = np.zeros((nit, nit), dtype=np.int32)
        n = self.nijt
        kind = np.zeros(n, np.int32)
        eps4 = np.zeros(n)
        sig6 = np.zeros(n)
        A = np.zeros(n)
        rhoi = np.zeros(n)
        C = np.zeros(n)
        ijt = 0
        for it in range(1, nit + 1):
            for jt in range(it, nit + 1):
                ijt += 1
                ijtype[it - 1, jt - 1] = ijt
                ijtype[jt - 1, it - 1] = ijt
                k, p = self.vdw.get((it, jt), ("zero", ()))
                if k == "LeJo":
                    kind[ijt - 1] = 1
                    eps4[ijt - 1] = 4.0 * p[1]
                    x = p[0]
                    sig6[ijt - 1] = x * x * x * x * x * x   # sig**6 as the reference evaluates it
                elif k == "Buck":
                    kind[ijt - 1] = 2
                    A[ijt - 1] = p[0]
                    rhoi[ijt - 1] = 1.0 / p[1] if p[1] != 0.0 else 0.0
                    C[ijt - 1] = p[2]
        return ijtype, kind, eps4, sig6, A, rhoi, C

    def rotation(self) -> np.ndarray:
        """Rotation matrices a(3,3) per molecule from quaternions (src/config.f90:307-317).
        Site r = R + a^T b  (src/config.f90:319-326)."""
        q1, q2, q3, q4 = (self.quat[:, k] for k in range(4))
        a = np.empty((self.nrig, 3, 3))
        a[:, 0, 0] = -q1 * q1 + q2 * q2 - q3 * q3 + q4 * q4
        a[:, 0, 1] = 2.0 * (q3 * q4 - q1 * q2)
        a[:, 0, 2] = 2.0 * (q2 * q3 + q1 * q4)
        a[:, 1, 0] = -2.0 * (q1 * q2 + q3 * q4)
        a[:, 1, 1] = q1 * q1 - q2 * q2 - q3 * q3 + q4 * q4
        a[:, 1, 2] = 2.0 * (q2 * q4 - q1 * q3)
        a[:, 2, 0] = 2.0 * (q2 * q3 - q1 * q4)
        a[:, 2, 1] = -2.0 * (q1 * q3 + q4 * q2)
        a[:, 2, 2] = -q1 * q1 - q2 * q2 + q3 * q3 + q4 * q4
        return a

    def site_positions(self) -> np.ndarray:
        """(ntota, 3) site positions, rigid sites rebuilt from COM + quaternion in the
        reference's operation order (src/config.f90:319-326)."""
        b = self.body_coords()
        a = self.rotation()
        pos = np.empty((self.ntota, 3))
        ns = self.water.ns
        r = np.empty((self.nrig, ns, 3))
        for j in range(ns):
            r[:, j, 0] = self.com[:, 0] + a[:, 0, 0] * b[j, 0] + a[:, 1, 0] * b[j, 1] + a[:, 2, 0] * b[j, 2]
            r[:, j, 1] = self.com[:, 1] + a[:, 0, 1] * b[j, 0] + a[:, 1, 1] * b[j, 1] + a[:, 2, 1] * b[j, 2]
            r[:, j, 2] = self.com[:, 2] + a[:, 0, 2] * b[j, 0] + a[:, 1, 2] * b[j, 1] + a[:, 2, 2] * b[j, 2]
        pos[:self.nriga] = r.reshape(-1, 3)
        pos[self.nriga:self.nmove] = self.ion_pos
        pos[self.nmove:] = self.wall_pos
        return pos


# ---------------------------------------------------------------------------
# builders
# ---------------------------------------------------------------------------
def spce_geometry() -> np.ndarray:
    """O at the origin, hydrogens in the xy plane symmetric about y (principal axes)."""
    half = np.deg2rad(109.47) / 2.0
    return np.array([[0.0, 0.0, 0.0],
                     [np.sin(half), np.cos(half), 0.0],
                     [-np.sin(half), np.cos(half), 0.0]])


def random_quaternions(rng: np.random.Generator, n: int) -> np.ndarray:
    """Uniform random unit quaternions (Shoemake's subgroup algorithm)."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    a, b = np.sqrt(1.0 - u1), np.sqrt(u1)
    q = np.stack([a * np.sin(2 * np.pi * u2), a * np.cos(2 * np.pi * u2),
                  b * np.sin(2 * np.pi * u3), b * np.cos(2 * np.pi * u3)], axis=1)
    return q


def _water_type(n: int) -> MolType:
    return MolType("spce", "rigmol", n,
                   [("O", 1, MASS_O, SPCE_Q_O), ("H", 2, MASS_H, SPCE_Q_H)],
                   geometry=spce_geometry())


def _crystal_type(nsites: int, a: float, lw: float) -> MolType:
    # two crystal site types (the reference requires exactly two, src/forcem.f90:178-179);
    # cx*cy*cz = a^3/2 makes sum(nss)/(2 cx cy cz) the rock-salt ion density 8/a^3
    return MolType("rocksalt", "crystal", 1,
                   [("Na", 4, MASS_CAT, 1.0), ("Cl", 4, MASS_AN, -1.0)],
                   nsites=nsites, cell=(a, a, 0.5 * a), lw=lw)


def _maxwell(rng, n, mass_kg_mol, T):
    # sigma^2 = kT/m in (A/fs)^2: R T / m [J/kg] -> (m/s)^2 -> 1e-10 A/fs per m/s
    s = np.sqrt(8.31441 * T / mass_kg_mol) * 1e-5
    v = rng.normal(0.0, s, size=(n, 3))
    return v - v.mean(0)


def bulk_water(seed: int = 1000, n_side: int = 6, box: float = 18.64, rcut: float = 9.0,
               kmax: int = 6, T: float = 298.0) -> System:
    """Bulk rigid SPC/E water: COMs on an n^3 lattice, random orientations, 3-D periodic.
    216 molecules in an 18.64 A cube is 1.0 g/cm3."""
    rng = np.random.default_rng(seed)
    n = n_side ** 3
    sp = box / n_side
    g = (np.arange(n_side) + 0.5) * sp
    com = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    com = com + rng.uniform(-0.05, 0.05, size=com.shape)
    quat = random_quaternions(rng, n)
    water = _water_type(n)
    # the reference needs a crystal type even when it has no sites (src/sysdef.f90:308)
    crystal = _crystal_type(0, 5.64, box)
    vdw = {(1, 1): ("LeJo", (SPCE_SIGMA, SPCE_EPS))}
    mw = MASS_O + 2 * MASS_H
    sysm = System("bulk%d" % n, (box, box, box), [water, crystal], vdw, com, quat,
                  np.zeros((0, 3)), np.zeros((0, 3)), np.zeros(0, np.int32),
                  rcut=rcut, kmax=kmax, skin=1.0, dt=2.0, temperature=T, z_periodic=True,
                  com_vel=_maxwell(rng, n, mw, T), ang_vel=np.zeros((n, 3)))
    sysm.meta.update(seed=seed, config="bulk SPC/E water")
    return sysm


def rocksalt_slab(nx: int, ny: int, nlayers: int, a: float = 5.64, z_top: float = 0.0):
    """Rock-salt (100) slab: nx x ny conventional cells laterally, nlayers atomic layers
    (spacing a/2) going down from z_top.  Returns positions and crystal site type
    (1 cation, 2 anion)."""
    h = 0.5 * a
    ii, jj, kk = np.meshgrid(np.arange(2 * nx), np.arange(2 * ny), np.arange(nlayers), indexing="ij")
    ii, jj, kk = ii.ravel(), jj.ravel(), kk.ravel()
    pos = np.stack([(ii + 0.25) * h, (jj + 0.25) * h, z_top - kk * h], axis=1)
    kind = np.where((ii + jj + kk) % 2 == 0, 1, 2).astype(np.int32)
    order = np.lexsort((ii, jj, kk))
    return pos[order], kind[order]


def water_film(rng, n: int, lx: float, ly: float, z0: float, spacing: float = 3.1):
    """n water COMs on a seeded lattice filling [0,lx)x[0,ly) from z0 upward."""
    nxl = max(1, int(round(lx / spacing)))
    nyl = max(1, int(round(ly / spacing)))
    per = nxl * nyl
    nz = -(-n // per)
    gx = (np.arange(nxl) + 0.5) * lx / nxl
    gy = (np.arange(nyl) + 0.5) * ly / nyl
    gz = z0 + np.arange(nz) * spacing
    grid = np.stack(np.meshgrid(gx, gy, gz, indexing="ij"), -1).reshape(-1, 3)
    pick = rng.permutation(len(grid))[:n]
    com = grid[np.sort(pick)]
    com = com + rng.uniform(-0.05, 0.05, size=com.shape)
    return com


def _buckingham_table(rng, n_ion_types: int):
    """Ion-oxygen Buckingham parameters in kJ/mol: A ~ U[1e4,1e5] kcal/mol, rho = 0.3 A,
    C ~ U[100,1000] kcal A^6/mol, seeded per ion type."""
    out = []
    for _ in range(n_ion_types):
        A = rng.uniform(1e4, 1e5) * KCAL
        C = rng.uniform(100.0, 1000.0) * KCAL
        out.append((A, 0.3, C))
    return out


def water_on_slab(n_water: int, ncell: int, nlayers: int, seed: int, rcut: float,
                  kmax: int, a: float = 5.64, lw: float = 40.0, T: float = 300.0,
                  n_mobile_ions: int = 0, name: str = "slab", gap: float = 2.8) -> System:
    """SPC/E water film above a frozen rock-salt (100) slab, slab-corrected 3-D Ewald.

    Lateral box = ncell * a; Lz = 3 Lx so that the reference's k-index rule gives
    kmax x kmax x 3 kmax (src/forcem.f90:673-678).  The crystal's top layer sits on the
    lower wall plane z = -lw/2 used by the reference's wall tail correction
    (src/forcem.f90:509-543); the water starts ``gap`` (2.8) A above it.
    Optional mobile ions (one cation and one anion type) are mixed into the film."""
    rng = np.random.default_rng(seed)
    lx = ncell * a
    lz = 3.0 * lx
    zwall = -0.5 * lw
    wall_pos, wall_kind = rocksalt_slab(ncell, ncell, nlayers, a, z_top=zwall)
    n_items = n_water + n_mobile_ions
    com_all = water_film(rng, n_items, lx, lx, zwall + gap)
    pick = rng.permutation(n_items)
    wsel = np.sort(pick[:n_water])
    isel = np.sort(pick[n_water:])
    com = com_all[wsel]
    ion_pos = com_all[isel]
    quat = random_quaternions(rng, n_water)
    moltypes = [_water_type(n_water)]
    vdw: Dict[Tuple[int, int], Tuple[str, Tuple[float, ...]]] = {(1, 1): ("LeJo", (SPCE_SIGMA, SPCE_EPS))}
    nion_t = 0
    if n_mobile_ions:
        ncat = n_mobile_ions // 2
        nan = n_mobile_ions - ncat
        moltypes.append(MolType("Na", "atmol", ncat, [("Na", 1, MASS_CAT, 1.0)]))
        moltypes.append(MolType("Cl", "atmol", nan, [("Cl", 1, MASS_AN, -1.0)]))
        nion_t = 2
        # cations first, then anions (site order follows molecule types)
        ion_pos = ion_pos[np.r_[np.arange(ncat), ncat + np.arange(nan)]]
    moltypes.append(_crystal_type(len(wall_pos), a, lw))
    buck = _buckingham_table(rng, 2)
    # types: 1 O, 2 H, [3 Na, 4 Cl mobile], last two crystal Na/Cl
    wt1 = 3 + nion_t
    for k in range(2):
        vdw[(1, wt1 + k)] = ("Buck", buck[k])
    if n_mobile_ions:
        mb = _buckingham_table(rng, 2)
        vdw[(1, 3)] = ("Buck", mb[0])
        vdw[(1, 4)] = ("Buck", mb[1])
        # ion-ion repulsion keeps mobile ions off the crystal and each other
        for it in (3, 4):
            for jt in (it, 4, wt1, wt1 + 1):
                if jt >= it:
                    vdw[(it, jt)] = ("LeJo", (2.6, 0.5))
    mw = MASS_O + 2 * MASS_H
    sysm = System(name, (lx, lx, lz), moltypes, vdw, com, quat, ion_pos, wall_pos, wall_kind,
                  rcut=rcut, kmax=kmax, skin=1.0, dt=2.0, temperature=T, z_periodic=False,
                  com_vel=_maxwell(rng, n_water, mw, T), ang_vel=np.zeros((n_water, 3)),
                  ion_vel=np.zeros((len(ion_pos), 3)))
    sysm.meta.update(seed=seed, ncell=ncell, nlayers=nlayers, lw=lw)
    return sysm


def rocksalt_bulk(ncell: int = 4, a: float = 5.64, rcut: float = 11.0, kmax: int = 10,
                  shake: float = 0.0, seed: int = 1200) -> System:
    """A periodic rock-salt crystal of mobile +-1 e ions (ncell^3 conventional cells, no
    water, no van der Waals), 3-D periodic: its Coulomb energy per ion pair is the
    Madelung constant times -facq / (a/2).  ``shake`` displaces every ion by a seeded
    uniform amount in [-shake, shake] per axis."""
    rng = np.random.default_rng(seed)
    h = 0.5 * a
    n = 2 * ncell
    ii, jj, kk = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    pos = np.stack([(ii + 0.25) * h, (jj + 0.25) * h, (kk + 0.25) * h], -1).reshape(-1, 3)
    cation = ((ii + jj + kk) % 2 == 0).ravel()
    pos = np.concatenate([pos[cation], pos[~cation]])
    if shake:
        pos = pos + rng.uniform(-shake, shake, size=pos.shape)
    box = n * h
    ncat = int(cation.sum())
    moltypes = [_water_type(0),
                MolType("Na", "atmol", ncat, [("Na", 1, MASS_CAT, 1.0)]),
                MolType("Cl", "atmol", len(pos) - ncat, [("Cl", 1, MASS_AN, -1.0)]),
                _crystal_type(0, a, box)]
    sysm = System("nacl%d" % ncell, (box, box, box), moltypes, {}, np.zeros((0, 3)), np.zeros((0, 4)),
                  pos, np.zeros((0, 3)), np.zeros(0, np.int32), rcut=rcut, kmax=kmax, skin=1.0,
                  z_periodic=True)
    sysm.meta.update(seed=seed, config="rock-salt crystal of mobile ions")
    return sysm


def replica_set(n: int = 64, seed: int = 3000, a: float = 5.64):
    """Independent interface systems of uneven size for batched stepping: per replica (seeded
    with seed + r) a rock-salt slab of 6-12 cells laterally (4 ncell^2 ions per layer, 256-2048
    ions in all), 512-4096 waters (fewer when the film would not fit the box), Lz = 3 Lx,
    |k| range of the 2048-water system (k-index 8 at 12 cells), rc 10 A, a temperature in
    280-360 K for the velocities."""
    out = []
    for r in range(n):
        rng = np.random.default_rng(seed + r)
        ncell = int(rng.integers(6, 13))
        lx = ncell * a
        per_layer_ions = 4 * ncell * ncell
        nlayers = int(np.clip(round(rng.uniform(256, 2048) / per_layer_ions), 2, 8))
        while per_layer_ions * nlayers > 2048 and nlayers > 1:
            nlayers -= 1
        per_layer_w = max(1, int(round(lx / 3.1))) ** 2
        lz = 3.0 * lx
        max_w = per_layer_w * int((0.5 * lz - 8.0) / 3.1)
        n_water = int(min(rng.integers(512, 4097), max_w))
        height = -(-n_water // per_layer_w) * 3.1
        lw = min(0.6 * lz, height + 20.0)
        kmax = max(4, int(round(8 * lx / (12 * a))))
        T = float(rng.uniform(280.0, 360.0))
        out.append(water_on_slab(n_water, ncell, nlayers, seed=seed + r, rcut=10.0, kmax=kmax, lw=lw, T=T,
                                 name="replica%02d" % r))
    return out


# ---------------------------------------------------------------------------
# the named workloads
# ---------------------------------------------------------------------------
def config(name: str) -> System:
    """Named workloads:

    * ``bulk216``  - 216 SPC/E waters, 18.64 A cube, rc 9, kmax 6, 3-D periodic;
    * ``pore_small`` - 96 waters + 8 mobile ions on a 2x2-cell, 4-layer slab (parity case);
    * ``pore_water`` - the same 96 waters without mobile ions (NVE drift case);
    * ``pore_sparse`` - 8 waters + 4 mobile ions in one layer 2.8 A above the same slab, the
      rest of the pore empty (Widom insertion case);
    * ``slab2048`` - 2048 waters on a 12x12x(8 layers) rock-salt slab (4608 ions),
      rc 10, k 8/8/24;
    * ``slab16k``  - 16384 waters on a 16x16x(6 layers) slab (6144 ions), rc 10,
      k 11/11/33 (|k|max of slab2048 kept);
    * ``slab16k_k8`` - the same with k 8/8/24;
    * ``nacl4``    - 512 mobile ions in a periodic 4x4x4-cell rock-salt crystal (closed-form
      Madelung energy);
    * ``slab512``  - 512 waters on a 6x6x(4 layers) slab (576 ions), rc 10, k 6/6/18 (small
      case of the surrogate with multiple time steps);
    * ``slab32k``  - 32768 waters on a 32x32x(3 layers) slab (12288 ions), box 180.48 x 180.48
      x 541.44 A, rc 10, k 21/21/63 (the surrogate's large interface).
    """
    if name == "bulk216":
        return bulk_water()
    if name == "pore_small":
        return water_on_slab(96, 2, 4, seed=1101, rcut=5.0, kmax=3, lw=20.0,
                             n_mobile_ions=8, name=name)
    if name == "pore_sparse":
        return water_on_slab(8, 2, 4, seed=1103, rcut=5.0, kmax=3, lw=20.0,
                             n_mobile_ions=4, name=name)
    if name == "pore_water":
        return water_on_slab(96, 2, 4, seed=1101, rcut=5.0, kmax=3, lw=20.0, name=name)
    if name == "slab2048":
        return water_on_slab(2048, 12, 8, seed=1001, rcut=10.0, kmax=8, name=name)
    if name == "slab16k":
        return water_on_slab(16384, 16, 6, seed=1002, rcut=10.0, kmax=11, lw=80.0, name=name)
    if name == "slab16k_k8":
        return water_on_slab(16384, 16, 6, seed=1002, rcut=10.0, kmax=8, lw=80.0, name=name)
    if name == "nacl4":
        return rocksalt_bulk(4)
    if name == "slab512":
        return water_on_slab(512, 6, 4, seed=1201, rcut=10.0, kmax=6, lw=40.0, name=name)
    if name == "slab32k":
        return water_on_slab(32768, 32, 3, seed=1004, rcut=10.0, kmax=21, lw=80.0, name=name)
    raise KeyError(name)


def libration_frames(system, n_steps, theta0=0.3, amp=0.2, seed=0):
    """Positions of a synthetic trajectory with rotations: every rigid molecule librates
    about its COM (angle theta0 sin(2 pi t / P + phi) about its own random axis, P in
    30..60 steps) while its COM oscillates (amplitude amp per axis, P in 40..80 steps);
    other moving sites oscillate like COMs.  Peak H speeds are ~0.05 A per step, about
    those of SPC/E water at 300 K with 2 fs steps.  Returns a list of (x, y, z, cx, cy, cz)."""
    from md_sim_amd import abi
    x0, y0, z0, cx0, cy0, cz0 = abi.positions(system)
    nr, nm = system.nrig, system.nmove
    rng = np.random.default_rng(seed)
    axis = rng.normal(size=(nr, 3))
    axis /= np.linalg.norm(axis, axis=1, keepdims=True)
    pr = rng.uniform(30, 60, nr)
    phr = rng.uniform(0, 2 * np.pi, nr)
    pt = rng.uniform(40, 80, (nm, 3))
    pht = rng.uniform(0, 2 * np.pi, (nm, 3))
    at = rng.uniform(-amp, amp, (nm, 3))
    com0 = np.stack([cx0, cy0, cz0], 1)
    site0 = np.stack([x0, y0, z0], 1)
    frames = []
    for t in range(n_steps):
        pos = site0.copy()
        th = theta0 * np.sin(2 * np.pi * t / pr + phr) - theta0 * np.sin(phr)
        c, s = np.cos(th), np.sin(th)
        shift = at * (np.sin(2 * np.pi * t / pt + pht) - np.sin(pht))
        com = com0 + shift[:3 * nr:3] if nr else com0
        for a in range(3):
            rel = site0[a:3 * nr:3] - com0
            kx = np.cross(axis, rel)
            kd = (axis * rel).sum(1, keepdims=True)
            rot = rel * c[:, None] + kx * s[:, None] + axis * kd * (1 - c)[:, None]
            pos[a:3 * nr:3] = com + rot
        if nm > 3 * nr:
            pos[3 * nr:nm] += shift[3 * nr:nm]
        frames.append((pos[:, 0].copy(), pos[:, 1].copy(), pos[:, 2].copy(),
                       com[:, 0].copy(), com[:, 1].copy(), com[:, 2].copy()))
    return frames
