"""Readers and writers for the reference's text formats.

* run parameters ``tot.inp`` - read by input (src/main.f90:122-155);
* force field ``.fld`` - read by sysdef_input (src/sysdef.f90:92-171);
* topology ``.top``    - read by sysdef_input (src/sysdef.f90:174-216);
* configuration ``.cfg`` - read by config_input (src/config.f90:174-242, record layouts
  src/config.f90:395-481).

The writers turn a generated ``systems.System`` into files the reference (and these readers)
accept; the readers turn files - the reference's own or ours - into a ``TopologyData`` with
the tables the reference derives from them, which drives the engine and the oracle.

Every record is read list-directed by the reference, so numbers are written with 17
significant digits and round-trip exactly.  Velocities and higher derivatives are written
in the reference's 1 fs units (config_init rescales them by dt, src/config.f90:75-90).
"""
from __future__ import annotations

import os
from typing import Dict

import numpy as np

from .systems import System


def _g(x: float) -> str:
    return "%.17g" % float(x)


def write_fld(s: System, path: str) -> None:
    """Force field: molecule/site blocks then the vdW table in canonical (it<=jt) order,
    since sysdef_init renumbers pairs that way (src/sysdef.f90:362-369)."""
    L = ["# MOLECULES", str(len(s.moltypes))]
    for it, m in enumerate(s.moltypes, start=1):
        L += ["# Molecule", "%d %s %s %d" % (it, m.name, m.kind, m.count), "# sites", str(len(m.sites))]
        for name, nss, mass, q in m.sites:
            L += ["%s %d" % (name, nss), "%s %s" % (_g(mass), _g(q))]
        if m.kind == "rigmol":
            L.append("# geometry")
            for r in m.geometry:
                L.append(" ".join(_g(v) for v in r))
        elif m.kind == "crystal":
            L += [str(m.nsites), " ".join(_g(c) for c in m.cell), _g(m.lw), "0"]
    L.append("# VDW INTERACTIONS")
    names = []
    for it, m in enumerate(s.moltypes, start=1):
        for sname, *_ in m.sites:
            names.append((m.name, sname))
    nit = s.nit
    ijt = 0
    for it in range(1, nit + 1):
        for jt in range(it, nit + 1):
            ijt += 1
            kind, p = s.vdw.get((it, jt), ("zero", ()))
            L += [str(ijt), "%d %s %s" % (it, *names[it - 1]), "%d %s %s" % (jt, *names[jt - 1]), kind]
            if kind == "LeJo":
                L.append("%s %s" % (_g(p[0]), _g(p[1])))
            elif kind == "Buck":
                L.append("%s %s %s" % (_g(p[0]), _g(p[1]), _g(p[2])))
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")


def write_top(s: System, path: str) -> None:
    L = ["# Atom connections, specific interactions"]
    for it, m in enumerate(s.moltypes, start=1):
        L.append("%d %s %s %d" % (it, m.name, m.kind, m.count))
        if m.kind == "rigmol":
            L.append("sites: %d" % m.ns)
        elif m.kind == "atmol":
            L.append("no constraints")
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")


def write_cfg(s: System, path: str) -> None:
    """Configuration: box, molecule records (rigid: index + site types, COM and four
    derivatives, quaternion and four derivatives, angular velocity and four derivatives),
    ion records, crystal records (index, crystal site type, position), thermostat."""
    dt = s.dt
    L = ["# SIMULATION BOX SIZE", " ".join(_g(v) for v in s.box), "",
         "# NUMBER OF PARTICLE TYPES", str(len(s.moltypes)), "", "# PARTICLE COORDINATES"]
    zero3 = "0 0 0"
    zero4 = "0 0 0 0"
    site = 0
    for it, m in enumerate(s.moltypes, start=1):
        L.append("%s %s %d" % (m.name, m.kind, m.count))
        if m.kind == "rigmol":
            types = []
            for ist, (_, nss, _, _) in enumerate(m.sites):
                types += [ist + 1] * nss
            v = s.com_vel if s.com_vel is not None else np.zeros((m.count, 3))
            w = s.ang_vel if s.ang_vel is not None else np.zeros((m.count, 3))
            for i in range(m.count):
                L.append("%d %s" % (i + 1, " ".join(str(t) for t in types)))
                L.append(" ".join(_g(c) for c in s.com[i]))
                L.append(" ".join(_g(c) for c in v[i]))
                L += [zero3] * 3
                L.append(" ".join(_g(c) for c in s.quat[i]))
                L += [zero4] * 4
                L.append(" ".join(_g(c) for c in w[i]))
                L += [zero3] * 4
            site += m.count * m.ns
        elif m.kind == "atmol":
            ion0 = site - s.nriga
            v = s.ion_vel if s.ion_vel is not None else np.zeros((len(s.ion_pos), 3))
            for k in range(m.count * m.ns):
                L.append("%d %d" % (site + 1, 1))
                L.append(" ".join(_g(c) for c in s.ion_pos[ion0 + k]))
                L.append(" ".join(_g(c) for c in v[ion0 + k]))
                L += [zero3] * 3
                site += 1
        else:
            for k in range(s.nwall):
                L.append("%d %d" % (s.nmove + k + 1, int(s.wall_kind[k])))
                L.append(" ".join(_g(c) for c in s.wall_pos[k]))
        L.append("")
    L += ["# THERMOSTAT", "1 0 0 0 0", "1 0 0 0 0"]
    _ = dt
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")


def write_harness_input(s: System, path: str, base: str) -> None:
    """Run parameters for the reference harness (the subset of tot.inp that the force path
    reads, src/main.f90:122-155)."""
    with open(path, "w") as f:
        f.write("%s\n%s %s %s %s %d\n" % (base, _g(s.dt), _g(s.rcut), _g(s.skin), _g(s.rdel), s.kmax))


def write_tot_inp(s: System, path: str, base: str, ensemble: str = "nve", timeq: float = 0.0,
                  timrun: float = 0.0) -> None:
    """Run parameters in the reference's tot.inp layout (src/main.f90:122-155); every optional
    analysis switch off."""
    L = ["ensemble: %s" % ensemble]
    if ensemble == "nvt":
        L.append("temper: %s" % _g(s.temperature))
    L += ["time: %s %s %s" % (_g(timeq), _g(timrun), _g(s.dt)), "tprn: %s" % _g(s.dt),
          "rcut: %s" % _g(s.rcut), "kmax: %d" % s.kmax, "skin: %s" % _g(s.skin), "rdel: %s" % _g(s.rdel),
          "l2d: F", "lhb: F", "ldiff: F", "lcouett: F", "lfext: F", "lchem: F", "filname: %s" % base]
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")


def write_reference_inputs(s: System, directory: str, base: str = "sys") -> Dict[str, str]:
    os.makedirs(directory, exist_ok=True)
    p = {k: os.path.join(directory, base + "." + k) for k in ("fld", "top", "cfg")}
    write_fld(s, p["fld"])
    write_top(s, p["top"])
    write_cfg(s, p["cfg"])
    p["inp"] = os.path.join(directory, "harness.inp")
    write_harness_input(s, p["inp"], base)
    p["tot"] = os.path.join(directory, "tot.inp")
    write_tot_inp(s, p["tot"], base)
    return p


# ---------------------------------------------------------------------------
# readers: the reference's input files -> TopologyData
# ---------------------------------------------------------------------------
def _f32w(v: float) -> float:
    """A default-real (FP32) literal widened to REAL*8, as the reference's constants are."""
    return float(np.float32(v))


class _ListInput:
    """Fortran list-directed input over a text: every read starts at a new record; a read
    that needs values skips blank records; a read with no items consumes one record."""

    def __init__(self, text: str):
        self.lines = text.splitlines()
        self.i = 0

    def skip(self):
        self.i += 1

    def tokens(self, n: int):
        out = []
        first = True
        while len(out) < n:
            if self.i >= len(self.lines):
                raise ValueError("unexpected end of input")
            toks = self.lines[self.i].replace(",", " ").split()
            self.i += 1
            if not toks and first:
                continue
            first = False
            out += toks
        return out[:n]

    def reals(self, n: int):
        return [float(t.replace("d", "e").replace("D", "e")) for t in self.tokens(n)]

    def ints(self, n: int):
        return [int(t) for t in self.tokens(n)]


def parse_inp(text: str) -> Dict[str, object]:
    """tot.inp (src/main.f90:122-155): ensemble, temperature, times, print interval, rcut,
    kmax, skin, rdel, optional analysis switches, output name."""
    r = _ListInput(text)
    out: Dict[str, object] = {}
    ens = r.tokens(2)[1]
    out["ensemble"] = ens
    if ens == "nvt":
        out["temper"] = float(r.tokens(2)[1])
    t = r.tokens(4)
    out["timeq"], out["timrun"], out["dt"] = float(t[1]), float(t[2]), float(t[3])
    out["timprn"] = float(r.tokens(2)[1])
    out["rcut"] = float(r.tokens(2)[1])
    out["kmax"] = int(r.tokens(2)[1])
    out["skin"] = float(r.tokens(2)[1])
    out["rdel"] = float(r.tokens(2)[1])

    def flag():
        v = r.tokens(2)[1].strip(".").upper()
        return v.startswith("T")
    l2d, lhb = flag(), flag()
    if l2d or lhb:
        r.tokens(3)
    if lhb:
        r.tokens(2)
    if flag():
        r.tokens(4)
    if flag():
        r.tokens(4)
    if flag():
        r.tokens(4)
    if flag():
        r.tokens(2)
        r.tokens(5)
    out["filname"] = r.tokens(2)[1]
    return out


def parse_fld(text: str):
    """Force field (src/sysdef.f90:92-171): molecule types with their sites, rigid geometry,
    surface-group constraint lengths and angle parameters, the crystal block, and the vdW
    table keyed by pair number."""
    r = _ListInput(text)
    r.tokens(1)
    ntype = r.ints(1)[0]
    mols = []
    for it in range(1, ntype + 1):
        r.tokens(1)
        t = r.tokens(4)
        if int(t[0]) != it:
            raise ValueError("wrong molecule number %s" % t[0])
        m = dict(name=t[1], kind=t[2], count=int(t[3]), sites=[], geometry=None, lengths=[], angles=[])
        r.tokens(1)
        nst = r.ints(1)[0]
        for _ in range(nst):
            nm, nss = r.tokens(2)
            mass, q = r.reals(2)
            m["sites"].append((nm, int(nss), mass, q))
        ns = sum(s[1] for s in m["sites"])
        if m["kind"] == "rigmol":
            r.tokens(1)
            m["geometry"] = np.array([r.reals(3) for _ in range(ns)])
        elif m["kind"] == "surfmol":
            r.tokens(1)
            m["lengths"] = [r.reals(1)[0] for _ in range(r.ints(1)[0])]
            r.tokens(1)
            m["angles"] = [tuple(r.reals(2)) for _ in range(r.ints(1)[0])]
        elif m["kind"] == "crystal":
            m["nsites"] = r.ints(1)[0]
            m["cell"] = tuple(r.reals(3))
            m["lw"] = r.reals(1)[0]
            m["qshift"] = [r.reals(1)[0] for _ in range(r.ints(1)[0])]
        elif m["kind"] != "atmol":
            raise ValueError("unknown molecule type %s" % m["kind"])
        mols.append(m)
    r.tokens(1)
    ntx = sum(len(m["sites"]) for m in mols)
    vdw = {}
    for _ in range(ntx * (ntx + 1) // 2):
        ijt = r.ints(1)[0]
        r.tokens(1)
        r.tokens(1)
        kind = r.tokens(1)[0]
        if kind == "zero":
            vdw[ijt] = ("zero", ())
        elif kind == "LeJo":
            vdw[ijt] = ("LeJo", tuple(r.reals(2)))
        elif kind == "12-6":
            # sigma, eps given as the 12-6 coefficients (src/sysdef.f90:149-156)
            a, b = r.reals(2)
            vdw[ijt] = ("LeJo", ((a / b) ** _f32w(1.0 / 6.0), b * b / (4.0 * a)))
        elif kind == "Buck":
            vdw[ijt] = ("Buck", tuple(r.reals(3)))
        else:
            raise ValueError("unknown interaction type %s" % kind)
    return mols, vdw


def parse_top(text: str, mols):
    """Topology (src/sysdef.f90:174-216): per surface-group molecule its constrained site
    pairs and angle triples (1-based site numbers)."""
    r = _ListInput(text)
    r.skip()
    cons, angs = [], []
    for it, m in enumerate(mols, start=1):
        t = r.tokens(4)
        if int(t[0]) != it or t[2] != m["kind"] or int(t[3]) != m["count"]:
            raise ValueError("topology does not match the force field at molecule type %d" % it)
        if m["kind"] == "rigmol":
            r.tokens(2)
        elif m["kind"] == "atmol":
            r.tokens(1)
        elif m["kind"] == "surfmol":
            nsx = int(r.tokens(2)[1])
            for _ in range(m["count"]):
                cons.append([(tuple(r.ints(2)), m["lengths"][k]) for k in range(nsx)])
            nsx = int(r.tokens(2)[1])
            for _ in range(m["count"] if nsx > 0 else 0):
                angs.append([(tuple(r.ints(3)), m["angles"][k], it) for k in range(nsx)])
    return cons, angs


def parse_cfg(text: str, mols):
    """Configuration (src/config.f90:174-242,395-481): box, rigid molecules as centre of mass
    + quaternion, atom and surface-site positions, crystal sites with their types."""
    r = _ListInput(text)
    r.tokens(1)
    box = tuple(r.reals(3))
    r.skip()
    r.tokens(1)
    if r.ints(1)[0] != len(mols):
        raise ValueError("wrong number of types")
    r.skip()
    r.tokens(1)
    out = dict(box=box, com=[], quat=[], atoms={}, crystal={})
    site = 0
    for m in mols:
        t = r.tokens(3)
        if t[1] != m["kind"] or int(t[2]) != m["count"]:
            raise ValueError("configuration does not match the force field at %s" % t[0])
        ns = sum(s[1] for s in m["sites"])
        if m["kind"] == "rigmol":
            for _ in range(m["count"]):
                r.tokens(1 + ns)
                out["com"].append(r.reals(3))
                for _ in range(4):
                    r.reals(3)
                out["quat"].append(r.reals(4))
                for _ in range(4):
                    r.reals(4)
                for _ in range(5):
                    r.reals(3)
            site += m["count"] * ns
        elif m["kind"] in ("atmol", "surfmol"):
            nrec = 5 if m["kind"] == "atmol" else 2
            for _ in range(m["count"] * ns):
                i, _t = r.ints(2)
                out["atoms"][i] = r.reals(3)
                for _ in range(nrec - 1):
                    r.reals(3)
            site += m["count"] * ns
        else:
            for _ in range(m["count"] * m["nsites"]):
                i, ty = r.ints(2)
                out["crystal"][i] = (ty, r.reals(3))
        r.skip()
    return out


def load_reference_inputs(fld: str, top: str, cfg: str, inp: str, name: str = "reference"):
    """The reference's input texts -> TopologyData, deriving the tables the way sysdef_init
    and config_init do (src/sysdef.f90:356-457, src/config.f90:55-95,303-329)."""
    from .topology import TopologyData
    run = parse_inp(inp)
    mols, vdw = parse_fld(fld)
    cons, angs = parse_top(top, mols)
    cf = parse_cfg(cfg, mols)
    kinds = [m["kind"] for m in mols]
    if kinds[-1] != "crystal":
        raise ValueError("the last molecule type must be the crystal")
    nltype = max([it for it, k in enumerate(kinds, 1) if k in ("rigmol", "atmol")], default=0)
    nstype = max([it for it, k in enumerate(kinds, 1) if k == "surfmol"], default=0)
    cry = mols[-1]
    nit = sum(len(m["sites"]) for m in mols)
    nijt = nit * (nit + 1) // 2
    # canonical pair numbering and the internal vdW forms
    ijtype = np.zeros((nit, nit), np.int32)
    k = 0
    for it in range(nit):
        for jt in range(it, nit):
            k += 1
            ijtype[it, jt] = ijtype[jt, it] = k
    kind = np.zeros(nijt, np.int32)
    eps4, sig6, A, rhoi, C = (np.zeros(nijt) for _ in range(5))
    for ijt in range(1, nijt + 1):
        kd, p = vdw.get(ijt, ("zero", ()))
        if kd == "LeJo":
            kind[ijt - 1] = 1
            eps4[ijt - 1] = 4.0 * p[1]
            x = p[0]
            sig6[ijt - 1] = x * x * x * x * x * x   # the reference's sig**6, left to right
        elif kd == "Buck":
            kind[ijt - 1] = 2
            A[ijt - 1], C[ijt - 1] = p[0], p[2]
            rhoi[ijt - 1] = 1.0 / p[1] if p[1] != 0.0 else 0.0
    # sites of the moving molecules, type by type (src/sysdef.f90:383-404)
    itype, q, nsi = [], [], np.zeros(nit, np.int32)
    jjt = 0
    for it in range(1, max(nltype, nstype) + 1):
        m = mols[it - 1]
        for _ in range(m["count"]):
            for ist, (sn, nss, mass, qq) in enumerate(m["sites"], start=1):
                itype += [jjt + ist] * nss
                q += [qq] * nss
        for ist in range(1, len(m["sites"]) + 1):
            if m["count"] == 0:
                continue   # set inside the reference's per-molecule loop: stays 0 without molecules
            if m["kind"] == "rigmol":
                nsi[jjt + ist - 1] = sum(s[1] for s in m["sites"]) if ist == 1 else 0
            else:
                nsi[jjt + ist - 1] = 1
        jjt += len(m["sites"])
    nmove = len(itype)
    nriga = sum(m["count"] * sum(s[1] for s in m["sites"]) for m in mols if m["kind"] == "rigmol")
    nliqa = sum(m["count"] * sum(s[1] for s in m["sites"]) for it, m in enumerate(mols, 1) if it <= nltype)
    nrig = sum(m["count"] for m in mols if m["kind"] == "rigmol")
    # crystal sites: type 1/2, or a charge-shift entry when negative (src/sysdef.f90:406-419)
    nss1 = cry["sites"][0][1]
    for i in range(nmove + 1, nmove + cry["nsites"] + 1):
        ty, _ = cf["crystal"][i]
        if ty > 0:
            q.append(cry["sites"][ty - 1][3])
            is_ = ty
        else:
            q.append(cry["qshift"][-ty - 1])
            is_ = 1 if -ty <= nss1 else 2
        itype.append(jjt + is_)
    ntota = len(itype)
    pos = np.zeros((ntota, 3))
    # rigid sites from COM + quaternion with body coordinates about the mass-weighted centre
    # (src/sysdef.f90:421-441, src/config.f90:303-329)
    com = np.array(cf["com"], np.float64).reshape(-1, 3)
    if nrig:
        w = mols[0]
        geo = w["geometry"]
        v_unit = _f32w(1.0e-10) / (run["dt"] * _f32w(1.0e-15))
        mi = [s[2] * v_unit ** 2 / 1000.0 for s in w["sites"] for _ in range(s[1])]
        body = np.empty_like(geo)
        for a in range(3):
            num = 0.0
            den = 0.0
            for j in range(len(mi)):
                num += geo[j, a] * mi[j]
                den += mi[j]
            body[:, a] = geo[:, a] - num / den
        qt = np.array(cf["quat"], np.float64)
        q1, q2, q3, q4 = (qt[:, k] for k in range(4))
        a11 = -q1 * q1 + q2 * q2 - q3 * q3 + q4 * q4
        a12 = 2.0 * (q3 * q4 - q1 * q2)
        a13 = 2.0 * (q2 * q3 + q1 * q4)
        a21 = -2.0 * (q1 * q2 + q3 * q4)
        a22 = q1 * q1 - q2 * q2 - q3 * q3 + q4 * q4
        a23 = 2.0 * (q2 * q4 - q1 * q3)
        a31 = 2.0 * (q2 * q3 - q1 * q4)
        a32 = -2.0 * (q1 * q3 + q4 * q2)
        a33 = -q1 * q1 - q2 * q2 + q3 * q3 + q4 * q4
        for j in range(3):
            pos[j:nriga:3, 0] = com[:, 0] + a11 * body[j, 0] + a21 * body[j, 1] + a31 * body[j, 2]
            pos[j:nriga:3, 1] = com[:, 1] + a12 * body[j, 0] + a22 * body[j, 1] + a32 * body[j, 2]
            pos[j:nriga:3, 2] = com[:, 2] + a13 * body[j, 0] + a23 * body[j, 1] + a33 * body[j, 2]
    for i, p in cf["atoms"].items():
        pos[i - 1] = p
    for i, (_, p) in cf["crystal"].items():
        pos[i - 1] = p
    constraints = None
    if cons:
        flat = [c for mol in cons for c in mol]
        constraints = (np.array([c[0][0] - 1 for c in flat], np.int32), np.array([c[0][1] - 1 for c in flat], np.int32),
                       np.array([c[1] for c in flat], np.float64))
    angles = angle_type = None
    if angs:
        flat = [a for mol in angs for a in mol]
        pi_w = _f32w(3.1415926535898)
        angles = (np.array([a[0][0] - 1 for a in flat], np.int32), np.array([a[0][1] - 1 for a in flat], np.int32),
                  np.array([a[0][2] - 1 for a in flat], np.int32), np.array([a[1][0] for a in flat], np.float64),
                  np.array([a[1][1] * pi_w / 180.0 for a in flat], np.float64))
        angle_type = np.array([a[2] for a in flat], np.int32)
    return TopologyData(name=name, box=cf["box"], rcut=run["rcut"], skin=run["skin"], kmax=run["kmax"],
                        site_type=np.array(itype, np.int32), charge=np.array(q, np.float64), nsi=nsi,
                        ijtype=ijtype, vdw_kind=kind, vdw_eps4=eps4, vdw_sig6=sig6, vdw_a=A, vdw_rhoi=rhoi,
                        vdw_c=C, n_move=nmove, n_rigid_sites=nriga, n_liquid_sites=nliqa, n_rig=nrig,
                        n_ltype=nltype, wall_lw=cry["lw"], wall_cell=cry["cell"],
                        wall_nss=(cry["sites"][0][1], cry["sites"][1][1]), pos=pos, com=com,
                        constraints=constraints, angles=angles, angle_type=angle_type, dt=run["dt"],
                        meta=dict(run=run))


def load_reference_files(directory: str, base: str = None, inp: str = "tot.inp"):
    """Read tot.inp (or the named run file) and the .fld/.top/.cfg it names."""
    run_text = open(os.path.join(directory, inp)).read()
    if base is None:
        base = parse_inp(run_text)["filname"]
    rd = lambda ext: open(os.path.join(directory, base + ext)).read()
    return load_reference_inputs(rd(".fld"), rd(".top"), rd(".cfg"), run_text, name=base)
