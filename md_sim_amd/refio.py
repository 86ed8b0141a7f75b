"""Readers and writers for the reference's text formats.

* force field ``.fld`` - read by sysdef_input (src/sysdef.f90:92-171);
* topology ``.top``    - read by sysdef_input (src/sysdef.f90:174-216);
* configuration ``.cfg`` - read by config_input (src/config.f90:174-242, record layouts
  src/config.f90:395-481).

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


def write_reference_inputs(s: System, directory: str, base: str = "sys") -> Dict[str, str]:
    os.makedirs(directory, exist_ok=True)
    p = {k: os.path.join(directory, base + "." + k) for k in ("fld", "top", "cfg")}
    write_fld(s, p["fld"])
    write_top(s, p["top"])
    write_cfg(s, p["cfg"])
    p["inp"] = os.path.join(directory, "harness.inp")
    write_harness_input(s, p["inp"], base)
    return p
