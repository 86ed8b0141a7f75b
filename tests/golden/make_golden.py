"""Write the golden fixtures in this directory from the reference itself.

Runs the reference force path (compiled from the reference's Fortran sources by
oracle/build_ref.sh into oracle/_ref, development container only) on seeded systems and
stores its inputs and outputs:

* bulk216.npz    - 216 SPC/E waters, 3-D periodic variant (src/forcem_back.f90);
* pore_small.npz - 96 waters + 8 mobile ions over a 4-layer rock-salt slab, slab variant
                   (src/forcem.f90) with slab correction and wall tail correction;
* slab2048.npz   - 2048 waters on a 4608-ion slab (the 16k workload's smaller sibling):
                   energies, pair-list checksums and every 7th site force.

Usage:  python tests/golden/make_golden.py
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from md_sim_amd import systems  # noqa: E402
from oracle.refrun import run_reference  # noqa: E402

SCALARS = ("uewa", "uself", "urf", "ulr", "uconf", "press", "ewqWW", "ewqLL", "ewqWL", "uavex",
           "alfa", "facq", "spii", "rskinsq")


def pair_digest(i, j):
    """Order-sensitive digest of a 1-based pair list."""
    h = hashlib.sha256()
    h.update(np.asarray(i, np.int64).tobytes())
    h.update(np.asarray(j, np.int64).tobytes())
    return h.hexdigest()


def main():
    for name in ("bulk216", "pore_small", "slab2048"):
        s = systems.config(name)
        d = run_reference(s)
        out = {k: d[k] for k in SCALARS}
        for k in ("elj", "eqq", "fwsx", "fwsy", "fwsz", "kmaxes", "counts"):
            out[k] = d[k]
        out["list_digest"] = np.array(pair_digest(d["list1"], d["list2"]))
        out["wlist_digest"] = np.array(pair_digest(d["listw1"], d["listw2"]))
        if name == "slab2048":
            sel = np.arange(0, s.nmove, 7)
            out["sel"] = sel
            for k in ("f_site_x", "f_site_y", "f_site_z", "f_real_x", "f_real_y", "f_real_z"):
                out[k] = d[k][sel]
            msel = np.arange(0, s.nrig, 7)
            out["msel"] = msel
            for k in ("fcx", "fcy", "fcz", "tx", "ty", "tz"):
                out[k] = d[k][msel]
            out["pos_digest"] = np.array(hashlib.sha256(np.concatenate([d["x0"], d["y0"], d["z0"]]).tobytes()).hexdigest())
            kk = np.arange(0, len(d["k_ak"]), 11)
            out["ksel"] = kk
            for k in ("k_ak", "k_sre", "k_sim"):
                out[k] = d[k][kk]
            out["k_vec"] = d["k_vec"].reshape(-1, 3)[kk]
        else:
            for k in ("x0", "y0", "z0", "qs", "itype", "cx0", "cy0", "cz0", "list1", "list2", "listw1", "listw2",
                      "f_real_x", "f_real_y", "f_real_z", "f_site_x", "f_site_y", "f_site_z",
                      "fcx", "fcy", "fcz", "tx", "ty", "tz", "k_ak", "k_sre", "k_sim", "k_vec"):
                out[k] = d[k]
            for k in ("list1", "list2", "listw1", "listw2", "itype"):
                out[k] = out[k].astype(np.int32)
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, **out)
        print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
