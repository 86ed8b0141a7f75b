"""Write tests/golden/nve_pore.npz: a relaxed start state of the pore_water system (96 SPC/E
waters over a frozen rock-salt slab) and the reference's own NVE trajectory from it.

1. Relax: rigid-body steepest descent of the seeded configuration on the CPU oracle's
   energy (COMs along the COM forces, quaternions along the energy gradient obtained from
   the site forces, mobile ions along their forces), so the NVE run starts near a minimum
   instead of converting lattice strain into heat.
2. Run the reference's time loop with its own force module (oracle/_ref/md_nve_slab,
   built from the reference's sources by oracle/build_ref.sh: Gear predictor-corrector,
   quaternions, thermostat off) for NSTEPS steps of DT fs and store htot, uconf, ke and the
   temperature per step.

3. Run the same loop with the engine's drop-in force module bound to the CPU oracle
   (oracle/_ref/md_nve_oracle): the reference's dynamics without its stale molecule-1
   offset, the trajectory the engine must follow.

The GPU test runs the loop with the drop-in module on the engine from the stored start
state and compares trajectories and energy conservation.

Usage:  python tests/golden/make_nve.py
"""
import dataclasses
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from md_sim_amd import abi, systems  # noqa: E402
from oracle import pyoracle, refrun  # noqa: E402

NSTEPS = 2000
DT = 1.0


def _site_jacobian(s, eps=1e-7):
    """d(site positions)/d(quaternion component): (nrig, 4, 3 sites, 3)."""
    base = s.site_positions()[:s.nriga].reshape(s.nrig, 3, 3)
    out = np.empty((s.nrig, 4, 3, 3))
    q0 = s.quat.copy()
    for k in range(4):
        q = q0.copy()
        q[:, k] += eps
        p = dataclasses.replace(s, quat=q).site_positions()[:s.nriga].reshape(s.nrig, 3, 3)
        out[:, k] = (p - base) / eps
    return out


def relax(s, n_iter=600, compute=None):
    """compute(x, y, z, cx, cy, cz) -> force results; the CPU oracle's by default."""
    if compute is None:
        compute = pyoracle.Oracle(s).compute

    def energy(sys_):
        r = compute(*abi.positions(sys_))
        return r["e_conf"] + r["e_wall_lr"], r

    e, r = energy(s)
    step = 0.02   # max site displacement per iteration (A)
    for it in range(n_iter):
        f = r["f_site"]                                   # (3, nmove)
        g_com = -r["f_com"].T                             # (nrig, 3)
        J = _site_jacobian(s)
        fs = f[:, :s.nriga].T.reshape(s.nrig, 3, 3)      # (nrig, site, xyz)
        g_q = -np.einsum("mksx,msx->mk", J, fs)
        g_ion = -f[:, s.nriga:s.nmove].T
        # largest site displacement of a unit step
        disp = max(np.abs(g_com).max(initial=0.0), 2.0 * np.abs(g_q).max(initial=0.0),
                   np.abs(g_ion).max(initial=0.0))
        if disp == 0.0:
            break
        h = step / disp
        q = s.quat - h * g_q
        q /= np.linalg.norm(q, axis=1, keepdims=True)
        trial = dataclasses.replace(s, com=s.com - h * g_com, quat=q, ion_pos=s.ion_pos - h * g_ion)
        et, rt = energy(trial)
        if et < e:
            s, e, r = trial, et, rt
            step = min(step * 1.2, 0.1)
        else:
            step *= 0.5
            if step < 1e-6:
                break
    return s, e


def drift(traj):
    """htot change over a run: mean of the last tenth minus mean of the first tenth (the
    first step is skipped: its htot pairs the kinetic energy with no potential energy)."""
    h = traj[1:, 1]
    m = max(1, len(h) // 10)
    return float(h[-m:].mean() - h[:m].mean())


def main():
    s0 = systems.config("pore_water")
    e0 = pyoracle.Oracle(s0).compute_system()
    s, e = relax(s0)
    print("relaxed: e_conf + e_wall_lr %.3f -> %.3f" % (e0["e_conf"] + e0["e_wall_lr"], e))
    s = dataclasses.replace(s, dt=DT)
    traj = refrun.run_nve(s, NSTEPS, variant="slab", timeout=1200)
    print("reference NVE: T %.1f -> %.1f K, htot %.4f -> %.4f" % (traj[0, 4], traj[-1, 4], traj[1, 1], traj[-1, 1]))
    print("drift (last 10%% - first 10%% of htot): %.4f kJ/mol" % drift(traj))
    # the same loop on the oracle through the engine's drop-in force module: the reference's
    # dynamics without its stale rsx(1) offset (src/forcem.f90:1006-1008)
    traj_o = refrun.run_nve(s, NSTEPS, variant="oracle", timeout=1200)
    print("oracle NVE: drift %.4f kJ/mol, step-2 uconf %.6f (reference %.6f)" % (drift(traj_o), traj_o[1, 2],
                                                                                  traj[1, 2]))
    path = os.path.join(HERE, "nve_pore.npz")
    np.savez_compressed(path, com=s.com, quat=s.quat, ion_pos=s.ion_pos, com_vel=s.com_vel,
                        ang_vel=s.ang_vel, ion_vel=s.ion_vel, dt=np.array(DT), nsteps=np.array(NSTEPS),
                        traj=traj, traj_oracle=traj_o)
    print("wrote", path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
