"""NVE energy drift, end to end: the reference's own time loop (Gear predictor-corrector for
COMs and quaternions, src/move.f90, thermostat off) driving the engine through its drop-in
Fortran force module (md_sim_amd/fortran/force_mdf.f90), against the same loop driving the
CPU oracle through that module and the reference's own force module (tests/golden/nve_pore.npz,
made by tests/golden/make_nve.py).

The drivers are built from the reference's sources into oracle/_ref by oracle/build_ref.sh;
the test is skipped where that build is absent."""
import dataclasses
import os
import sys
import time

import numpy as np
import pytest

from md_sim_amd import systems
from helpers import golden

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "oracle", "_ref", "md_nve_gpu")


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _drift(traj):
    from golden.make_nve import drift
    return drift(traj)


@pytest.fixture(scope="module")
def runs():
    if not os.access(EXE, os.X_OK):
        pytest.skip("NVE driver with the engine's force module not built (oracle/build_ref.sh)")
    from oracle import refrun
    g = golden("nve_pore")
    s = dataclasses.replace(systems.config("pore_water"), com=g["com"], quat=g["quat"], ion_pos=g["ion_pos"],
                            com_vel=g["com_vel"], ang_vel=g["ang_vel"], ion_vel=g["ion_vel"], dt=float(g["dt"]))
    n = int(g["nsteps"])
    _say("case nve pore_water fp64, %d steps" % n)
    t64 = refrun.run_nve(s, n, variant="gpu", env={"MDF_PRECISION": "fp64"}, timeout=300)
    _say("case nve pore_water fp32, %d steps" % n)
    t32 = refrun.run_nve(s, n, variant="gpu", env={"MDF_PRECISION": "fp32"}, timeout=300)
    return g["traj"], g["traj_oracle"], t64, t32


def test_engine_follows_oracle_trajectory(runs):
    """The same integrator on the engine and on the oracle: the trajectories coincide to
    the force path's tolerance for the first 1000 steps (1 ps), before chaos separates
    them."""
    _, orc, t64, _ = runs
    assert t64.shape == orc.shape
    n = 1000
    assert np.abs(t64[:n, 2] - orc[:n, 2]).max() < 1e-9 * np.abs(orc[:n, 2]).max()
    assert np.abs(t64[:n, 1] - orc[:n, 1]).max() < 1e-9 * np.abs(orc[:n, 1]).max()


def test_reference_agrees_until_its_stale_offset(runs):
    """Step 1 evaluates identical positions in the reference and the engine.  From step 2
    the reference's force_sum has overwritten molecule 1's lead offset (rsx(1),
    src/forcem.f90:1006-1008), which the engine does not replicate."""
    ref, orc, t64, _ = runs
    assert abs(t64[0, 2] - ref[0, 2]) < 1e-9 * abs(ref[0, 2])
    assert abs(t64[1, 1] - ref[1, 1]) < 1e-9 * abs(ref[1, 1])


def test_fp64_drift_within_reference(runs):
    ref, orc, t64, _ = runs
    d_ref, d_orc, d64 = _drift(ref), _drift(orc), _drift(t64)
    _say("  htot drift over the run: reference %.3f, oracle %.3f, engine FP64 %.3f kJ/mol" % (d_ref, d_orc, d64))
    assert np.all(np.isfinite(t64[:, 1]))
    assert abs(d64) <= 2.0 * abs(d_ref) + 1.0


def test_fp32_drift_within_twice_fp64(runs):
    _, _, t64, t32 = runs
    d64, d32 = _drift(t64), _drift(t32)
    _say("  htot drift over the run: engine FP64 %.3f, FP32 %.3f kJ/mol" % (d64, d32))
    assert np.all(np.isfinite(t32[:, 1]))
    assert abs(d32) <= 2.0 * abs(d64) + 1.0


# ---- the 16k-water interface workload (slab16k: 16384 SPC/E waters on 6144 frozen ions) ----
N16K = int(os.environ.get("MDF_NVE16K_STEPS", "10000"))


@pytest.fixture(scope="module")
def start16k():
    """The slab16k workload's sizes (16384 SPC/E waters on a 16x16x6 rock-salt slab, rc 10,
    k 11) drawn with seed 1006, relaxed by rigid-body steepest descent on the engine's FP64
    forces, with the 1 fs step of the small-system drift test.

    The benchmark's draw (seed 1002) gives the O-Cl Buckingham term A = 6.5e4 kJ/mol (the
    low end of the 1e4-1e5 kcal/mol range): its barrier sits near 1 A, and behind it the
    H-Cl Coulomb term has no repulsion at all, so waters fall onto the anions within tens of
    steps (the reference's own loop on the oracle heats from 150 to 1000 K in 12 steps from
    the lattice start; a relaxed start lasts about 30 steps).  Seed 1006 draws A = 3.7e5 and
    2.6e5 kJ/mol, whose barriers hold."""
    if not os.access(EXE, os.X_OK):
        pytest.skip("NVE driver with the engine's force module not built (oracle/build_ref.sh)")
    from golden.make_nve import relax
    from md_sim_amd.engine import ForceEngine
    s0 = systems.water_on_slab(16384, 16, 6, seed=1006, rcut=10.0, kmax=11, lw=80.0, name="slab16k_md")
    eng = ForceEngine(s0)
    _say("case nve slab16k_md: relaxing on the engine")
    t0 = time.time()
    s, e = relax(s0, n_iter=400, compute=eng.force_get)
    eng.close()
    _say("  relaxed e_conf + e_wall_lr %.1f kJ/mol in %.0f s" % (e, time.time() - t0))
    return dataclasses.replace(s, dt=1.0)


@pytest.fixture(scope="module")
def run16k_fp64(start16k):
    from oracle import refrun
    _say("case nve slab16k_md fp64, %d steps" % N16K)
    return refrun.run_nve(start16k, N16K, variant="gpu", env={"MDF_PRECISION": "fp64"}, timeout=900)


def test_interface_fp64_run_conserves_energy(run16k_fp64):
    """10k steps (10 ps) of the reference's Gear loop at 16k waters on the FP64 engine: every
    step finite, and the htot drift small against the kinetic energy's fluctuation."""
    t = run16k_fp64
    assert t.shape[0] == N16K and np.all(np.isfinite(t[:, 1]))
    d = _drift(t)
    _say("  slab16k_md FP64: T %.1f -> %.1f K, htot drift %.3f kJ/mol, ke std %.3f" % (t[1, 4], t[-1, 4], d,
                                                                                  t[1:, 3].std()))
    assert abs(d) < t[1:, 3].std()


def test_interface_fp32_drift_within_twice_fp64(start16k, run16k_fp64):
    """Config 2's precision check at its size: FP32 pair and structure-factor arithmetic
    (FP64 accumulation, FP64 integrator) drifts within twice the FP64 drift over the same
    10k steps from the same start."""
    from oracle import refrun
    _say("case nve slab16k_md fp32, %d steps" % N16K)
    t32 = refrun.run_nve(start16k, N16K, variant="gpu", env={"MDF_PRECISION": "fp32"}, timeout=900)
    d64, d32 = _drift(run16k_fp64), _drift(t32)
    _say("  slab16k_md htot drift: FP64 %.3f, FP32 %.3f kJ/mol; FP32 T %.1f -> %.1f K" % (d64, d32, t32[1, 4],
                                                                                       t32[-1, 4]))
    assert np.all(np.isfinite(t32[:, 1]))
    assert abs(d32) <= 2.0 * abs(d64) + 1.0
