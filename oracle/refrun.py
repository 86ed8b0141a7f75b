"""Run the reference golden-vector harness (oracle/_ref, built by oracle/build_ref.sh) on a
System and parse its dump.  Test infrastructure, never part of the product path: it makes
and checks golden vectors in the development container, and bench.py's cpu_baseline leg
times it (the harness binaries built here travel with the tree; the reference sources do
not) on the host cores wherever it runs, the GPU box included."""
from __future__ import annotations

import os
import re
import shutil
import subprocess
import tempfile
from typing import Dict

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")


def ref_available(fast: bool = False) -> bool:
    sfx = "_fast" if fast else ""
    return all(os.access(os.path.join(REF_DIR, "ref_" + v + sfx), os.X_OK) for v in ("slab", "bulk"))


def parse_dump(path: str) -> Dict[str, np.ndarray]:
    raw = open(path, "rb").read()
    out: Dict[str, np.ndarray] = {}
    pos = 0
    while pos < len(raw):
        name = raw[pos:pos + 16].decode().strip()
        kind, cnt = np.frombuffer(raw, dtype=np.int64, count=2, offset=pos + 16)
        pos += 32
        dt = np.float64 if kind == 1 else np.int64
        out[name] = np.frombuffer(raw, dtype=dt, count=int(cnt), offset=pos).copy()
        pos += 8 * int(cnt)
    return out


MPIEXEC = os.environ.get("MDF_MPIEXEC", "/opt/conda/bin/mpiexec")


def mpi_available() -> bool:
    return os.access(MPIEXEC, os.X_OK)


def run_reference(system, workdir: str = None, timeout: float = 600.0, nrep: int = 0,
                  nproc: int = 1, env=None, fast: bool = False) -> Dict[str, np.ndarray]:
    """Write the reference's input files for `system`, run the matching harness variant
    (bulk = 3-D periodic, slab otherwise) and return the parsed dump.  nproc > 1 runs the
    harness on that many MPI ranks (the reference's own decomposition of the force path)
    for timing only: the result then holds the timing and no dump.  env adds environment
    variables (REF_CHEMPOT="ichmax lwmin lwmax lchmax ilmid ncall" runs chempot ncall times
    after the step and dumps chem_pave, chem_inave, chem_seed and chem_par).  fast runs the
    timing build with the reference's own compiler flags (-O2 -ffast-math)."""
    from md_sim_amd.refio import write_reference_inputs

    variant = "bulk" if system.z_periodic else "slab"
    exe = os.path.join(REF_DIR, "ref_" + variant + ("_fast" if fast else ""))
    if not os.access(exe, os.X_OK):
        raise FileNotFoundError(exe)
    if nproc > 1 and nrep <= 0:
        raise ValueError("several ranks only time: give nrep")
    own = workdir is None
    d = tempfile.mkdtemp(prefix="mdref_") if own else workdir
    write_reference_inputs(system, d)
    out = os.path.join(d, "dump.bin")
    args = [exe, out] + ([str(nrep)] if nrep > 0 else [])
    if nproc > 1:
        args = [MPIEXEC, "-n", str(int(nproc))] + args
    e = dict(os.environ)
    if env:
        e.update(env)
    r = subprocess.run(args, cwd=d, capture_output=True, text=True, timeout=timeout, env=e)
    if r.returncode != 0 or (nproc == 1 and not os.path.exists(out)):
        raise RuntimeError("reference harness failed (%d):\n%s\n%s" % (r.returncode, r.stdout[-2000:], r.stderr[-2000:]))
    dump = parse_dump(out) if nproc == 1 else {}
    dump["_stdout"] = np.array(r.stdout)
    for line in r.stdout.splitlines():
        m = re.match(r"\s*TIMING nrep=\s*(\d+)\s+seconds=\s*(\S+)", line)
        if m:
            dump["_timing"] = np.array([float(m.group(1)), float(m.group(2))])
    if own:
        shutil.rmtree(d, ignore_errors=True)
    return dump


def parse_nve_obs(path: str) -> Dict[str, np.ndarray]:
    """The measurement arrays md_nve writes with MD_NVE_OBS: name/count headers, one value a
    line (ngc and inave as int64, the rest float64; NaN kept)."""
    out = {}
    lines = open(path).read().split("\n")
    i = 0
    while i < len(lines):
        head = lines[i].split()
        i += 1
        if not head:
            continue
        name, n = head[0], int(head[1])
        vals = [v.strip() for v in lines[i:i + n]]
        i += n
        if name in ("ngc", "inave"):
            out[name] = np.array([int(v) for v in vals], np.int64)
        else:
            out[name] = np.array([float(v) for v in vals], np.float64)
    return out


def run_nve(system, nsteps: int, variant: str = None, env=None, workdir: str = None,
            timeout: float = 600.0, obs: bool = False):
    """Run an NVE trajectory driver (oracle/md_nve.f90) on `system` for `nsteps` steps of
    system.dt: variant 'slab'/'bulk' is the reference's own force module, 'gpu' the engine's
    drop-in module.  Returns (nsteps, 6): step, htot, uconf, ke, tempnow, ulr; with obs also
    the measurement arrays after the last step (parse_nve_obs)."""
    from md_sim_amd.refio import write_reference_inputs

    if variant is None:
        variant = "bulk" if system.z_periodic else "slab"
    exe = os.path.join(REF_DIR, "md_nve_" + variant)
    if not os.access(exe, os.X_OK):
        raise FileNotFoundError(exe)
    own = workdir is None
    d = tempfile.mkdtemp(prefix="mdnve_") if own else workdir
    write_reference_inputs(system, d)
    out = os.path.join(d, "traj.txt")
    e = dict(os.environ)
    if env:
        e.update(env)
    if obs:
        e["MD_NVE_OBS"] = "obs.txt"
    r = subprocess.run([exe, out, str(int(nsteps))], cwd=d, capture_output=True, text=True, timeout=timeout,
                       env=e)
    if r.returncode != 0 or not os.path.exists(out):
        raise RuntimeError("NVE driver failed (%d):\n%s\n%s" % (r.returncode, r.stdout[-2000:], r.stderr[-2000:]))
    t = np.loadtxt(out, ndmin=2)
    o = parse_nve_obs(os.path.join(d, "obs.txt")) if obs else None
    if own:
        shutil.rmtree(d, ignore_errors=True)
    return (t, o) if obs else t
