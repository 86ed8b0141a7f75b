"""The Fortran ISO_C_BINDING path: the interface module and program build and link against
the C ABI (CPU), and a force evaluation from Fortran matches the oracle (GPU)."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from md_sim_amd import systems

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FDIR = os.path.join(ROOT, "md_sim_amd", "fortran")
PROG = os.path.join(FDIR, "mdf_force_once")


def _have_fortran():
    return subprocess.run(["which", "amdflang"], capture_output=True).returncode == 0


@pytest.mark.skipif(not _have_fortran(), reason="no Fortran compiler")
def test_fortran_interface_builds_and_links():
    subprocess.check_call(["bash", os.path.join(FDIR, "build.sh")])
    assert os.access(PROG, os.X_OK)
    out = subprocess.run(["ldd", PROG], capture_output=True, text=True).stdout
    assert "libmdforce.so" in out


@pytest.mark.gpu
def test_fortran_force_evaluation_matches_oracle(oracle_lib):
    from md_sim_amd.fortran.sysfile import read_forces, write_sysfile
    assert os.access(PROG, os.X_OK), "Fortran program not built"
    s = systems.config("pore_small")
    d = tempfile.mkdtemp()
    sf, ff = os.path.join(d, "sys.bin"), os.path.join(d, "f.bin")
    write_sysfile(s, sf)
    r = subprocess.run([PROG, sf, ff], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    got = read_forces(ff, s.nmove, s.nrig, s.nijt)
    ref = oracle_lib.Oracle(s).compute_system()
    err = np.abs(got["f_site"] - ref["f_site"]).max() / np.abs(ref["f_site"]).max()
    assert err < 1e-10, err
    assert abs(got["e_conf"] - ref["e_conf"]) < 1e-10 * abs(ref["e_conf"])
    assert np.abs(got["torque"] - ref["torque"]).max() < 1e-10 * np.abs(ref["torque"]).max()
