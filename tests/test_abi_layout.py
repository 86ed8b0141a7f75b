"""The ctypes mirrors of include/mdforce.h's structs have the C layout: size and every field's
offset, read from a probe program that gcc compiles against the header (CPU only)."""
import ctypes as C
import os
import shutil
import subprocess
import tempfile

import pytest

from md_sim_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
STRUCTS = (("mdf_topology", abi.Topology), ("mdf_params", abi.Params), ("mdf_outputs", abi.Outputs),
           ("mdf_mts_params", abi.MtsParams), ("mdf_mts_stats", abi.MtsStats),
           ("mdf_chempot_params", abi.ChempotParams))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="no C compiler")
def test_struct_layouts_match_header():
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "mdforce.h"', "int main(void) {"]
    for cname, py in STRUCTS:
        lines.append('printf("%s %%zu\\n", sizeof(%s));' % (cname, cname))
        for f, _ in py._fields_:
            lines.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (cname, f, cname, f))
    lines.append("return 0; }")
    d = tempfile.mkdtemp(prefix="abi_")
    try:
        src, exe = os.path.join(d, "probe.c"), os.path.join(d, "probe")
        open(src, "w").write("\n".join(lines) + "\n")
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), src, "-o", exe], check=True)
        out = dict(l.split() for l in subprocess.run([exe], capture_output=True, text=True,
                                                     check=True).stdout.splitlines())
    finally:
        shutil.rmtree(d, ignore_errors=True)
    for cname, py in STRUCTS:
        assert int(out[cname]) == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert int(out[cname + "." + f]) == getattr(py, f).offset, (cname, f)
