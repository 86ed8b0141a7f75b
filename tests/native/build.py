"""Build the test-only HIP probe library (tests/native/libmdf_math_probe.so, gfx950) with the
engine's own compiler flags and runtime link."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
LIB = os.path.join(HERE, "libmdf_math_probe.so")


def build() -> str:
    if ROOT not in sys.path:
        sys.path.insert(0, ROOT)
    from md_sim_amd import build as hipbuild
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    obj = os.path.join(objdir, "math_probe.o")
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    r = subprocess.run([hipcc, *hipbuild.FLAGS, "-c", os.path.join(HERE, "math_probe.hip"), "-o", obj],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("math_probe.hip:\n" + r.stdout)
    return hipbuild.link([obj], LIB, objdir)


if __name__ == "__main__":
    print(build())
