"""Build the force engine's HIP library (libmdforce.so, gfx950) in-tree with hipcc."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libmdforce.so")
SOURCES = ["nlist.hip", "pair.hip", "recip.hip", "recip64.hip", "epilogue.hip", "surrogate.hip", "chempot.hip", "api.hip"]
ARCH = os.environ.get("MDF_OFFLOAD_ARCH", "gfx950")
FLAGS = ["-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
         # no packed FP32 instructions anywhere in the device code
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         # matrix-core accumulators in ordinary VGPRs: on gfx950 v_mfma_f64_16x16x4_f64 with its
         # accumulators in AGPRs was measured to sustain ~45 TFLOP/s, in VGPRs ~75 (of 78.6)
         "-mllvm", "-amdgpu-mfma-vgpr-form",
         "-Wno-unused-result"]


def build(verbose: bool = False, jobs: int = 5) -> str:
    """Compile every source (in parallel) and link the shared library.  Always rebuilds."""
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    procs = []
    objs = []
    for src in SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        objs.append(obj)
        cmd = [hipcc, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        if len([p for p in procs if p[1].poll() is None]) >= jobs:
            procs[0][1].wait()
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed.append((src, out))
        elif verbose and out.strip():
            print(out)
    if failed:
        raise RuntimeError("\n".join("%s:\n%s" % f for f in failed))
    return link(objs, LIB, objdir)


def link(objs, out: str, objdir: str) -> str:
    """Link HIP objects into a shared library against a symbol-less stub whose soname is the
    generic "libamdhip64.so": the library then needs the HIP runtime under that name, found
    through the rpath in /opt/rocm/lib when the process has none, and shared with the runtime
    a host process (PyTorch-ROCm bundles one under the same name) has already loaded - one
    HIP runtime per process instead of two."""
    stubdir = os.path.join(objdir, "stub")
    os.makedirs(stubdir, exist_ok=True)
    stub_src = os.path.join(stubdir, "stub.c")
    with open(stub_src, "w") as f:
        f.write("/* link-time stand-in carrying the soname of the HIP runtime */\n")
    subprocess.check_call(["gcc", "-shared", "-fPIC", "-Wl,-soname,libamdhip64.so", "-o",
                           os.path.join(stubdir, "libamdhip64.so"), stub_src])
    rocm_lib = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib")
    tmp = out + ".tmp"
    subprocess.check_call(["g++", "-shared", "-fPIC", "-o", tmp, *objs, "-L" + stubdir, "-Wl,--no-as-needed", "-lamdhip64", "-Wl,--as-needed",
                           "-Wl,--disable-new-dtags", "-Wl,-rpath," + rocm_lib])
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
