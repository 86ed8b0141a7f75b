"""Python handle on the CPU oracle (oracle/oracle.c) - TEST INFRASTRUCTURE ONLY.

The oracle is the checker for the HIP engine: tests, the smoke check and the CPU
baseline leg of bench.py use it; the product path never does."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from md_sim_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")


def build(force: bool = False) -> str:
    src = os.path.join(HERE, "oracle.c")
    if force or not os.path.exists(LIB) or os.path.getmtime(LIB) < os.path.getmtime(src):
        subprocess.check_call(["gcc", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-o", LIB + ".tmp",
                               src, "-lm"])
        os.replace(LIB + ".tmp", LIB)
    return LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.orc_create.restype = C.c_int
        L.orc_compute.restype = C.c_int
        L.orc_last_error.restype = C.c_char_p
        _lib = L
    return _lib


class Oracle:
    """force_init / force_get of the reference, restated on the CPU."""

    def __init__(self, system, **kw):
        self.system = system
        self.sa = abi.build_abi(system, **kw)
        x, y, z, cx, cy, cz = abi.positions(system)
        self._h = C.c_void_p()
        L = lib()
        dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
        rc = L.orc_create(C.byref(self.sa.top), C.byref(self.sa.par), dp(x), dp(y), dp(z), dp(cx), dp(cy), dp(cz),
                          C.byref(self._h))
        if rc:
            raise RuntimeError("oracle create failed: %s" % L.orc_last_error().decode())

    def compute(self, x, y, z, cx, cy, cz) -> abi.ForceResult:
        s = self.system
        o, arrs = abi.make_outputs(s.nmove, s.nrig, s.nijt)
        dp = lambda a: np.ascontiguousarray(a, np.float64).ctypes.data_as(C.POINTER(C.c_double))
        keep = [np.ascontiguousarray(a, np.float64) for a in (x, y, z, cx, cy, cz)]
        rc = lib().orc_compute(self._h, *[k.ctypes.data_as(C.POINTER(C.c_double)) for k in keep], C.byref(o))
        if rc:
            raise RuntimeError("oracle compute failed (%d): %s" % (rc, lib().orc_last_error().decode()))
        return abi.collect(o, arrs, s.nmove, s.nrig)

    def compute_system(self) -> abi.ForceResult:
        return self.compute(*abi.positions(self.system))

    def pairs(self):
        n1, n2 = C.c_int32(), C.c_int32()
        cap1 = cap2 = 1
        while True:
            a = np.zeros(cap1, np.int32); b = np.zeros(cap1, np.int32)
            c = np.zeros(cap2, np.int32); d = np.zeros(cap2, np.int32)
            pi = lambda v: v.ctypes.data_as(C.POINTER(C.c_int32))
            rc = lib().orc_get_pairs(self._h, pi(a), pi(b), cap1, C.byref(n1), pi(c), pi(d), cap2, C.byref(n2))
            if rc == 0:
                return a[:n1.value], b[:n1.value], c[:n2.value], d[:n2.value]
            cap1, cap2 = max(n1.value, 1), max(n2.value, 1)

    def kvectors(self):
        n = C.c_int32()
        lib().orc_get_kvectors(self._h, None, None, 0, C.byref(n))
        kv = np.zeros(3 * n.value, np.int32)
        ak = np.zeros(n.value)
        lib().orc_get_kvectors(self._h, kv.ctypes.data_as(C.POINTER(C.c_int32)),
                               ak.ctypes.data_as(C.POINTER(C.c_double)), n.value, C.byref(n))
        sf = np.zeros(2 * n.value)
        lib().orc_get_structure_factor(self._h, sf.ctypes.data_as(C.POINTER(C.c_double)), n.value)
        return kv.reshape(-1, 3), ak, sf[0::2] + 1j * sf[1::2]

    def constants(self):
        out = np.zeros(8)
        lib().orc_constants(self._h, out.ctypes.data_as(C.POINTER(C.c_double)))
        return dict(alfa=out[0], facq=out[1], spii=out[2], rskinsq=out[3], pi=out[4],
                    klmax=int(out[5]), kmmax=int(out[6]), knmax=int(out[7]))

    def __del__(self):
        try:
            if self._h:
                lib().orc_destroy(self._h)
        except Exception:
            pass
