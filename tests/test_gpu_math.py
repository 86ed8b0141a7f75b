"""The pair kernel's FP64 elementary functions against correctly rounded values.

The FP64 form evaluates e^x (Tang's table method, degree-5 Taylor on |r| <= ln2/128), 1/sqrt(x)
and 1/x (hardware estimate + one third-order step) with the inline device functions of
md_sim_amd/csrc/mdf_math.h; the reference's force loop calls Fortran's exp and sqrt
(src/forcem.f90:333-340).  A test-only probe library (tests/native/math_probe.hip) runs the
same device functions on host arrays; the reference values here are 40-digit decimal
evaluations, so each result's error is measured in ulps of the correctly rounded value:
e^x within 2 ulp, 1/sqrt(x) and 1/x within 1 ulp, over the arguments the pair kernel takes
(erfc and Buckingham exponents 0..50, squared distances and 1 + p a r) and well beyond them.
The hardware estimates the refinement starts from must be within 2^-20 relative, so that the
cubic error term of the one step is below 2^-58.
"""
import ctypes
import importlib.util
import math
import os
from decimal import Decimal, getcontext

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "native", "libmdf_math_probe.so")


def _lib():
    if not os.path.exists(LIB):
        spec = importlib.util.spec_from_file_location("mdf_probe_build", os.path.join(HERE, "native", "build.py"))
        nb = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(nb)
        nb.build()
    lib = ctypes.CDLL(LIB)
    lib.mdf_math_probe.restype = ctypes.c_int
    lib.mdf_math_probe.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    return lib


def test_probe_library_exports_the_probe():
    """CPU: the probe library builds and exports its one entry point (no GPU call)."""
    assert hasattr(_lib(), "mdf_math_probe")


def _probe(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros((5, x.size), dtype=np.float64)
    rc = _lib().mdf_math_probe(x.ctypes.data, x.size, out.ctypes.data)
    assert rc == 0, rc
    return out


def _ulps(y, exact):
    """|y - exact| in ulps of the double nearest `exact` (a Decimal)."""
    return float(abs(Decimal(float(y)) - exact)) / math.ulp(float(exact))


@pytest.mark.gpu
def test_exp_rsqrt_rcp_against_correctly_rounded():
    getcontext().prec = 40
    rng = np.random.default_rng(11)
    ln2_128 = math.log(2.0) / 128.0
    # e^-v: the kernel's range (0..50), reduction boundaries (odd multiples of ln2/128, where
    # |r| is largest, and their neighbours), then far out to the subnormal edge
    v_exp = np.concatenate([
        rng.uniform(0.0, 50.0, 3000),
        np.array([k * ln2_128 for k in range(1, 400, 2)]),
        np.nextafter(np.array([k * ln2_128 for k in range(1, 400, 2)]), 0.0),
        rng.uniform(50.0, 700.0, 500),
        [0.0, 5e-324, 1e-300, 1e-17, 1e-10, 0.5, 1.0, 700.0],
    ])
    # 1/sqrt(x), 1/x: squared distances (A^2), 1 + p a r, and six more decades either side
    v_pos = np.concatenate([
        np.exp(rng.uniform(math.log(1e-6), math.log(1e8), 3000)),
        rng.uniform(1.0, 3.0, 500),
        2.0 ** np.arange(-40, 41),
        [1.0, 3.0, 0.1, 7.0],
    ])
    out_e = _probe(v_exp)
    out_p = _probe(v_pos)

    worst_e = max(_ulps(y, Decimal(-float(v)).exp()) for y, v in zip(out_e[0], v_exp))
    worst_rs = max(_ulps(y, 1 / Decimal(float(v)).sqrt()) for y, v in zip(out_p[1], v_pos))
    worst_rc = max(_ulps(y, 1 / Decimal(float(v))) for y, v in zip(out_p[2], v_pos))
    seed_rs = max(abs(float(Decimal(float(y)) * Decimal(float(v)).sqrt() - 1)) for y, v in zip(out_p[3], v_pos))
    seed_rc = max(abs(float(Decimal(float(y)) * Decimal(float(v)) - 1)) for y, v in zip(out_p[4], v_pos))
    print("worst ulps: exp %.3f, rsqrt %.3f, rcp %.3f; hardware estimates: rsq %.2e (2^%.1f), rcp %.2e (2^%.1f)"
          % (worst_e, worst_rs, worst_rc, seed_rs, math.log2(seed_rs or 2 ** -60), seed_rc,
             math.log2(seed_rc or 2 ** -60)), flush=True)
    assert worst_e <= 2.0
    assert worst_rs <= 1.0
    assert worst_rc <= 1.0
    assert seed_rs < 2.0 ** -20 and seed_rc < 2.0 ** -20

    # far arguments: below the subnormal range the result is 0, e^-800 and beyond exactly 0
    far = _probe(np.array([746.0, 800.0, 1e4, 1e300]))[0]
    assert np.all(far == 0.0), far
    assert np.isnan(_probe(np.array([np.nan]))[0][0])
