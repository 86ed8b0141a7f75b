"""HIP engine behaviour beyond single-configuration parity: full-size workload, FP32 mode,
list rebuilds along a trajectory, device-resident inputs, graph replays, error paths."""
import os
import sys

import numpy as np
import pytest

from md_sim_amd import abi, systems
from helpers import rel_err, rms_rel

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _engine(s, **kw):
    from md_sim_amd.engine import ForceEngine
    return ForceEngine(s, **kw)


def test_full_size_16k_matches_oracle(oracle_lib):
    """The benchmark workload itself (16384 waters + 6144 ions, k 11/11/33) at 1e-10."""
    _say("case slab16k fp64")
    s = systems.config("slab16k")
    pos = abi.positions(s)
    r = _engine(s).force_get(*pos)
    ro = oracle_lib.Oracle(s).compute(*pos)
    for a in range(3):
        assert rel_err(r["f_site"][a], ro["f_site"][a]) < 1e-10
    assert rel_err(r["torque"], ro["torque"]) < 1e-10
    assert rel_err(r["f_com"], ro["f_com"]) < 1e-10
    assert abs(r["e_conf"] - ro["e_conf"]) < 1e-10 * abs(ro["e_conf"])
    assert abs(r["e_recip"] - ro["e_recip"]) < 1e-10 * abs(ro["e_recip"])
    assert r["n_pairs_listed"] == ro["n_pairs_listed"]
    assert r["n_wall_pairs_listed"] == ro["n_wall_pairs_listed"]


@pytest.mark.parametrize("name", ["bulk216", "slab2048"])
def test_fp32_force_error_bound(oracle_lib, name):
    """FP32 arithmetic with FP64 accumulation: force RMS error <= 1e-3 relative to FP64."""
    _say("case fp32", name)
    s = systems.config(name)
    pos = abi.positions(s)
    r = _engine(s, precision=abi.FP32_ACC64).force_get(*pos)
    ro = oracle_lib.Oracle(s).compute(*pos)
    err = rms_rel(r["f_site"], ro["f_site"])
    _say("  fp32 force rms rel err %.3e" % err)
    assert err < 1e-3
    assert rms_rel(r["torque"], ro["torque"]) < 1e-3
    assert abs(r["e_conf"] - ro["e_conf"]) < 1e-4 * abs(ro["e_recip"]) + 1e-4 * abs(ro["e_conf"])


def _mma_system(name):
    # slab512 at k-index 11, 14, 18 and 21: knmax 33, 42, 54 and 63, 5, 6, 7 and 8 kn tiles of
    # 16 (the kernels' register layouts and launch attributes change with the tile count; 11
    # and 21 are the k-sets of slab16k and slab32k)
    if name.startswith("slab512_k"):
        return systems.water_on_slab(512, 6, 4, seed=1201, rcut=10.0, kmax=int(name[9:]), lw=40.0, name=name)
    return systems.config(name)


@pytest.mark.parametrize("name", ["slab2048", "slab512", "bulk216", "slab512_k11", "slab512_k14", "slab512_k18",
                                  "slab512_k21"])
def test_fp32_structure_factor_on_matrix_cores(name, monkeypatch):
    """The FP32 form's reciprocal sum on the matrix cores (FP16 hi/lo split products, FP32
    tile sums, FP64 across blocks and rows): S(k) against the FP64 form's and the FP32
    vector-ALU kernel's, every k-vector written (no stale partials); forces against the
    vector-ALU kernels."""
    _say("case fp32 S(k)", name)
    s = _mma_system(name)
    pos = abi.positions(s)
    e64 = _engine(s, precision=abi.FP64)
    e64.force_get(*pos)
    _, _, s64 = e64.kvectors()
    e32 = _engine(s, precision=abi.FP32_ACC64)
    assert e32.reciprocal_on_matrix_cores()
    r32 = e32.force_get(*pos)
    _, _, s32 = e32.kvectors()
    monkeypatch.setenv("MDF_SF_MMA", "0")
    monkeypatch.setenv("MDF_BP_MMA", "0")
    ev = _engine(s, precision=abi.FP32_ACC64)
    rv = ev.force_get(*pos)
    _, _, sv = ev.kvectors()
    scale = np.abs(s64).max()
    err_m = np.abs(s32 - s64).max() / scale
    err_v = np.abs(sv - s64).max() / scale
    _say("  max |S32 - S64| / max |S64|: matrix cores %.2e, vector ALU %.2e" % (err_m, err_v))
    assert err_m < 2e-5 and err_v < 2e-5
    # the whole reciprocal path (structure factor and back-projection on the matrix cores)
    # against the vector-ALU kernels: site forces, COM forces and torques
    ef = rms_rel(r32["f_site"], rv["f_site"])
    _say("  force rms rel diff, matrix cores vs vector ALU: %.2e" % ef)
    assert ef < 1e-5
    assert rms_rel(r32["f_com"], rv["f_com"]) < 1e-5 and rms_rel(r32["torque"], rv["torque"]) < 1e-5


def test_rebuilds_along_trajectory_match_oracle(oracle_lib):
    """Rigid translations of every molecule step by step: the skin test fires at the same
    steps as the reference rule and the rebuilt pair sets stay identical."""
    _say("case trajectory slab2048")
    s = systems.config("slab2048")
    x, y, z, cx, cy, cz = abi.positions(s)
    rng = np.random.default_rng(7)
    v = rng.normal(0, 0.05, size=(s.nrig, 3))
    e = _engine(s)
    o = oracle_lib.Oracle(s)
    builds = []
    for t in range(1, 9):
        xs, ys, zs = x.copy(), y.copy(), z.copy()
        d = np.repeat(v * t, 3, axis=0)
        xs[:s.nriga] += d[:, 0]; ys[:s.nriga] += d[:, 1]; zs[:s.nriga] += d[:, 2]
        c = (cx + v[:, 0] * t, cy + v[:, 1] * t, cz + v[:, 2] * t)
        r = e.force_get(xs, ys, zs, *c)
        ro = o.compute(xs, ys, zs, *c)
        assert r["list_rebuilt"] == ro["list_rebuilt"], t
        builds.append(r["list_rebuilt"])
        assert rel_err(r["f_site"], ro["f_site"]) < 1e-10, t
        l1, l2, w1, w2 = e.pairs()
        m1, m2, u1, u2 = o.pairs()
        assert np.array_equal(l1, m1) and np.array_equal(l2, m2) and np.array_equal(w1, u1), t
    assert any(builds), "the trajectory never triggered a rebuild"


def test_device_inputs_and_graph_replays_match_host_path():
    """mdf_compute_device on torch tensors, replayed from captured graphs several times with
    host work in between, equals mdf_compute on host arrays and the direct-launch path."""
    import torch
    _say("case device/graph pore_small")
    s = systems.config("pore_small")
    pos = abi.positions(s)
    ref = _engine(s).force_get(*pos)
    e = _engine(s)
    dev = [torch.tensor(a[:s.nmove] if k < 3 else a, dtype=torch.float64, device="cuda:0")
           for k, a in enumerate(pos)]
    for rep in range(5):
        e.compute_device(*[t.data_ptr() for t in dev])
        r = e.read_outputs()
        _ = torch.ones(1000, device="cuda:0").sum().item()  # eager work between replays
        assert np.array_equal(r["f_site"], ref["f_site"]), rep
        assert r["e_conf"] == ref["e_conf"], rep
    os.environ["MDF_GRAPHS"] = "0"
    try:
        r2 = _engine(s).force_get(*pos)
    finally:
        del os.environ["MDF_GRAPHS"]
    assert np.array_equal(r2["f_site"], ref["f_site"])


@pytest.mark.parametrize("prec", [abi.FP64, abi.FP32_ACC64], ids=["fp64", "fp32"])
def test_forked_kspace_graph_equals_one_stream(prec, monkeypatch):
    """The captured step graph runs the k-space chain (structure factor, reduction,
    back-projection) on a second stream beside the skin test, list build and pair kernel,
    joined before the epilogue.  Over a trajectory of rigid translations with list rebuilds,
    device-resident and replayed from graphs, every step equals the one-stream graph
    (MDF_FORK_KSPACE=0) and the direct launches bit for bit."""
    import torch
    _say("case forked k-space slab2048", prec)
    s = systems.config("slab2048")
    x, y, z, cx, cy, cz = abi.positions(s)
    rng = np.random.default_rng(11)
    v = rng.normal(0, 0.05, size=(s.nrig, 3))
    frames = []
    for t in range(10):
        xs, ys, zs = x.copy(), y.copy(), z.copy()
        d = np.repeat(v * t, 3, axis=0)
        xs[:s.nriga] += d[:, 0]; ys[:s.nriga] += d[:, 1]; zs[:s.nriga] += d[:, 2]
        frames.append([torch.tensor(a[:s.nmove] if k < 3 else a, dtype=torch.float64, device="cuda:0")
                       for k, a in enumerate((xs, ys, zs, cx + v[:, 0] * t, cy + v[:, 1] * t, cz + v[:, 2] * t))])

    def run(env):
        for k, val in env.items():
            monkeypatch.setenv(k, val)
        e = _engine(s, precision=prec)
        out = []
        for f in frames:
            e.compute_device(*[a.data_ptr() for a in f])
            out.append(e.read_outputs())
        e.close()
        for k in env:
            monkeypatch.delenv(k)
        return out

    forked = run({"MDF_FORK_KSPACE": "1"})   # slab2048 is below the default size for the fork
    one = run({"MDF_FORK_KSPACE": "0"})
    direct = run({"MDF_GRAPHS": "0"})
    assert any(r["list_rebuilt"] for r in forked[1:]), "the trajectory never triggered a rebuild"
    for t, (a, b, c) in enumerate(zip(forked, one, direct)):
        for key in ("f_site", "torque", "f_com"):
            assert np.array_equal(a[key], b[key]) and np.array_equal(a[key], c[key]), (t, key)
        assert a["e_conf"] == b["e_conf"] == c["e_conf"], t
        assert a["e_recip"] == b["e_recip"] == c["e_recip"], t


def test_list_capacity_overflow_reports_error():
    """A fixed, too small list capacity stops like the reference ('nnlist too big')."""
    from md_sim_amd.engine import MDForceError
    _say("case overflow")
    s = systems.config("pore_small")
    with pytest.raises(MDForceError) as ei:
        _engine(s, max_neighbors=4)
    assert ei.value.code == 3


def test_deterministic_repeat():
    """Same input, same bits (no atomics in any force or energy sum)."""
    _say("case determinism slab2048")
    s = systems.config("slab2048")
    pos = abi.positions(s)
    e = _engine(s)
    a = e.force_get(*pos)
    b = e.force_get(*pos)
    assert np.array_equal(a["f_site"], b["f_site"]) and a["e_conf"] == b["e_conf"]
    assert np.array_equal(a["e_vdw"], b["e_vdw"]) and np.array_equal(a["virial"], b["virial"])


@pytest.mark.parametrize("shape", [(8, False, 1024, 512), (32, False, 8192, 2048), (16, True, 2048, 1024),
                                   (32, True, 4096, 1024)])
def test_launch_shapes_keep_parity(oracle_lib, shape):
    """Every launch shape the autotuner can pick gives the FP64 result."""
    _say("case launch shape", shape)
    s = systems.config("slab2048")
    pos = abi.positions(s)
    e = _engine(s)
    e.set_launch_shapes(*shape)
    r = e.force_get(*pos)
    ro = oracle_lib.Oracle(s).compute(*pos)
    assert rel_err(r["f_site"], ro["f_site"]) < 1e-10
    assert abs(r["e_conf"] - ro["e_conf"]) < 1e-10 * abs(ro["e_conf"])


def test_autotune_then_parity(oracle_lib):
    _say("case autotune slab2048")
    s = systems.config("slab2048")
    pos = abi.positions(s)
    e = _engine(s)
    ch = e.autotune(reps=3)
    _say("  chosen", ch)
    assert ch["pair_lanes"] in (8, 16, 32) and ch["sf_blocks"] > 0 and ch["bp_blocks"] > 0
    r = e.force_get(*pos)
    ro = oracle_lib.Oracle(s).compute(*pos)
    assert rel_err(r["f_site"], ro["f_site"]) < 1e-10


def test_fp16_split_range_of_large_structure_factors():
    """The matrix-core back-projection splits T = A_k conj S(k) and kn T into FP16 hi/lo
    parts; FP16 stops at 65504.  With every charge of slab512 scaled by 5000, max |T| and
    max |kn T| are about 1.6e5 and 2.5e5 (32 and 50 unscaled); the per-row power-of-two scaling keeps the FP32 form
    within its usual error of the FP64 form (unscaled, hi would saturate and the forces
    would be wrong by far more)."""
    _say("case fp16 range slab512 x5000")
    s = systems.config("slab512")
    q0 = s.charges
    s.charges = lambda: q0() * 5000.0
    pos = abi.positions(s)
    e64 = _engine(s, precision=abi.FP64)
    r64 = e64.force_get(*pos)
    kv, ak, s64 = e64.kvectors()
    t = ak * np.abs(s64)
    kn = np.abs(np.asarray(kv).reshape(-1, 3)[:, 2])
    _say("  max |T| %.3e, max |kn T| %.3e" % (t.max(), (kn * t).max()))
    assert t.max() > 2 * 65504 and (kn * t).max() > 2 * 65504
    e32 = _engine(s, precision=abi.FP32_ACC64)
    assert e32.reciprocal_on_matrix_cores()
    r32 = e32.force_get(*pos)
    err = rms_rel(r32["f_site"], r64["f_site"])
    _say("  fp32 (matrix cores) force rms rel err vs fp64 %.3e" % err)
    assert err < 1e-5
    assert rms_rel(r32["torque"], r64["torque"]) < 1e-5
    assert abs(r32["e_recip"] - r64["e_recip"]) < 1e-5 * abs(r64["e_recip"])


@pytest.mark.parametrize("name", ["bulk216", "pore_small", "slab2048", "slab512_k11", "slab512_k21"])
def test_fp64_matrix_cores_against_vector_alu(oracle_lib, name, monkeypatch):
    """The FP64 form's reciprocal sum on the FP64 matrix cores (k_sf64, k_sf64_reduce, k_bp64:
    kn symmetry, sorted rows, branch-free product sequences per tile width) against the
    vector-ALU kernels and the oracle: S(k), site forces, energies and virial at FP64
    rounding; kn tiles 5 and 8 cover the widths of slab16k and slab32k."""
    _say("case fp64 matrix cores", name)
    s = _mma_system(name)
    pos = abi.positions(s)
    e = _engine(s, precision=abi.FP64)
    assert e.reciprocal_on_matrix_cores()
    r = e.force_get(*pos)
    _, _, sm = e.kvectors()
    monkeypatch.setenv("MDF_MMA64", "0")
    ev = _engine(s, precision=abi.FP64)
    assert not ev.reciprocal_on_matrix_cores()
    rv = ev.force_get(*pos)
    _, _, sv = ev.kvectors()
    es = np.abs(sm - sv).max() / np.abs(sv).max()
    ef = rel_err(r["f_site"], rv["f_site"])
    _say("  matrix vs vector: S(k) %.2e, forces %.2e" % (es, ef))
    assert es < 1e-13 and ef < 1e-12
    assert abs(r["e_recip"] - rv["e_recip"]) < 1e-13 * abs(rv["e_recip"])
    assert np.abs(r["virial"] - rv["virial"]).max() < 1e-12 * np.abs(rv["virial"]).max()
    if name in ("bulk216", "pore_small", "slab512_k11"):
        ro = oracle_lib.Oracle(s).compute(*pos)
        assert rel_err(r["f_site"], ro["f_site"]) < 1e-10
        assert abs(r["e_conf"] - ro["e_conf"]) < 1e-10 * abs(ro["e_conf"])
