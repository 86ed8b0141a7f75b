"""Batches of independent systems: one launch per stage over every system of the batch."""
import sys

import numpy as np
import pytest

from md_sim_amd import abi, systems
from helpers import rel_err

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _dev(s, pos):
    import torch
    return [torch.tensor(a[:s.nmove] if k < 3 else (a if len(a) else np.zeros(1)), dtype=torch.float64,
                         device="cuda:0") for k, a in enumerate(pos)]


def _ptrs(ts, s):
    return tuple(t.data_ptr() for t in ts[:3]) + (tuple(t.data_ptr() for t in ts[3:]) if s.nrig else (0, 0, 0))


def _mixed():
    syss = [systems.config("pore_small"), systems.config("bulk216"), systems.rocksalt_bulk(2, rcut=5.5, kmax=5)]
    syss += systems.replica_set(3, seed=4100)
    kws = [{}, {}, dict(z_mode=abi.Z_PERIODIC)] + [{}] * 3
    return syss, kws


@pytest.mark.parametrize("precision", [abi.FP64, abi.FP32_ACC64])
@pytest.mark.parametrize("kernels", ["matrix", "valu"])
def test_batch_equals_single_handles(precision, kernels, monkeypatch):
    """A mixed batch (slab, bulk 3-D form, water-free crystal, replicas of uneven size) gives
    bit for bit what each handle gives alone.  By default the handles run their k-sum on the
    matrix cores (FP64 products in the FP64 form, FP16 hi/lo products in the FP32 form) and
    so does the batch, one launch per kernel template over the systems that share it; with
    the vector-ALU k-sum everywhere (MDF_SF_MMA=0 MDF_BP_MMA=0 MDF_MMA64=0) the batch runs
    that one."""
    from md_sim_amd.engine import ForceBatch, ForceEngine
    _say("case batch vs single, precision", precision, kernels)
    syss, kws = _mixed()
    if kernels == "valu":
        monkeypatch.setenv("MDF_SF_MMA", "0")
        monkeypatch.setenv("MDF_BP_MMA", "0")
        monkeypatch.setenv("MDF_MMA64", "0")
    engines, singles, dev = [], [], []
    for s, kw in zip(syss, kws):
        pos = abi.positions(s)
        e = ForceEngine(s, precision=precision, **kw)
        assert e.reciprocal_on_matrix_cores() == (kernels == "matrix")
        singles.append(e.force_get(*pos))
        engines.append(e)
        dev.append(_dev(s, pos))
    b = ForceBatch(engines)
    for rep in range(3):
        b.compute_device([_ptrs(d, s) for d, s in zip(dev, syss)])
        outs = b.read_outputs()
        for i, (o, r) in enumerate(zip(outs, singles)):
            assert np.array_equal(o["f_site"], r["f_site"]), (rep, i)
            assert np.array_equal(o["torque"], r["torque"]), (rep, i)
            assert o["e_conf"] == r["e_conf"] and o["e_recip"] == r["e_recip"], (rep, i)
            assert np.array_equal(o["e_vdw"], r["e_vdw"]) and np.array_equal(o["e_coul_real"], r["e_coul_real"])
            assert o["n_pairs_listed"] == r["n_pairs_listed"], (rep, i)
    b.close()
    if kernels == "matrix":
        # the same handles batched with the vector-ALU k-sum forced: forces no longer equal
        # the handles' bit for bit (so the batch above did run the matrix-core kernels), and
        # agree to FP64 rounding (FP64 form) or at the FP32 form's budget
        fscale = max(np.abs(r["f_site"]).max() for r in singles)
        monkeypatch.setenv("MDF_BATCH_VALU", "1")
        bv = ForceBatch(engines)
        monkeypatch.delenv("MDF_BATCH_VALU")
        bv.compute_device([_ptrs(d, s) for d, s in zip(dev, syss)])
        outs = bv.read_outputs()
        assert any(not np.array_equal(o["f_site"], r["f_site"]) for o, r in zip(outs, singles))
        for i, (o, r) in enumerate(zip(outs, singles)):
            d = np.abs(r["f_site"] - o["f_site"]).max()
            assert d <= (1e-12 if precision == abi.FP64 else 1e-5) * fscale, (i, d, fscale)
        bv.close()
    for e in engines:
        e.close()


@pytest.mark.parametrize("precision", [abi.FP64, abi.FP32_ACC64], ids=["fp64", "fp32"])
def test_batch_forked_kspace_equals_single_handles(precision, monkeypatch):
    """With MDF_FORK_KSPACE_BATCH=1 the batch runs its k-space kernels on a second stream
    beside the skin test, list build and pair kernel (joined before the epilogue); the
    results stay bit for bit those of the handles alone, over graph replays."""
    from md_sim_amd.engine import ForceBatch, ForceEngine
    _say("case forked batch, precision", precision)
    syss, kws = _mixed()
    engines, singles, dev = [], [], []
    for s, kw in zip(syss, kws):
        pos = abi.positions(s)
        e = ForceEngine(s, precision=precision, **kw)
        singles.append(e.force_get(*pos))
        engines.append(e)
        dev.append(_dev(s, pos))
    monkeypatch.setenv("MDF_FORK_KSPACE_BATCH", "1")
    b = ForceBatch(engines)
    monkeypatch.delenv("MDF_FORK_KSPACE_BATCH")
    for rep in range(4):
        b.compute_device([_ptrs(d, s) for d, s in zip(dev, syss)])
        outs = b.read_outputs()
        for i, (o, r) in enumerate(zip(outs, singles)):
            assert np.array_equal(o["f_site"], r["f_site"]), (rep, i)
            assert o["e_conf"] == r["e_conf"] and o["e_recip"] == r["e_recip"], (rep, i)
    b.close()
    for e in engines:
        e.close()


def test_batch_trajectory_matches_oracle(oracle_lib):
    """Replicas moving independently: rebuilds fire per replica as the reference rule says and
    every replica's forces stay at the FP64 bar."""
    import torch
    from md_sim_amd.engine import ForceBatch, ForceEngine
    _say("case batch trajectory")
    syss = systems.replica_set(4, seed=4200)
    engines = [ForceEngine(s) for s in syss]
    oracles = [oracle_lib.Oracle(s) for s in syss]
    b = ForceBatch(engines)
    rng = np.random.default_rng(5)
    vel = [rng.normal(0, 0.06, size=(s.nrig, 3)) for s in syss]
    builds = np.zeros(len(syss), int)
    for t in range(1, 7):
        devs, hosts = [], []
        for s, v in zip(syss, vel):
            x, y, z, cx, cy, cz = abi.positions(s)
            d = np.repeat(v * t, 3, axis=0)
            x, y, z = x.copy(), y.copy(), z.copy()
            x[:s.nriga] += d[:, 0]; y[:s.nriga] += d[:, 1]; z[:s.nriga] += d[:, 2]
            c = (cx + v[:, 0] * t, cy + v[:, 1] * t, cz + v[:, 2] * t)
            hosts.append((x, y, z) + c)
            devs.append(_dev(s, hosts[-1]))
        b.compute_device([_ptrs(d, s) for d, s in zip(devs, syss)])
        outs = b.read_outputs()
        for i, (o, s) in enumerate(zip(outs, syss)):
            ro = oracles[i].compute(*hosts[i])
            assert o["list_rebuilt"] == ro["list_rebuilt"], (t, i)
            builds[i] += o["list_rebuilt"]
            assert rel_err(o["f_site"], ro["f_site"]) < 1e-10, (t, i)
            assert abs(o["e_conf"] - ro["e_conf"]) < 1e-10 * abs(ro["e_conf"]), (t, i)
        torch.cuda.synchronize()
    assert builds.sum() > 0
    b.close()


def test_batch_list_overflow_reports_error():
    """A system of the batch with a fixed, too small list capacity stops the batch read with
    the list-overflow status."""
    from md_sim_amd.engine import ForceBatch, ForceEngine, MDForceError
    _say("case batch overflow")
    s_ok = systems.config("pore_small")
    e_ok = ForceEngine(s_ok)
    s_bad = systems.replica_set(1, seed=4300)[0]
    e_bad = ForceEngine(s_bad, max_neighbors=1024)   # ample at creation ...
    b = ForceBatch([e_ok, e_bad])
    # ... then squeeze the second system's molecules together so its rows outgrow the cap
    x, y, z, cx, cy, cz = abi.positions(s_bad)
    f = 0.45
    xs, ys, zs = x.copy(), y.copy(), z.copy()
    off = np.repeat(np.stack([cx, cy, cz], 1) * (f - 1.0), 3, axis=0)
    xs[:s_bad.nriga] += off[:, 0]; ys[:s_bad.nriga] += off[:, 1]; zs[:s_bad.nriga] += off[:, 2]
    bad = _dev(s_bad, (xs, ys, zs, cx * f, cy * f, cz * f))
    ok = _dev(s_ok, abi.positions(s_ok))
    b.compute_device([_ptrs(ok, s_ok), _ptrs(bad, s_bad)])
    with pytest.raises(MDForceError) as ei:
        b.read_outputs()
    assert ei.value.code == 3
    b.close()


def test_batch_autotune_keeps_results():
    """The batch tuner changes launch shapes only: after it, the batch still gives each
    system's forces and energies at the FP64 bar."""
    from md_sim_amd.engine import ForceBatch, ForceEngine
    _say("case batch autotune")
    syss = systems.replica_set(4, seed=4400)
    engines = [ForceEngine(s) for s in syss]
    dev = [_dev(s, abi.positions(s)) for s in syss]
    ref = [e.force_get(*abi.positions(s)) for e, s in zip(engines, syss)]
    b = ForceBatch(engines)
    b.compute_device([_ptrs(d, s) for d, s in zip(dev, syss)])
    b.read_outputs()
    ch = b.autotune(reps=2)
    assert ch["pair_lanes"] in (8, 16, 32) and ch["sf_blocks"] > 0 and ch["bp_blocks"] > 0
    for _ in range(2):
        b.compute_device([_ptrs(d, s) for d, s in zip(dev, syss)])
        outs = b.read_outputs()
        for i, (o, r) in enumerate(zip(outs, ref)):
            assert rel_err(o["f_site"], r["f_site"]) < 1e-12, i
            assert abs(o["e_conf"] - r["e_conf"]) < 1e-12 * abs(r["e_conf"]), i
            assert o["n_pairs_listed"] == r["n_pairs_listed"], i
    b.close()
