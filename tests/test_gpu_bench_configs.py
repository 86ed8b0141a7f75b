"""The configurations the benchmark runs, each checked against the FP64 form and the oracle.

* slab16k in the FP32 form: 16384 waters + 6144 ions, k 11/11/33, so the matrix-core
  reciprocal kernels run with 5 kn tiles of 16 (k_sf_mma<5>, k_bp_mma<5>);
* slab32k in the FP32 form: 32768 waters + 12288 ions, k 21/21/63, 8 kn tiles (k_sf_mma<8>,
  k_bp_mma<8>, which also change their launch attributes), the reference that the
  surrogate's errors are measured against;
* the 64-replica batch (replica_set(64, seed=3000)): FP64 batch against the oracle on a
  sample of replicas and against FP64 single handles in full; FP32 batch against FP64 handles.

The reference FP64 CPU path is the oracle (oracle/oracle.c, pinned to the reference build's
golden vectors); at slab32k the oracle would need minutes of CPU time, so the FP64 engine,
pinned to the oracle at slab16k at 1e-10, is the comparison there.
"""
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


def _n_tiles(e):
    kv, _, _ = e.kvectors()
    return (2 * int(np.abs(kv[:, 2]).max()) + 1 + 15) // 16


def _fp32_vs(ref_f, ref_s, s, pos, nt_expected, monkeypatch):
    e32 = _engine(s, precision=abi.FP32_ACC64)
    assert e32.reciprocal_on_matrix_cores()
    assert _n_tiles(e32) == nt_expected
    r32 = e32.force_get(*pos)
    _, _, s32 = e32.kvectors()
    e32.close()
    # the FP32 form on the vector ALU, for scale: the matrix cores must be as accurate
    monkeypatch.setenv("MDF_SF_MMA", "0")
    monkeypatch.setenv("MDF_BP_MMA", "0")
    ev = _engine(s, precision=abi.FP32_ACC64)
    rv = ev.force_get(*pos)
    ev.close()
    monkeypatch.delenv("MDF_SF_MMA")
    monkeypatch.delenv("MDF_BP_MMA")
    err_f = rms_rel(r32["f_site"], ref_f["f_site"])
    err_v = rms_rel(rv["f_site"], ref_f["f_site"])
    err_s = np.abs(s32 - ref_s).max() / np.abs(ref_s).max()
    de, dv = abs(r32["e_conf"] - ref_f["e_conf"]), abs(rv["e_conf"] - ref_f["e_conf"])
    _say("  fp32 force rms rel err: matrix cores %.3e, vector ALU %.3e; max |S32 - S64| / max |S64| %.2e; "
         "e_conf err %.3e / %.3e" % (err_f, err_v, err_s, de, dv))
    assert err_f < 1e-5            # FP32-class: the force RMS error budget of the FP32 form is 1e-3
    assert err_f < 2.0 * err_v + 1e-7
    assert err_s < 2e-5
    assert rms_rel(r32["torque"], ref_f["torque"]) < 1e-5
    assert rms_rel(r32["f_com"], ref_f["f_com"]) < 1e-5
    # e_conf sums ~1e7 site-pair terms of either sign, each rounded to FP32
    assert de < 1e-5 * (abs(ref_f["e_conf"]) + abs(ref_f["e_recip"]))
    assert r32["n_pairs_listed"] == ref_f["n_pairs_listed"]
    return r32


def test_slab16k_fp32_matrix_cores_against_oracle(oracle_lib, monkeypatch):
    """The benchmark's 16k-water workload in the FP32 form (5 kn tiles) against the oracle."""
    _say("case slab16k fp32 vs oracle")
    s = systems.config("slab16k")
    pos = abi.positions(s)
    ro = oracle_lib.Oracle(s).compute(*pos)
    e64 = _engine(s, precision=abi.FP64)
    r64 = e64.force_get(*pos)
    _, _, s64 = e64.kvectors()
    assert rel_err(r64["f_site"], ro["f_site"]) < 1e-10
    _fp32_vs(ro, s64, s, pos, 5, monkeypatch)


def test_slab32k_fp32_matrix_cores_against_fp64(monkeypatch):
    """The surrogate's 32k-water interface in the FP32 form (8 kn tiles) against FP64."""
    _say("case slab32k fp32 vs fp64")
    s = systems.config("slab32k")
    pos = abi.positions(s)
    e64 = _engine(s, precision=abi.FP64)
    r64 = e64.force_get(*pos)
    _, _, s64 = e64.kvectors()
    _fp32_vs(r64, s64, s, pos, 8, monkeypatch)


@pytest.fixture(scope="module")
def replicas():
    return systems.replica_set(64, seed=3000)


def _batch_step(syss, prec):
    import torch
    from md_sim_amd.engine import ForceBatch
    engines = [_engine(s, precision=prec) for s in syss]
    dev = []
    for s in syss:
        pos = abi.positions(s)
        dev.append([torch.tensor(a[:s.nmove] if k < 3 else a, dtype=torch.float64, device="cuda:0")
                    for k, a in enumerate(pos)])
    b = ForceBatch(engines)
    b.compute_device([[t.data_ptr() for t in d] for d in dev])
    out = b.read_outputs()
    b.close()
    return engines, out


def test_replica_batch_64_fp64_against_oracle_and_handles(oracle_lib, replicas):
    """The benchmark's 64-replica batch in FP64: every replica equals its own handle bit for
    bit (both run the k-sum on the FP64 matrix cores), and a sample of replicas (smallest,
    largest, one between) matches the oracle."""
    _say("case 64-replica batch fp64")
    engines, out = _batch_step(replicas, abi.FP64)
    for i, (s, e) in enumerate(zip(replicas, engines)):
        assert e.reciprocal_on_matrix_cores(), i
        r = e.force_get(*abi.positions(s))
        assert np.array_equal(out[i]["f_site"], r["f_site"]), i
        assert out[i]["e_conf"] == r["e_conf"], i
    sizes = [s.nrig for s in replicas]
    sample = sorted({int(np.argmin(sizes)), int(np.argmax(sizes)), 17})
    for i in sample:
        s = replicas[i]
        _say("  replica %d: %d waters, %d ions vs oracle" % (i, s.nrig, s.nwall))
        ro = oracle_lib.Oracle(s).compute(*abi.positions(s))
        assert rel_err(out[i]["f_site"], ro["f_site"]) < 1e-10, i
        assert rel_err(out[i]["torque"], ro["torque"]) < 1e-10, i
        assert abs(out[i]["e_conf"] - ro["e_conf"]) < 1e-10 * abs(ro["e_conf"]), i
    for e in engines:
        e.close()


def test_replica_batch_64_fp32_against_fp64(replicas):
    """The same batch in the FP32 form against FP64 single handles, every replica, and bit
    for bit against its own FP32 handles (the FP16 hi/lo matrix-core k-sum in both)."""
    _say("case 64-replica batch fp32")
    engines, out = _batch_step(replicas, abi.FP32_ACC64)
    worst = 0.0
    for i, s in enumerate(replicas):
        r32 = engines[i].force_get(*abi.positions(s))
        assert engines[i].reciprocal_on_matrix_cores(), i
        assert np.array_equal(out[i]["f_site"], r32["f_site"]) and out[i]["e_conf"] == r32["e_conf"], i
        r64 = _engine(s, precision=abi.FP64).force_get(*abi.positions(s))
        err = rms_rel(out[i]["f_site"], r64["f_site"])
        worst = max(worst, err)
        assert err < 1e-5, (i, err)
        assert abs(out[i]["e_conf"] - r64["e_conf"]) < 1e-5 * (abs(r64["e_conf"]) + abs(r64["e_recip"])), i
    _say("  worst fp32 force rms rel err %.3e" % worst)
    for e in engines:
        e.close()
