"""A short, fixed run of the force step for hardware-counter profiling.

``python -m md_sim_amd.profile_step [--workload slab16k] [--precision fp64] [--steps 5]``
runs ``--steps`` force evaluations on device-resident positions with direct kernel launches
(no graph replay, so a profiler attributes counters to each dispatch), then one plain
device-to-device copy of ``--copy-mib`` MiB.  The copy is the calibration for byte
counters: it fetches and writes exactly that many bytes, so the ratio of its measured
FETCH_SIZE / WRITE_SIZE to the true size scales the counters of the engine's kernels.

``--mts N`` runs the k-space surrogate with the full sum every N-th step instead, on a
librating water trajectory: ``--mts-train`` full-sum training steps, then ``--steps`` steps
of the scheme.

``--rebuild-every-step`` builds the pair list on every step (per-stage timing mode 2), so
that every dispatch of the list-build kernels is a full build.

``--chempot "N WMIN WMAX BINS MID"`` adds the Widom insertions after every step (tot.inp's
ichmax, lwmin, lwmax, lchmax, ilmid).

``--pair-lanes G [--pair-strided]`` fixes the pair kernel's launch shape (else the default).
"""
from __future__ import annotations

import argparse
import os
import sys


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="slab16k")
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64")
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--copy-mib", type=int, default=256)
    ap.add_argument("--mts", type=int, default=0, help="full k-sum every N-th step with the surrogate between")
    ap.add_argument("--mts-model", choices=["network", "held"], default="network")
    ap.add_argument("--mts-train", type=int, default=8)
    ap.add_argument("--rebuild-every-step", action="store_true")
    ap.add_argument("--chempot", default="", help="Widom insertions: 'ichmax lwmin lwmax lchmax ilmid'")
    ap.add_argument("--pair-lanes", type=int, default=0, help="lanes per lead of the pair kernel (8, 16, 32)")
    ap.add_argument("--pair-strided", action="store_true", help="strided row pass instead of the compacted one")
    a = ap.parse_args(argv)
    os.environ["MDF_GRAPHS"] = "0"
    import torch

    from . import abi, systems
    from .engine import ForceEngine

    s = systems.config(a.workload)
    prec = abi.FP32_ACC64 if a.precision == "fp32" else abi.FP64
    pos = abi.positions(s)
    eng = ForceEngine(s, precision=prec, virial=False)
    if a.pair_lanes:
        eng.set_launch_shapes(pair_lanes=a.pair_lanes, pair_strided=a.pair_strided, sf_blocks=8192, bp_blocks=2048)
    if a.rebuild_every_step:
        eng.set_stage_timing(True, rebuild_every_step=True)
    if a.mts > 1:
        frames = systems.libration_frames(s, a.mts_train + a.steps, seed=7)
        eng.force_get(*frames[0])
        eng.mts_enable(interval=a.mts, model=a.mts_model, train_frames=a.mts_train, tol=0.5)
    else:
        frames = [pos] * a.steps
    if a.chempot:
        v = a.chempot.split()
        eng.chempot_enable(n_insert=int(v[0]), wall_min=float(v[1]), wall_max=float(v[2]), n_bins=int(v[3]),
                           bin_mid=int(v[4]))
    devf = [[torch.tensor(v[:s.nmove] if k < 3 else v, dtype=torch.float64, device="cuda:0")
             for k, v in enumerate(f)] for f in frames]
    for d in devf:
        eng.compute_device(*[t.data_ptr() for t in d])
    r = eng.read_outputs()
    if a.mts > 1:
        print("scheme: %s" % eng.mts_stats(), flush=True)
    if a.chempot:
        c = eng.chempot()
        print("insertions: %d orientations, mu %.4f kJ/mol" % (int(c["inave"].sum()), c["mu"]), flush=True)
    print("steps done: e_conf %.6f" % r["e_conf"], flush=True)
    n = a.copy_mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    print("copy done: %d MiB" % a.copy_mib, flush=True)
    eng.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
