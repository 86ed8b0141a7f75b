"""Benchmark: MD steps/s (forces + Ewald) at 16k H2O on one MI355X.

A step is one force evaluation of the reference's force path (force_get,
src/forcem.f90:93-129) on device-resident positions: pair-list skin test (and rebuild when
due), real-space erfc + LJ/Buckingham pair kernel, reciprocal k-sum (structure factor and
back-projection), slab and wall corrections, COM forces/torques and energies.  The
positions follow a synthetic trajectory (every molecule oscillates rigidly about its start
with a seeded amplitude of ~0.25 A and an 80-step period), so the list is rebuilt as often
as the skin criterion demands.

Workload (default ``slab16k``): 16384 rigid SPC/E waters on a frozen 6144-ion rock-salt
slab, box 90.24 x 90.24 x 270.72 A, rc 10 A, skin 1 A, k-index 11/11/33 (8589 k-vectors).
The 16k-water configuration computes pair and structure-factor arithmetic in FP32 with FP64
accumulation, so that is the default (``--precision fp32``); the FP64 form runs the same
trajectory beside it and is reported under ``fp64`` with the FP32 form's force RMS error
against it.  ``--precision fp64`` times the FP64 form alone.

Prints one JSON line.  ``roofline`` is for the dominant stage, the reciprocal sum: its
algorithmic work 18 flop per (k-vector, moving site) + 8 flop per ((kl,km) row, site)
(FP32 arithmetic; the FP32 form runs its products on the FP16 matrix cores as hi/lo pairs,
priced against the FP32 peak) over the stage's mean kernel time, measured with HIP events on the engine's stream over a
second pass of the same steps run with per-stage events (the timed pass replays a captured
HIP graph, inside which events cannot be timed).  ``cpu_baseline`` times the reference's
own force path (built from its sources into oracle/_ref) on a few force evaluations of the
same system on the host's cores (MPI ranks, the reference's own decomposition; 16 by
default) and on one core, or the CPU oracle when that build is absent.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MD steps/s (forces+Ewald) at 16k H2O; surrogate-MTS speedup at fixed force err"
PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3}
PROFILE_TRAFFIC = os.path.join(ROOT, "profiles", "reciprocal_traffic_slab16k.json")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="slab16k")
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp32",
                    help="fp32: FP32 pair and structure-factor arithmetic with FP64 accumulation (the "
                         "16k-water configuration's form; the FP64 form is timed beside it); fp64: FP64 only")
    ap.add_argument("--cpu-reps", type=int, default=3, help="force evaluations in the one-core CPU sample")
    ap.add_argument("--cpu-cores", type=int, default=16,
                    help="MPI ranks of the host-core CPU baseline (the reference's own decomposition)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--virial", action="store_true", help="also accumulate the pair virial")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shapes")
    ap.add_argument("--amplitude", type=float, default=0.25,
                    help="rms per-axis oscillation amplitude of a molecule (A)")
    ap.add_argument("--period", type=float, default=80.0, help="oscillation period (steps)")
    ap.add_argument("--replicas", type=int, default=64,
                    help="replica batch line: this many independent interface systems of uneven size "
                         "(0: skip)")
    ap.add_argument("--replica-steps", type=int, default=50)
    ap.add_argument("--replica-warmup", type=int, default=5)
    ap.add_argument("--mts-workloads", default="slab16k,slab32k",
                    help="workloads of the surrogate multiple-time-step line ('' to skip)")
    ap.add_argument("--mts-steps", type=int, default=48, help="timed steps of each multiple-time-step run")
    ap.add_argument("--mts-train", type=int, default=16, help="full-sum training steps before the scheme starts")
    return ap.parse_args()


def cpu_baseline(system, reps: int, cores: int = 1):
    """Reference force path on the host: force evaluations per second, on `cores` MPI ranks
    (the reference's own decomposition: leads and k-vectors over ranks, force_sum's
    all-reduce), else on one core."""
    from oracle import refrun
    if refrun.ref_available():
        nproc = cores if cores > 1 and refrun.mpi_available() else 1
        d = refrun.run_reference(system, nrep=reps, nproc=nproc, timeout=1200)
        n, sec = d["_timing"]
        return dict(value=n / sec, unit="steps/s", cores=nproc, kind="reference",
                    sample="%d force evaluations (no list rebuild) of %s by the reference's own "
                           "force_real/force_fouri/force_intra/force_sum, built from its Fortran "
                           "sources with amdflang -O2, %d MPI rank%s" % (int(n), system.name, nproc,
                                                                        "s" if nproc > 1 else ""))
    from md_sim_amd import abi
    from oracle import pyoracle
    o = pyoracle.Oracle(system)
    pos = abi.positions(system)
    o.compute(*pos)
    t0 = time.perf_counter()
    for _ in range(reps):
        o.compute(*pos)
    sec = time.perf_counter() - t0
    return dict(value=reps / sec, unit="steps/s", cores=1, kind="port",
                sample="%d force evaluations (no list rebuild) of %s by the C restatement of the "
                       "reference (oracle/oracle.c), one thread" % (reps, system.name))


def _trajectory(s, args, dev, n_frames=None):
    """Device frames of the synthetic trajectory: every molecule (and mobile ion) oscillates
    rigidly about its start, x(t) = x0 + a sin(2 pi t / P), a ~ N(0, amp) per axis, so the
    skin test fires and the list is rebuilt repeatedly while molecules never collide."""
    import numpy as np
    import torch
    from md_sim_amd import abi
    x, y, z, cx, cy, cz = abi.positions(s)
    nm, nr = s.nmove, s.nrig
    rng = np.random.default_rng(12345)
    mol_a = np.clip(rng.normal(0.0, args.amplitude, size=(nr, 3)), -3 * args.amplitude, 3 * args.amplitude)
    site_a = np.zeros((nm, 3))
    site_a[:3 * nr] = np.repeat(mol_a, 3, axis=0)
    site_a[3 * nr:] = np.clip(rng.normal(0.0, args.amplitude, size=(nm - 3 * nr, 3)), -3 * args.amplitude,
                              3 * args.amplitude)
    base = torch.tensor(np.stack([x[:nm], y[:nm], z[:nm]]), dtype=torch.float64, device=dev)
    amp = torch.tensor(site_a.T.copy(), dtype=torch.float64, device=dev)
    cbase = torch.tensor(np.stack([cx, cy, cz]), dtype=torch.float64, device=dev)
    camp = torch.tensor(mol_a.T.copy(), dtype=torch.float64, device=dev)
    n_frames = n_frames or args.warmup + args.steps
    frames = torch.empty((n_frames, 3, nm), dtype=torch.float64, device=dev)
    cframes = torch.empty((n_frames, 3, max(nr, 1)), dtype=torch.float64, device=dev)
    for t in range(n_frames):
        ph = float(np.sin(2.0 * np.pi * t / args.period))
        frames[t] = base + ph * amp
        cframes[t, :, :nr] = cbase + ph * camp
    torch.cuda.synchronize()
    return frames, cframes


def _run(eng, frames, cframes, args):
    """Warm-up, then the timed pass (graph replays) and a second pass with per-stage events.
    Returns (seconds, outputs after warm-up, outputs at the end, per-stage means)."""
    import torch

    def step(t):
        f, c = frames[t], cframes[t]
        eng.compute_device(f[0].data_ptr(), f[1].data_ptr(), f[2].data_ptr(),
                           c[0].data_ptr(), c[1].data_ptr(), c[2].data_ptr())

    n_frames = args.warmup + args.steps
    for t in range(args.warmup):
        step(t)
    first = eng.read_outputs()
    torch.cuda.synchronize()
    eng.reset_stage_times()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for t in range(args.warmup, n_frames):
        step(t)
    last = eng.read_outputs()
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    step_gpu_ms = eng.stage_times()["step"]
    # kernel durations: the same trajectory once more with HIP events between the stages
    # (direct launches; events inside a replayed graph cannot be timed)
    eng.set_stage_timing(True)
    eng.reset_stage_times()
    for t in range(args.warmup, n_frames):
        step(t)
    eng.read_outputs()
    stages = eng.stage_times()
    eng.set_stage_timing(False)
    stages["step_graph"] = step_gpu_ms
    return sec, first, last, stages


def replica_bench(args, prec):
    """Replica batch: independent interface systems of uneven size (512-4096 waters on
    256-2048 frozen ions, each its own box, list and k-set, seeded per replica) stepped by one
    batch (one launch per stage over all of them, one graph per step), against the same
    systems stepped as independent handles (one graph per system per step, each on its own
    stream).  Both run the same oscillating trajectories; each picks its launch shapes with
    its autotuner first."""
    import numpy as np
    import torch
    from md_sim_amd import systems
    from md_sim_amd.engine import ForceBatch, ForceEngine
    dev = torch.device("cuda:0")
    syss = systems.replica_set(args.replicas, seed=3000)
    nf = args.replica_warmup + args.replica_steps
    traj = [_trajectory(s, args, dev, nf) for s in syss]
    engines = [ForceEngine(s, precision=prec) for s in syss]

    def ptrs(t):
        return [(f[t][0].data_ptr(), f[t][1].data_ptr(), f[t][2].data_ptr(),
                 c[t][0].data_ptr(), c[t][1].data_ptr(), c[t][2].data_ptr()) for f, c in traj]

    # independent handles
    for e, p in zip(engines, ptrs(0)):
        e.compute_device(*p)
        e.read_outputs()
        e.autotune(reps=5)
    for t in range(args.replica_warmup):
        for e, p in zip(engines, ptrs(t)):
            e.compute_device(*p)
    for e in engines:
        e.read_outputs()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for t in range(args.replica_warmup, nf):
        for e, p in zip(engines, ptrs(t)):
            e.compute_device(*p)
    single = [e.read_outputs() for e in engines]
    torch.cuda.synchronize()
    sec_single = time.perf_counter() - t0

    # one batch
    b = ForceBatch(engines)
    b.compute_device(ptrs(0))
    b.read_outputs()
    shapes = b.autotune(reps=5)
    for t in range(args.replica_warmup):
        b.compute_device(ptrs(t))
    b.read_outputs()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for t in range(args.replica_warmup, nf):
        b.compute_device(ptrs(t))
    batch = b.read_outputs()
    torch.cuda.synchronize()
    sec_batch = time.perf_counter() - t0
    b.close()
    dif = max(abs(o["e_conf"] - r["e_conf"]) / abs(r["e_conf"]) for o, r in zip(batch, single))
    ferr = max(float(np.sqrt(((o["f_site"] - r["f_site"]) ** 2).mean() / (r["f_site"] ** 2).mean()))
               for o, r in zip(batch, single))
    nw = [s.nrig for s in syss]
    ni = [s.nwall for s in syss]
    out = {
        "workload": "%d replicas: %d-%d SPC/E waters on %d-%d frozen ions (%d waters, %d ions in all), "
                    "T 280-360 K, own box, list and k-set each" % (len(syss), min(nw), max(nw), min(ni), max(ni),
                                                                   sum(nw), sum(ni)),
        "steps": args.replica_steps,
        "batch_steps_per_s": args.replica_steps / sec_batch,
        "batch_system_steps_per_s": args.replica_steps * len(syss) / sec_batch,
        "independent_handles_steps_per_s": args.replica_steps / sec_single,
        "batch_speedup_vs_independent_handles": sec_single / sec_batch,
        "batch_launch_shapes": shapes,
        "max_e_conf_rel_diff_batch_vs_handles": dif,
        "max_force_rms_rel_diff_batch_vs_handles": ferr,
    }
    for e in engines:
        e.close()
    del traj
    torch.cuda.empty_cache()
    return out


def mts_bench(args, workload):
    """Surrogate k-space force with multiple time steps against the plain FP32 form, on a
    librating water trajectory (every molecule librates about its COM at H speeds of water
    at 300 K with 2 fs steps while the COM oscillates).  Each run: `mts_train` full-sum
    steps that train the network, 8 untimed steps, then `mts_steps` timed steps on
    device-resident positions; a second pass on the host gives the force RMS error of every
    step against the plain FP32 form, relative to the RMS site force.  Ablation: the same
    scheme with the held force instead of the network.  The judge's tolerance is set to 0.05
    so that the scheme stays on; its judged errors are reported."""
    import numpy as np
    import torch
    from md_sim_amd import abi, systems
    from md_sim_amd.engine import ForceEngine
    dev = torch.device("cuda:0")
    s = systems.config(workload)
    n_warm = args.mts_train + 8
    n_all = n_warm + args.mts_steps
    frames = systems.libration_frames(s, n_all, seed=7)
    nm, nr = s.nmove, s.nrig
    dframes = [[torch.tensor(a[:nm] if k < 3 else a, dtype=torch.float64, device=dev) for k, a in enumerate(f)]
               for f in frames]
    torch.cuda.synchronize()

    def timed(e):
        for t in range(n_warm):
            e.compute_device(*[a.data_ptr() for a in dframes[t]])
        e.read_outputs()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for t in range(n_warm, n_all):
            e.compute_device(*[a.data_ptr() for a in dframes[t]])
        e.read_outputs()
        torch.cuda.synchronize()
        return args.mts_steps / (time.perf_counter() - t0)

    e = ForceEngine(s, precision=abi.FP32_ACC64, virial=False)
    e.force_get(*frames[0])
    shapes = e.autotune(reps=5)
    plain_sps = timed(e)
    ref = [e.force_get(*f) for f in frames[n_warm:]]
    e.close()
    runs = []
    for interval, model in ((4, "network"), (2, "network"), (4, "held")):
        cfg = dict(interval=interval, model=model, train_frames=args.mts_train, train_iters=2, lr=1e-3, tol=0.05)
        e = ForceEngine(s, precision=abi.FP32_ACC64, virial=False)
        e.force_get(*frames[0])
        e.set_launch_shapes(**shapes)
        e.mts_enable(**cfg)
        sps = timed(e)
        st = e.mts_stats()
        e.close()
        e = ForceEngine(s, precision=abi.FP32_ACC64, virial=False)
        e.set_launch_shapes(**shapes)
        e.mts_enable(**cfg)
        outs = [e.force_get(*f) for f in frames]
        e.close()
        errs = [float(np.sqrt(((o["f_site"] - r["f_site"]) ** 2).mean() / (r["f_site"] ** 2).mean()))
                for o, r in zip(outs[n_warm:], ref)]
        runs.append({"interval": interval, "model": model, "steps_per_s": sps, "speedup_vs_plain_fp32": sps / plain_sps,
                     "force_rms_rel_err_mean": float(np.mean(errs)), "force_rms_rel_err_max": float(np.max(errs)),
                     "judged_err_mean": st["err_mean"], "judged_err_max": st["err_max"],
                     "fallbacks": st["n_fallback"]})
    del dframes
    torch.cuda.empty_cache()
    return {"workload": "%s: %d SPC/E waters + %d frozen ions, k-index %d, FP32 with FP64 accumulation"
                        % (workload, nr, s.nwall, s.kmax),
            "plain_fp32_steps_per_s": plain_sps, "runs": runs}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from md_sim_amd import abi, systems
    from md_sim_amd.engine import ForceEngine

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda:0")
    s = systems.config(args.workload)
    nm = s.nmove
    frames, cframes = _trajectory(s, args, dev)
    prec = abi.FP32_ACC64 if args.precision == "fp32" else abi.FP64
    # the reference computes no virial; it is an opt-in extra of this engine
    eng = ForceEngine(s, precision=prec, virial=args.virial)
    shapes = eng.autotune(reps=10) if not args.no_autotune else None
    sec, first, last, stages = _run(eng, frames, cframes, args)
    kv, ak, _ = eng.kvectors()
    on_mma = eng.reciprocal_on_matrix_cores()
    eng.close()

    # the FP64 form on the same trajectory: its throughput, and the reduced-precision
    # form's force error against it on the last frame
    fp64 = None
    if prec != abi.FP64:
        e64 = ForceEngine(s, precision=abi.FP64, virial=args.virial)
        shapes64 = e64.autotune(reps=10) if not args.no_autotune else None
        sec64, _, last64, stages64 = _run(e64, frames, cframes, args)
        e64.close()
        f32, f64 = last["f_site"], last64["f_site"]
        err = float(np.sqrt(((f32 - f64) ** 2).mean()) / np.sqrt((f64 ** 2).mean()))
        fp64 = {"steps_per_s": args.steps / sec64, "ms_per_step": sec64 / args.steps * 1e3,
                "stages_ms": stages64, "launch_shapes": shapes64,
                "force_rms_rel_err_vs_fp64": err,
                "e_conf_rel_diff_vs_fp64": abs(last["e_conf"] - last64["e_conf"]) / abs(last64["e_conf"])}

    n_k = len(ak)
    n_rows = len(set(map(tuple, kv[:, :2].tolist())))
    flops = 18.0 * n_k * nm + 8.0 * n_rows * nm
    t_recip_ms = stages["sf"] + stages["sf_reduce"] + stages["backproj"]
    achieved = flops / (t_recip_ms * 1e-3) / 1e12
    peak = PEAK_TFLOPS[args.precision]
    traffic = None
    if os.path.exists(PROFILE_TRAFFIC):
        try:
            traffic = json.load(open(PROFILE_TRAFFIC)).get(args.precision)
        except Exception:
            traffic = None

    cpu = cpu1 = None
    if not args.no_cpu_baseline:
        cores = max(1, min(args.cpu_cores, os.cpu_count() or 1))
        try:
            cpu = cpu_baseline(s, 4 * args.cpu_reps if cores > 1 else args.cpu_reps, cores)
        except Exception as e:  # the GPU figure stands on its own
            cpu = {"error": str(e)[:200]}
        if cores > 1:
            try:
                cpu1 = cpu_baseline(s, args.cpu_reps, 1)
            except Exception as e:
                cpu1 = {"error": str(e)[:200]}

    prec_txt = "FP64" if args.precision == "fp64" else "FP32 arithmetic with FP64 accumulation"
    out = {
        "metric": METRIC,
        "value": args.steps / sec,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else (
            "f32 (f64 accumulation; reciprocal-sum products as FP16 hi/lo pairs on the matrix cores)"
            if on_mma else "f32 (f64 accumulation)"),
        "data": "synthetic (seeded lattice water on a rock-salt slab, oscillating trajectory)",
        "config": {
            "workload": "%s: %d SPC/E waters + %d frozen ions, box %.2f x %.2f x %.2f A, rc %.1f A, "
                        "skin %.1f A, k-index %d (%d k-vectors), %s; one force evaluation per step"
                        % (s.name, s.nrig, s.nwall, s.box[0], s.box[1], s.box[2], s.rcut, s.skin, s.kmax, n_k,
                           prec_txt),
            "global_batch": 1,
            "parallelism": "single GPU",
        },
        "roofline": {
            "kernel": "reciprocal sum (k_sf + k_sf_reduce + k_bp)",
            "bound": "mfma",
            "pipe": "FP64 vector ALU (MI355X FP64 vector peak = FP64 matrix peak)" if args.precision == "fp64"
                    else ("FP16 matrix cores (v_mfma_f32_16x16x32_f16), each FP32 operand split into FP16 hi + lo "
                          "parts, 3 matrix products per FP32 product; peak is MI355X's FP32 peak, the accuracy "
                          "class the results keep" if on_mma
                          else "FP32 vector ALU (peak assumes packed FP32, which this build does not emit)"),
            "achieved": achieved,
            "peak": peak,
            "unit": "TFLOP/s",
            "frac": achieved / peak,
            "traffic": traffic,
            "algorithmic_flop_per_launch": flops,
            "mean_ms": t_recip_ms,
        },
        "cpu_baseline": cpu,
        "cpu_baseline_one_core": cpu1,
        "stages_ms": stages,
        "launch_shapes": shapes,
        "list_builds_timed": last["n_list_builds"] - first["n_list_builds"],
        "e_conf_last": last["e_conf"],
    }
    if fp64 is not None:
        out["fp64"] = fp64
    if args.replicas > 0:
        out["replica_batch"] = replica_bench(args, prec)
    mts = [w for w in args.mts_workloads.split(",") if w]
    if mts:
        out["surrogate_mts"] = [mts_bench(args, w) for w in mts]
    if cpu and "value" in cpu:
        out["speedup_vs_cpu"] = out["value"] / cpu["value"]
        if fp64 is not None:
            out["fp64"]["speedup_vs_cpu"] = fp64["steps_per_s"] / cpu["value"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
