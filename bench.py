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
The headline ``value`` is the FP64 form, the reference's arithmetic.  The FP32 form (pair
and structure-factor arithmetic in FP32 with FP64 accumulation; reciprocal-sum products as
FP16 hi/lo pairs on the matrix cores) runs the same trajectory beside it and is reported
under ``fp32`` with its force RMS error against FP64.

``roofline`` is the FP64 step's dominant kernel; ``rooflines`` holds every hot-path kernel
of both forms.  Algorithmic work per launch (one launch per step):
* pair kernel: 44 flop per Coulomb site pair within rc, 13 per Lennard-Jones and 17 per
  Buckingham lead pair within rc, 17 per listed lead pair (minimum image, r^2, cutoff
  test), with the counts taken from the pair list and the last frame's positions;
* structure factor: 8 flop per (k-vector, moving site) + 8 per ((kl, km) row, moving site);
* back-projection: 10 flop per (k-vector, moving site);
* list build: 24 B per site (positions) + 8 B per site (cell keys) + 4 B per listed pair.
Each is divided by the kernel's mean duration, measured in a second pass over the same steps
with direct launches (the timed pass replays a captured HIP graph, inside which events
cannot be timed): the pair and reciprocal-sum kernels by start and stop events that their
own dispatch records on the engine's stream (hipExtLaunchKernelGGL), the list build (six
kernels) by HIP events around the stage in a third pass that rebuilds the list on every
step.  ``traffic`` is the HBM
bytes per launch that rocprofv3's FETCH_SIZE and WRITE_SIZE counters give for the kernel
(profiles/traffic_<workload>.json), when that file covers the workload.

``cpu_baseline`` times the reference's own force path, built from its Fortran sources with
its own flags (-O2 -ffast-math, src/Makefile:2) into oracle/_ref, on a bounded sample of
force evaluations of the same system: on the host cores this process may use (MPI ranks,
the reference's own decomposition; at most 16, one GPU's share of the host) and on one
core.  Without that build the CPU oracle (oracle/oracle.c) is timed instead.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "MD steps/s (forces+Ewald) at 16k H2O; surrogate-MTS speedup at fixed force err"
PEAK = {"fp64": 78.6, "fp32": 157.3, "fp16_mfma": 2500.0, "hbm": 8000.0}   # TFLOP/s, GB/s
MAX_CPU_RANKS = 16


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="slab16k")
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32 form beside the FP64 headline")
    ap.add_argument("--build-steps", type=int, default=40, help="steps of the pass that rebuilds the list every step")
    ap.add_argument("--cpu-reps", type=int, default=3, help="force evaluations in the one-core CPU sample")
    ap.add_argument("--cpu-reps-all", type=int, default=24, help="force evaluations in the all-core CPU sample")
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
    ap.add_argument("--mts-tol", type=float, default=1e-3,
                    help="the scheme falls back to the full sum when its judged force RMS error exceeds this")
    ap.add_argument("--mts-offline-frames", type=int, default=400,
                    help="frames per training trajectory of the off-line trained network (0: skip it)")
    return ap.parse_args()


def host_cpu():
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return model, os.cpu_count() or 1, avail


def cpu_baseline(system, reps: int, cores: int = 1):
    """Reference force path on the host: force evaluations per second, on `cores` MPI ranks
    (the reference's own decomposition: leads and k-vectors over ranks, force_sum's
    all-reduce), else on one core."""
    from oracle import refrun
    model, n_all, n_avail = host_cpu()
    host = "%s; %d logical CPUs, %d available to this process" % (model, n_all, n_avail)
    for fast in (True, False):
        if refrun.ref_available(fast=fast):
            nproc = cores if cores > 1 and refrun.mpi_available() else 1
            d = refrun.run_reference(system, nrep=reps, nproc=nproc, timeout=1200, fast=fast)
            n, sec = d["_timing"]
            flags = "-O2 -ffast-math (the reference's own flags)" if fast else "-O2 -ffp-contract=off"
            return dict(value=n / sec, unit="steps/s", cores=nproc, kind="reference", host=host,
                        sample="%d force evaluations (no list rebuild) of %s by the reference's own "
                               "force_real/force_fouri/force_intra/force_sum, built from its Fortran "
                               "sources with amdflang %s, %d MPI rank%s"
                               % (int(n), system.name, flags, nproc, "s" if nproc > 1 else ""))
    from md_sim_amd import abi
    from oracle import pyoracle
    o = pyoracle.Oracle(system)
    pos = abi.positions(system)
    o.compute(*pos)
    t0 = time.perf_counter()
    for _ in range(reps):
        o.compute(*pos)
    sec = time.perf_counter() - t0
    return dict(value=reps / sec, unit="steps/s", cores=1, kind="port", host=host,
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
    """Warm-up, the timed pass (graph replays), a pass with per-stage events and a pass
    with per-stage events and a list build on every step."""
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
    eng.set_stage_timing(True)
    eng.reset_stage_times()
    for t in range(args.warmup, n_frames):
        step(t)
    eng.read_outputs()
    stages = eng.stage_times()
    # the list build: a build on every step
    eng.set_stage_timing(True, rebuild_every_step=True)
    eng.reset_stage_times()
    nb = min(args.build_steps, args.steps)
    for t in range(args.warmup, args.warmup + nb):
        step(t)
    eng.read_outputs()
    builds = eng.stage_times()
    eng.set_stage_timing(False)
    stages["step_graph"] = step_gpu_ms
    return dict(sec=sec, first=first, last=last, stages=stages, build_stages=builds, build_steps=nb)


def pair_counts(eng, s, xyz):
    """Site-pair counts of the pair kernel's algorithmic work on the frame `xyz` (moving
    sites, 3 x n_move) with the list as it stands: Coulomb site pairs, LJ and Buckingham
    lead pairs within rc (src/forcem.f90:288-403,418-498; the cutoff is on the lead
    sites), and listed lead pairs."""
    import numpy as np
    from md_sim_amd import abi
    i_mm, j_mm, i_mw, j_mw = eng.pairs()
    x, y, z, _, _, _ = abi.positions(s)
    p = [np.array(a, np.float64) for a in (x, y, z)]
    for k in range(3):
        p[k][:s.nmove] = xyz[k]

    def rsq(i, j):
        r2 = np.zeros(len(i))
        for k in range(3):
            d = p[k][i] - p[k][j]
            if k < 2 or s.z_periodic:
                d -= s.box[k] * np.rint(d / s.box[k])
            r2 += d * d
        return r2

    st = s.site_types() - 1
    nsi = s.nsi()
    ijtype, kind = s.vdw_tables()[:2]
    rc2 = s.rcut * s.rcut
    a, b = i_mm[rsq(i_mm, j_mm) <= rc2], j_mm[rsq(i_mm, j_mm) <= rc2]
    c, d = i_mw[rsq(i_mw, j_mw) <= rc2], j_mw[rsq(i_mw, j_mw) <= rc2]
    n_coul = int((nsi[st[a]] * nsi[st[b]]).sum() + nsi[st[c]].sum())
    kinds = np.concatenate([kind[ijtype[st[a], st[b]] - 1], kind[ijtype[st[c], st[d]] - 1]])
    return dict(listed_moving=len(i_mm), listed_frozen=len(i_mw), within_rc_moving=len(a), within_rc_frozen=len(c),
                coulomb_site_pairs=n_coul, lj_pairs=int((kinds == 1).sum()), buckingham_pairs=int((kinds == 2).sum()))


def pair_flops(pc):
    return (44.0 * pc["coulomb_site_pairs"] + 13.0 * pc["lj_pairs"] + 17.0 * pc["buckingham_pairs"]
            + 17.0 * (pc["listed_moving"] + pc["listed_frozen"]))


def load_traffic(workload):
    path = os.path.join(ROOT, "profiles", "traffic_%s.json" % workload)
    try:
        with open(path) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return None


def traffic_of(traffic, form, pattern):
    """HBM bytes per launch of the kernels matching `pattern` in the profile of `form`, summed."""
    if not traffic or form not in traffic:
        return None
    ks = [v["bytes"] for k, v in traffic[form]["kernels"].items() if re.search(pattern, k)]
    return float(sum(ks)) if ks else None


def roofline(kernel, bound, pipe, work, ms, peak, unit, traffic=None, **extra):
    scale = 1e12 if unit == "TFLOP/s" else 1e9
    achieved = work / (ms * 1e-3) / scale if ms > 0 else 0.0
    r = {"kernel": kernel, "bound": bound, "pipe": pipe, "achieved": achieved, "peak": peak, "unit": unit,
         "frac": achieved / peak, "traffic": traffic,
         ("algorithmic_flop_per_launch" if unit == "TFLOP/s" else "algorithmic_bytes_per_launch"): work,
         "mean_ms": ms}
    r.update(extra)
    return r


def form_rooflines(form, run, s, pc, n_k, n_rows, kernels, traffic):
    """Rooflines of one form's kernels from its per-stage means."""
    st, bs = run["stages"], run["build_stages"]
    # kernel means from the kernels' own dispatch events, the stage means where those are absent
    km = {k: (st.get(kt) or st[stg]) for k, kt, stg in (("pair", "k_pair", "pair"), ("sf", "k_sf", "sf"),
                                                        ("sf_reduce", "k_sf_reduce", "sf_reduce"),
                                                        ("backproj", "k_bp", "backproj"))}
    nm = s.nmove
    f64 = form == "fp64"
    vpeak = PEAK[form]
    out = {}
    pk = "mdf::k_pair<%s," % ("double" if f64 else "float")
    out["pair"] = roofline(
        "k_pair<%s>" % ("double" if f64 else "float"), "valu",
        "%s vector ALU" % ("FP64" if f64 else "FP32 (no packed FP32 in this build)"),
        pair_flops(pc), km["pair"], vpeak, "TFLOP/s", traffic_of(traffic, form, re.escape(pk)))
    sf_work = 8.0 * n_k * nm + 8.0 * n_rows * nm
    bp_work = 10.0 * n_k * nm
    if kernels == "mma64":
        sf_pipe = bp_pipe = "FP64 matrix cores (v_mfma_f64_16x16x4_f64)"
        sf_pat, bp_pat, red_pat = r"k_sf64<", r"k_bp64<", r"k_sf64_reduce"
        names = ("k_sf64", "k_bp64")
        mpeak, bound = PEAK["fp64"], "mfma"
    elif kernels == "mma16":
        sf_pipe = bp_pipe = ("FP16 matrix cores (v_mfma_f32_16x16x32_f16), each FP32 operand split into FP16 "
                             "hi + lo parts, 3 products per FP32 product")
        sf_pat, bp_pat, red_pat = r"k_sf_mma<", r"k_bp_mma<", r"k_sf_reduce_rows"
        names = ("k_sf_mma", "k_bp_mma")
        mpeak, bound = PEAK["fp16_mfma"], "mfma"
    else:
        sf_pipe = bp_pipe = "%s vector ALU" % ("FP64" if f64 else "FP32")
        sf_pat, bp_pat, red_pat = r"k_sf<", r"k_bp<", r"k_sf_reduce\("
        names = ("k_sf", "k_bp")
        mpeak, bound = vpeak, "valu"
    extra_sf, extra_bp = {}, {}
    if kernels == "mma16":   # the accuracy class the results keep
        extra_sf["frac_of_fp32_peak"] = sf_work / (km["sf"] * 1e-3) / 1e12 / PEAK["fp32"] if km["sf"] > 0 else 0.0
        extra_bp["frac_of_fp32_peak"] = bp_work / (km["backproj"] * 1e-3) / 1e12 / PEAK["fp32"] \
            if km["backproj"] > 0 else 0.0
    out["structure_factor"] = roofline(names[0], bound, sf_pipe, sf_work, km["sf"], mpeak, "TFLOP/s",
                                       traffic_of(traffic, form, sf_pat), **extra_sf)
    out["back_projection"] = roofline(names[1], bound, bp_pipe, bp_work, km["backproj"], mpeak, "TFLOP/s",
                                      traffic_of(traffic, form, bp_pat), **extra_bp)
    t_rec = km["sf"] + km["sf_reduce"] + km["backproj"]
    tr = [traffic_of(traffic, form, p) for p in (sf_pat, red_pat, bp_pat)]
    out["reciprocal_stage"] = roofline(
        "%s + reduce + %s" % names, bound, sf_pipe, sf_work + bp_work, t_rec, mpeak, "TFLOP/s",
        sum(tr) if all(v is not None for v in tr) else None)
    listed = pc["listed_moving"] + pc["listed_frozen"]
    out["list_build"] = roofline(
        "k_cell_count + k_cell_scan_build + k_cell_fill + k_cell_order + k_build", "hbm", "HBM",
        32.0 * s.ntota + 4.0 * listed, bs["build"], PEAK["hbm"], "GB/s",
        traffic_of(traffic, form, r"k_cell_count|k_cell_scan_build|k_cell_fill|k_cell_order|k_build\("),
        steps=run["build_steps"], note="timed in a pass that rebuilds the list on every step")
    return out


def recip_kind(eng, form):
    if not eng.reciprocal_on_matrix_cores():
        return "valu"
    return "mma64" if form == "fp64" else "mma16"


def form_bench(s, prec, form, args, frames, cframes):
    from md_sim_amd.engine import ForceEngine
    eng = ForceEngine(s, precision=prec, virial=args.virial)
    shapes = eng.autotune(reps=10) if not args.no_autotune else None
    run = _run(eng, frames, cframes, args)
    run["shapes"] = shapes
    run["kernels"] = recip_kind(eng, form)
    # the list as the last step of the build pass built it, on that step's positions
    run["pairs"] = pair_counts(eng, s, frames[args.warmup + run["build_steps"] - 1].cpu().numpy())
    kv, ak, _ = eng.kvectors()
    run["n_k"] = len(ak)
    run["n_rows"] = len(set(map(tuple, kv[:, :2].tolist())))
    eng.close()
    return run


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
        "precision": "fp64" if prec == 0 else "fp32 (f64 accumulation)",
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


def surrogate_offline(args):
    """Network weights trained off line on exact Ewald k-space forces of the interface and
    replica workloads (16k waters on 6144 ions, and four replicas of the 64-replica set),
    with the engine's own training kernels: each source system runs a librating trajectory
    whose every step is a full sum that trains the network (Adam, two iterations per frame
    on the change of the k-space force from the frame before), the weights carried from one
    source to the next.  Returns the weights and a description of the frames."""
    from md_sim_amd import abi, systems
    from md_sim_amd.engine import ForceEngine
    n = args.mts_offline_frames
    srcs = [(systems.config("slab16k"), 11), (systems.config("slab16k"), 12)]
    srcs += [(s, 20 + k) for k, s in enumerate(systems.replica_set(64, seed=3000)[:64:16])]
    w, frames_total, sites_total = None, 0, 0
    for s, seed in srcs:
        frames = systems.libration_frames(s, n, seed=seed)
        e = ForceEngine(s, precision=abi.FP32_ACC64, virial=False)
        e.force_get(*frames[0])
        e.mts_enable(interval=4, model="network", train_frames=n + 8, train_iters=2, lr=1e-3, tol=1.0)
        if w is not None:
            e.set_mts_weights(w)
        for f in frames:
            e.force_get(*f)
        st = e.mts_stats()
        assert st["n_surrogate"] == 0, st   # every step a full sum
        w = e.mts_weights()
        e.close()
        frames_total += len(frames)
        sites_total += len(frames) * s.nmove
        del frames
    return w, "%d frames (%.1f M site samples): 2 librating trajectories of slab16k and 1 of each of 4 " \
              "replicas of the 64-replica set, %d frames each" % (frames_total, sites_total / 1e6, n)


def mts_bench(args, workload, offline=None):
    """Surrogate k-space force with multiple time steps against the plain FP32 form, on a
    librating water trajectory (every molecule librates about its COM at H speeds of water
    at 300 K with 2 fs steps while the COM oscillates).  Each run: `mts_train` full-sum
    steps that train the network, 8 untimed steps, then `mts_steps` timed steps on
    device-resident positions; a second pass on the host gives the force RMS error of every
    step against the plain FP32 form, relative to the RMS site force.  The scheme falls back
    to the full sum on the steps after a judged error above `mts_tol`.  Ablation: the same
    scheme with the held force instead of the network."""
    import numpy as np
    import torch
    from md_sim_amd import abi, systems
    from md_sim_amd.engine import ForceEngine
    dev = torch.device("cuda:0")
    s = systems.config(workload)
    n_warm = args.mts_train + 8
    n_all = n_warm + args.mts_steps
    frames = systems.libration_frames(s, n_all, seed=7)
    nm = s.nmove
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
    variants = [(4, "network", None), (2, "network", None), (4, "held", None)]
    if offline is not None:
        variants += [(4, "network", offline[0]), (2, "network", offline[0])]
    for interval, model, w0 in variants:
        # with off-line weights: no training on this trajectory, the first full sums only
        # start the correction history
        cfg = dict(interval=interval, model=model, train_frames=args.mts_train, train_iters=2 if w0 is None else 0,
                   lr=1e-3, tol=args.mts_tol)
        e = ForceEngine(s, precision=abi.FP32_ACC64, virial=False)
        e.force_get(*frames[0])
        e.set_launch_shapes(**shapes)
        e.mts_enable(**cfg)
        if w0 is not None:
            e.set_mts_weights(w0)
        sps = timed(e)
        st = e.mts_stats()
        e.close()
        e = ForceEngine(s, precision=abi.FP32_ACC64, virial=False)
        e.set_launch_shapes(**shapes)
        e.mts_enable(**cfg)
        if w0 is not None:
            e.set_mts_weights(w0)
        outs = [e.force_get(*f) for f in frames]
        e.close()
        errs = [float(np.sqrt(((o["f_site"] - r["f_site"]) ** 2).mean() / (r["f_site"] ** 2).mean()))
                for o, r in zip(outs[n_warm:], ref)]
        runs.append({"interval": interval, "model": model, "tol": args.mts_tol,
                     "training": "on line, this trajectory" if w0 is None and model == "network" else
                                 ("off line: " + offline[1] if w0 is not None else "none"),
                     "steps_per_s": sps,
                     "speedup_vs_plain_fp32": sps / plain_sps,
                     "force_rms_rel_err_mean": float(np.mean(errs)), "force_rms_rel_err_max": float(np.max(errs)),
                     "judged_err_mean": st["err_mean"], "judged_err_max": st["err_max"],
                     "fallbacks": st["n_fallback"]})
    del dframes
    torch.cuda.empty_cache()
    return {"workload": "%s: %d SPC/E waters + %d frozen ions, k-index %d, FP32 with FP64 accumulation"
                        % (workload, s.nrig, s.nwall, s.kmax),
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

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda:0")
    s = systems.config(args.workload)
    frames, cframes = _trajectory(s, args, dev)
    traffic = load_traffic(args.workload)

    r64 = form_bench(s, abi.FP64, "fp64", args, frames, cframes)
    roof64 = form_rooflines("fp64", r64, s, r64["pairs"], r64["n_k"], r64["n_rows"], r64["kernels"], traffic)
    dominant = max(("pair", "structure_factor", "back_projection"), key=lambda k: roof64[k]["mean_ms"])

    fp32 = None
    if not args.no_fp32:
        r32 = form_bench(s, abi.FP32_ACC64, "fp32", args, frames, cframes)
        f32, f64 = r32["last"]["f_site"], r64["last"]["f_site"]
        fp32 = {
            "steps_per_s": args.steps / r32["sec"], "ms_per_step": r32["sec"] / args.steps * 1e3,
            "dtype": "f32 (f64 accumulation; reciprocal-sum products as FP16 hi/lo pairs on the matrix cores)"
                     if r32["kernels"] == "mma16" else "f32 (f64 accumulation)",
            "stages_ms": r32["stages"], "build_stages_ms": r32["build_stages"], "launch_shapes": r32["shapes"],
            "force_rms_rel_err_vs_fp64": float(np.sqrt(((f32 - f64) ** 2).mean()) / np.sqrt((f64 ** 2).mean())),
            "e_conf_rel_diff_vs_fp64": abs(r32["last"]["e_conf"] - r64["last"]["e_conf"]) / abs(r64["last"]["e_conf"]),
            "rooflines": form_rooflines("fp32", r32, s, r32["pairs"], r32["n_k"], r32["n_rows"], r32["kernels"],
                                        traffic),
        }

    cpu = cpu1 = None
    if not args.no_cpu_baseline:
        _, _, avail = host_cpu()
        cores = max(1, min(MAX_CPU_RANKS, avail))
        try:
            cpu = cpu_baseline(s, args.cpu_reps_all if cores > 1 else args.cpu_reps, cores)
        except Exception as e:  # the GPU figure stands on its own
            cpu = {"error": str(e)[:200]}
        if cores > 1:
            try:
                cpu1 = cpu_baseline(s, args.cpu_reps, 1)
            except Exception as e:
                cpu1 = {"error": str(e)[:200]}

    sec = r64["sec"]
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
        "dtype": "f64",
        "data": "synthetic (seeded lattice water on a rock-salt slab, oscillating trajectory)",
        "config": {
            "workload": "%s: %d SPC/E waters + %d frozen ions, box %.2f x %.2f x %.2f A, rc %.1f A, "
                        "skin %.1f A, k-index %d (%d k-vectors), FP64; one force evaluation per step"
                        % (s.name, s.nrig, s.nwall, s.box[0], s.box[1], s.box[2], s.rcut, s.skin, s.kmax,
                           r64["n_k"]),
            "global_batch": 1,
            "parallelism": "single GPU",
        },
        "roofline": dict(roof64[dominant], stage=dominant),
        "rooflines": {"fp64": roof64},
        "pair_counts": r64["pairs"],
        "cpu_baseline": cpu,
        "cpu_baseline_one_core": cpu1,
        "stages_ms": r64["stages"],
        "build_stages_ms": r64["build_stages"],
        "launch_shapes": r64["shapes"],
        "list_builds_timed": r64["last"]["n_list_builds"] - r64["first"]["n_list_builds"],
        "e_conf_last": r64["last"]["e_conf"],
    }
    if fp32 is not None:
        out["fp32"] = fp32
        out["rooflines"]["fp32"] = fp32.pop("rooflines")
    if cpu and "value" in cpu:
        out["speedup_vs_cpu"] = out["value"] / cpu["value"]
        if fp32 is not None:
            out["fp32"]["speedup_vs_cpu"] = fp32["steps_per_s"] / cpu["value"]
    del frames, cframes
    torch.cuda.empty_cache()
    if args.replicas > 0:
        out["replica_batch"] = replica_bench(args, abi.FP32_ACC64)
    mts = [w for w in args.mts_workloads.split(",") if w]
    if mts:
        off = surrogate_offline(args) if args.mts_offline_frames > 0 else None
        out["surrogate_mts"] = [mts_bench(args, w, off) for w in mts]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
