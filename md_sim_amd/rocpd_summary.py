"""Summaries of rocprofv3 databases (the rocpd SQLite files its runs write).

``python -m md_sim_amd.rocpd_summary kernels DB [--match REGEX] [--top N]``
    per kernel: calls, mean / min / max duration and total time, as a markdown table;

``python -m md_sim_amd.rocpd_summary traffic --fetch DB --write DB [--copy-mib 256]
[--match REGEX] [--json OUT [--key KEY]]``
    HBM bytes per dispatch of each kernel from a ``--pmc FETCH_SIZE`` run and a
    ``--pmc WRITE_SIZE`` run (KiB per dispatch, summed over the counter's hardware instances,
    averaged over the kernel's dispatches).  The device-to-device copy that
    ``md_sim_amd.profile_step`` runs last moves a known number of bytes, so each counter is
    scaled by (true bytes) / (what the counter reads for that copy).  --json writes the
    figures to OUT, under KEY when given (other keys of an existing OUT are kept).

``python -m md_sim_amd.rocpd_summary timeline DB [--last KERNEL] [--first KERNEL] [--forked]``
    the steps of a kernel trace of graph replays, each from the dispatch after one KERNEL
    (default ``k_final``) to the next: step spans, split into steps with and without a list
    build (``k_build`` longer than 20 us), and the kernel-by-kernel timeline (queue, start,
    end, duration in us from the step's first dispatch) of the median step of each kind;
    --forked keeps the steps whose kernels ran on more than one queue (forked graph replays);

``python -m md_sim_amd.rocpd_summary counters DB [--match REGEX]``
    per kernel, the mean per dispatch of every counter of a ``--pmc`` run (SQ counters summed
    over their hardware instances, GRBM counters the largest instance: each counts the whole
    GPU's cycles), and from them: waves, VALU, FP64 matrix and LDS instructions per wave, and
    the matrix pipe's busy cycles per matrix instruction.  GRBM_GUI_ACTIVE spans the whole
    dispatch window under the profiler, well beyond the kernel's own run, so busy fractions
    are left to the kernel's measured duration.
"""
from __future__ import annotations

import argparse
import collections
import json
import re
import sqlite3
import sys

COPY_KERNEL = "__amd_rocclr_copyBuffer"


def kernel_durations(db):
    """name -> list of durations (ns), one per dispatch."""
    out = collections.defaultdict(list)
    with sqlite3.connect(db) as c:
        for name, dur in c.execute("select name, duration from kernels"):
            out[name].append(dur)
    return out


def counter_per_dispatch(db, counter):
    """name -> list of the counter's value per dispatch (hardware instances summed)."""
    per = collections.defaultdict(float)
    names = {}
    with sqlite3.connect(db) as c:
        for name, disp, cn, cv in c.execute(
                "select name, dispatch_id, counter_name, counter_value from pmc_events where counter_name = ?",
                (counter,)):
            per[disp] += cv
            names[disp] = name
    out = collections.defaultdict(list)
    for disp, v in per.items():
        out[names[disp]].append(v)
    return out


def counters_per_dispatch(db):
    """name -> {counter -> list of per-dispatch values}: SQ and other block counters summed
    over their hardware instances, GRBM counters the largest instance."""
    per = collections.defaultdict(lambda: collections.defaultdict(float))
    names = {}
    with sqlite3.connect(db) as c:
        for name, disp, cn, cv in c.execute(
                "select name, dispatch_id, counter_name, counter_value from pmc_events"):
            names[disp] = name
            if cn.startswith("GRBM_"):
                per[disp][cn] = max(per[disp][cn], cv)
            else:
                per[disp][cn] += cv
    out = collections.defaultdict(lambda: collections.defaultdict(list))
    for disp, cs in per.items():
        for cn, v in cs.items():
            out[names[disp]][cn].append(v)
    return out


def short(name, n=90):
    return name if len(name) <= n else name[:n - 3] + "..."


def cmd_kernels(a):
    d = kernel_durations(a.db)
    total = sum(sum(v) for v in d.values()) or 1.0
    rows = sorted(d.items(), key=lambda kv: -sum(kv[1]))
    if a.match:
        rows = [r for r in rows if re.search(a.match, r[0])]
    print("| kernel | calls | mean (us) | min (us) | max (us) | total (ms) | share |")
    print("|---|---:|---:|---:|---:|---:|---:|")
    for name, v in rows[:a.top]:
        print("| `%s` | %d | %.1f | %.1f | %.1f | %.2f | %.1f%% |" % (
            short(name), len(v), sum(v) / len(v) / 1e3, min(v) / 1e3, max(v) / 1e3, sum(v) / 1e6,
            100.0 * sum(v) / total))


def _kshort(name):
    m = re.search(r"mdf::(k_\w+)", name)
    return m.group(1) if m else name[:30]


def cmd_timeline(a):
    with sqlite3.connect(a.db) as c:
        rows = c.execute("select name, start, end, queue_id from kernels order by start").fetchall()
    ends = [i for i, r in enumerate(rows) if _kshort(r[0]) == a.last]
    steps = [rows[i + 1:j + 1] for i, j in zip(ends, ends[1:])]
    steps = [st for st in steps if any(_kshort(r[0]) == a.first for r in st)]
    if a.forked:
        steps = [st for st in steps if len({r[3] for r in st}) > 1]
    if not steps:
        sys.exit("no steps between %s dispatches hold %s" % (a.last, a.first))
    span = lambda st: (max(r[2] for r in st) - min(r[1] for r in st)) / 1e3
    build = lambda st: any(_kshort(r[0]) == "k_build" and r[2] - r[1] > 20000 for r in st)
    for title, group in (("steps without a list build", [st for st in steps if not build(st)]),
                         ("steps with a list build", [st for st in steps if build(st)])):
        if not group:
            continue
        group = sorted(group, key=span)
        sp = [span(st) for st in group]
        print("%s: %d, span median %.1f us, min %.1f, max %.1f; the median step:\n" % (
            title, len(group), sp[len(sp) // 2], sp[0], sp[-1]))
        st = group[len(group) // 2]
        t0 = min(r[1] for r in st)
        print("| kernel | queue | start (us) | end (us) | duration (us) |")
        print("|---|---:|---:|---:|---:|")
        for r in st:
            print("| `%s` | %s | %.1f | %.1f | %.1f |" % (_kshort(r[0]), r[3], (r[1] - t0) / 1e3, (r[2] - t0) / 1e3,
                                                      (r[2] - r[1]) / 1e3))
        print()


def cmd_counters(a):
    d = counters_per_dispatch(a.db)
    names = sorted(d, key=lambda n: -sum(d[n].get("GRBM_GUI_ACTIVE", [0.0])))
    if a.match:
        names = [n for n in names if re.search(a.match, n)]
    cols = sorted({cn for n in names for cn in d[n]})
    mean = lambda v: sum(v) / len(v) if v else 0.0
    print("| kernel | dispatches | " + " | ".join(cols) + " |")
    print("|---|---:|" + "---:|" * len(cols))
    for n in names:
        print("| `%s` | %d | " % (short(n, 70), len(next(iter(d[n].values())))) +
              " | ".join("%.4g" % mean(d[n].get(cn, [])) for cn in cols) + " |")
    print()
    print("| kernel | waves | VALU instructions per wave | FP64 matrix instructions per wave | "
          "LDS instructions per wave | matrix busy cycles per matrix instruction |")
    print("|---|---:|---:|---:|---:|---:|")
    for n in names:
        c = {cn: mean(v) for cn, v in d[n].items()}
        waves = c.get("SQ_WAVES", 0.0)
        f = lambda num, den: ("%.1f" % (num / den)) if den > 0 else "-"
        print("| `%s` | %.0f | %s | %s | %s | %s |" % (
            short(n, 70), waves, f(c.get("SQ_INSTS_VALU", 0.0), waves), f(c.get("SQ_INSTS_VALU_MFMA_F64", 0.0), waves),
            f(c.get("SQ_INSTS_LDS", 0.0), waves),
            f(c.get("SQ_VALU_MFMA_BUSY_CYCLES", 0.0), c.get("SQ_INSTS_VALU_MFMA_F64", 0.0))))


def cmd_traffic(a):
    true_kib = a.copy_mib * 1024.0
    f = counter_per_dispatch(a.fetch, "FETCH_SIZE")
    w = counter_per_dispatch(a.write, "WRITE_SIZE")
    if COPY_KERNEL not in f or COPY_KERNEL not in w:
        sys.exit("the calibration copy (%s) is missing from a database" % COPY_KERNEL)
    # the calibration copy is the last (largest) one
    fs = true_kib / max(f[COPY_KERNEL])
    ws = true_kib / max(w[COPY_KERNEL])
    out = {"calibration": {"copy_mib": a.copy_mib, "fetch_kib_read": max(f[COPY_KERNEL]),
                           "write_kib_read": max(w[COPY_KERNEL]), "fetch_scale": fs, "write_scale": ws},
           "kernels": {}}
    print("Calibration: the %d MiB copy reads %.0f KiB by FETCH_SIZE and %.0f KiB by WRITE_SIZE: scales %.3f and "
          "%.3f.\n" % (a.copy_mib, max(f[COPY_KERNEL]), max(w[COPY_KERNEL]), fs, ws))
    print("| kernel | dispatches | fetch (MB) | write (MB) |")
    print("|---|---:|---:|---:|")
    for name in sorted(set(f) | set(w)):
        if name == COPY_KERNEL or (a.match and not re.search(a.match, name)):
            continue
        fv, wv = f.get(name, [0.0]), w.get(name, [0.0])
        fb = sum(fv) / len(fv) * fs * 1024.0
        wb = sum(wv) / len(wv) * ws * 1024.0
        out["kernels"][name] = {"dispatches": len(fv), "fetch_bytes": fb, "write_bytes": wb,
                                "bytes": fb + wb}
        print("| `%s` | %d | %.3f | %.3f |" % (short(name), len(fv), fb / 1e6, wb / 1e6))
    if a.json:
        doc = out
        if a.key:
            try:
                with open(a.json) as fh:
                    doc = json.load(fh)
            except (OSError, ValueError):
                doc = {}
            doc[a.key] = out
        with open(a.json, "w") as fh:
            json.dump(doc, fh, indent=1)


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    k = sub.add_parser("kernels")
    k.add_argument("db")
    k.add_argument("--match", default="")
    k.add_argument("--top", type=int, default=30)
    t = sub.add_parser("traffic")
    t.add_argument("--fetch", required=True)
    t.add_argument("--write", required=True)
    t.add_argument("--copy-mib", type=int, default=256)
    t.add_argument("--match", default="")
    t.add_argument("--json", default="")
    t.add_argument("--key", default="")
    tl = sub.add_parser("timeline")
    tl.add_argument("db")
    tl.add_argument("--last", default="k_final")
    tl.add_argument("--first", default="k_pair")
    tl.add_argument("--forked", action="store_true")
    cc = sub.add_parser("counters")
    cc.add_argument("db")
    cc.add_argument("--match", default="")
    a = ap.parse_args(argv)
    {"kernels": cmd_kernels, "traffic": cmd_traffic, "counters": cmd_counters, "timeline": cmd_timeline}[a.cmd](a)
    return 0


if __name__ == "__main__":
    sys.exit(main())
