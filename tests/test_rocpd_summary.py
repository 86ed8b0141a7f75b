"""CPU: the rocpd summaries behind profiles/ on small synthetic databases with known answers."""
import json
import sqlite3

from md_sim_amd import rocpd_summary


def _pmc_db(path, rows):
    with sqlite3.connect(path) as c:
        c.execute("create table pmc_events (name text, dispatch_id int, counter_name text, counter_value real)")
        c.executemany("insert into pmc_events values (?, ?, ?, ?)", rows)


def test_counters_sum_sq_instances_and_take_the_largest_grbm_instance(tmp_path):
    db = str(tmp_path / "c.db")
    rows = []
    for disp in (1, 2):
        for inst in range(8):   # eight hardware instances per counter
            rows += [("k_a", disp, "GRBM_GUI_ACTIVE", 1000.0 + inst), ("k_a", disp, "SQ_WAVES", 2.0),
                     ("k_a", disp, "SQ_INSTS_VALU", 50.0 * disp), ("k_a", disp, "SQ_INSTS_VALU_MFMA_F64", 4.0),
                     ("k_a", disp, "SQ_VALU_MFMA_BUSY_CYCLES", 256.0)]
    _pmc_db(db, rows)
    d = rocpd_summary.counters_per_dispatch(db)
    assert d["k_a"]["GRBM_GUI_ACTIVE"] == [1007.0, 1007.0]
    assert d["k_a"]["SQ_WAVES"] == [16.0, 16.0]
    assert d["k_a"]["SQ_INSTS_VALU"] == [400.0, 800.0]


def test_counters_table_per_wave_figures(tmp_path, capsys):
    db = str(tmp_path / "c.db")
    rows = [("k_a", 1, "SQ_WAVES", 10.0), ("k_a", 1, "SQ_INSTS_VALU", 300.0), ("k_a", 1, "SQ_INSTS_LDS", 20.0),
            ("k_a", 1, "SQ_INSTS_VALU_MFMA_F64", 40.0), ("k_a", 1, "SQ_VALU_MFMA_BUSY_CYCLES", 2560.0)]
    _pmc_db(db, rows)
    rocpd_summary.main(["counters", db])
    last = capsys.readouterr().out.strip().splitlines()[-1]
    # waves 10; 30 VALU, 4 matrix and 2 LDS instructions per wave; 64 busy cycles per matrix instruction
    assert last == "| `k_a` | 10 | 30.0 | 4.0 | 2.0 | 64.0 |", last


def test_traffic_scales_by_the_calibration_copy(tmp_path):
    fdb, wdb, out = str(tmp_path / "f.db"), str(tmp_path / "w.db"), str(tmp_path / "t.json")
    copy = rocpd_summary.COPY_KERNEL
    # the 256 MiB copy reads half its bytes by FETCH_SIZE and all of them by WRITE_SIZE
    _pmc_db(fdb, [("k_a", 1, "FETCH_SIZE", 100.0), ("k_a", 2, "FETCH_SIZE", 300.0),
                  (copy, 3, "FETCH_SIZE", 128.0 * 1024)])
    _pmc_db(wdb, [("k_a", 1, "WRITE_SIZE", 10.0), ("k_a", 2, "WRITE_SIZE", 30.0),
                  (copy, 3, "WRITE_SIZE", 256.0 * 1024)])
    rocpd_summary.main(["traffic", "--fetch", fdb, "--write", wdb, "--json", out, "--key", "fp64"])
    k = json.load(open(out))["fp64"]["kernels"]["k_a"]
    assert k["dispatches"] == 2
    assert k["fetch_bytes"] == 2 * 200.0 * 1024 and k["write_bytes"] == 20.0 * 1024


def test_timeline_splits_steps_at_the_last_kernel(tmp_path, capsys):
    db = str(tmp_path / "k.db")
    with sqlite3.connect(db) as c:
        c.execute("create table kernels (name text, start int, end int, queue_id int)")
        rows = []
        t = 0
        for step, build_ns in enumerate((1000, 30000, 1000)):
            for name, dur, q in (("mdf::k_gather(x)", 4000, 2), ("mdf::k_build(x)", build_ns, 4),
                                 ("mdf::k_pair(x)", 70000, 4), ("mdf::k_final(x)", 8000, 2)):
                rows.append((name, t, t + dur, q))
                t += dur
        c.executemany("insert into kernels values (?, ?, ?, ?)", rows)
    rocpd_summary.main(["timeline", db])
    out = capsys.readouterr().out
    # three steps; the first ends with its k_final and only the next two start after one
    assert "steps without a list build: 1, span median 83.0 us" in out
    assert "steps with a list build: 1, span median 112.0 us" in out
    assert "| `k_build` | 4 | 4.0 | 34.0 | 30.0 |" in out
