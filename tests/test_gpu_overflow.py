"""A pair-list overflow on the device-resident path is never lost.

The reference stops at the first overflowing build ('nnlist_make(): nnlist too big!',
src/nnlist.f90:136-139,157-160).  With mdf_compute_device steps are enqueued without the host
reading the flags, so an overflow at a rebuild step followed by steps that do not rebuild
(and run on the same truncated rows) must still be reported by mdf_read_outputs, and a later
mdf_compute_device refuses to enqueue more steps on them.  A system is crowded by moving
whole molecules next to one lead, which both triggers a rebuild and overfills its row."""
import sys

import numpy as np
import pytest

from md_sim_amd import abi, systems
from helpers import rel_err

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _row_counts(e, n_lead_sites):
    """Full-list row lengths per lead site: each i<j pair counts for both ends."""
    i, j, wi, _ = e.pairs()
    mm = np.bincount(np.concatenate([i, j]), minlength=n_lead_sites)
    mw = np.bincount(wi, minlength=n_lead_sites)
    return mm, mw


def _crowd(s, pos, lead_mol, n_in, seed=7):
    """Move n_in rigid molecules far from molecule lead_mol to random spots 3-9 A from its
    COM (no two oxygens closer than 2.2 A): that molecule's row gains about n_in partners."""
    x, y, z, cx, cy, cz = [np.array(a, np.float64) for a in pos]
    rng = np.random.default_rng(seed)
    com = np.stack([cx, cy, cz], 1)
    c0 = com[lead_mol]
    d = com - c0
    d[:, 0] -= s.box[0] * np.round(d[:, 0] / s.box[0])
    d[:, 1] -= s.box[1] * np.round(d[:, 1] / s.box[1])
    far = np.argsort(-np.linalg.norm(d, axis=1))[:n_in]
    ox = np.stack([x[0:3 * s.nrig:3], y[0:3 * s.nrig:3], z[0:3 * s.nrig:3]], 1)
    for m in far:
        for _ in range(10000):
            v = rng.normal(size=3)
            v *= rng.uniform(3.0, 9.0) / np.linalg.norm(v)
            t = c0 + v
            dd = ox - (t + (ox[m] - com[m]))
            if np.min(np.einsum("ij,ij->i", dd, dd)[np.arange(s.nrig) != m]) > 2.2 ** 2:
                break
        shift = t - com[m]
        for a in range(3):
            x[3 * m + a] += shift[0]; y[3 * m + a] += shift[1]; z[3 * m + a] += shift[2]
        cx[m] += shift[0]; cy[m] += shift[1]; cz[m] += shift[2]
        com[m] = t
        ox[m] = [x[3 * m], y[3 * m], z[3 * m]]
    return x, y, z, cx, cy, cz


def _device(s, pos):
    import torch
    return [torch.tensor(a[:s.nmove] if k < 3 else a, dtype=torch.float64, device="cuda:0")
            for k, a in enumerate(pos)]


def _ptrs(ts):
    return [t.data_ptr() for t in ts]


def _busiest(s):
    from md_sim_amd.engine import ForceEngine
    e = ForceEngine(s)
    mm, mw = _row_counts(e, s.nmove)
    e.close()
    return int(np.argmax(mm)) // 3, int(mm.max()), int(mw.max())


def test_fixed_capacity_overflow_survives_a_step_without_rebuild():
    """Fixed capacity: an overflowing build, then a step that does not rebuild; the read
    reports 'nnlist too big', and the next step rebuilds and succeeds once the rows fit."""
    import torch
    from md_sim_amd.engine import ForceEngine, MDForceError
    s = systems.config("slab2048")
    pos0 = abi.positions(s)
    mol, mm_max, mw_max = _busiest(s)
    cap = max(mm_max, mw_max) + 6
    _say("case fixed capacity", cap, "lead molecule", mol)
    crowd = _crowd(s, pos0, mol, 40)
    e = ForceEngine(s, max_neighbors=cap)
    dev = _device(s, crowd)
    e.compute_device(*_ptrs(dev))   # rebuild step: the crowded row overflows
    e.compute_device(*_ptrs(dev))   # same positions: no rebuild, same truncated rows
    with pytest.raises(MDForceError) as ei:
        e.read_outputs()
    assert ei.value.code == 3 and "later steps ran on the truncated rows" in str(ei.value)
    # the rows are still truncated: the next step rebuilds them, and they fit at pos0
    r = e.force_get(*pos0)
    ref = ForceEngine(s).force_get(*pos0)
    assert rel_err(r["f_site"], ref["f_site"]) < 1e-12
    torch.cuda.synchronize()


def test_growing_capacity_reports_steps_on_truncated_rows():
    """Default (growing) capacity: the overflow of an earlier step is reported after the
    last step has been redone on grown rows; that step's outputs are valid.  Once the host
    can see the pending overflow, mdf_compute_device refuses to enqueue."""
    import torch
    from md_sim_amd.engine import ForceEngine, MDForceError
    s = systems.config("slab2048")
    pos0 = abi.positions(s)
    mol, mm_max, _ = _busiest(s)
    n_in = 256 - mm_max + 60
    _say("case growing capacity: row", mm_max, "+", n_in)
    crowd = _crowd(s, pos0, mol, n_in)
    ref = ForceEngine(s, positions=crowd).force_get(*crowd)
    e = ForceEngine(s)
    dev = _device(s, crowd)
    e.compute_device(*_ptrs(dev))
    e.compute_device(*_ptrs(dev))
    torch.cuda.synchronize()   # the second step has published the pending overflow
    with pytest.raises(MDForceError) as ei:
        e.compute_device(*_ptrs(dev))
    assert ei.value.code == 3 and "not enqueued" in str(ei.value)
    with pytest.raises(MDForceError) as ei:
        e.read_outputs()
    assert ei.value.code == 3 and "truncated rows" in str(ei.value)
    r = ei.value.result
    assert rel_err(r["f_site"], ref["f_site"]) < 1e-12
    assert abs(r["e_conf"] - ref["e_conf"]) <= 1e-12 * abs(ref["e_conf"])
    # grown rows: the handle goes on normally
    e.compute_device(*_ptrs(dev))
    r2 = e.read_outputs()
    assert rel_err(r2["f_site"], ref["f_site"]) < 1e-12


def test_batch_reports_overflow_of_an_earlier_step():
    """The same for a batch: one system overflows, a second batch step does not rebuild,
    the read reports it and carries valid outputs of the last step for every system."""
    from md_sim_amd.engine import ForceBatch, ForceEngine, MDForceError
    s = systems.config("slab2048")
    p = systems.config("pore_small")
    pos0, posp = abi.positions(s), abi.positions(p)
    mol, mm_max, _ = _busiest(s)
    crowd = _crowd(s, pos0, mol, 256 - mm_max + 60)
    _say("case batch overflow")
    ref = ForceEngine(s, positions=crowd).force_get(*crowd)
    refp = ForceEngine(p).force_get(*posp)
    e0, e1 = ForceEngine(s), ForceEngine(p)
    b = ForceBatch([e0, e1])
    d0, d1 = _device(s, crowd), _device(p, posp)
    b.compute_device([_ptrs(d0), _ptrs(d1)])
    b.compute_device([_ptrs(d0), _ptrs(d1)])
    with pytest.raises(MDForceError) as ei:
        b.read_outputs()
    assert ei.value.code == 3 and "system 0" in str(ei.value)
    r0, r1 = ei.value.result
    assert rel_err(r0["f_site"], ref["f_site"]) < 1e-12
    assert rel_err(r1["f_site"], refp["f_site"]) < 1e-12
    b.compute_device([_ptrs(d0), _ptrs(d1)])
    r0, r1 = b.read_outputs()
    assert rel_err(r0["f_site"], ref["f_site"]) < 1e-12
    b.close()
