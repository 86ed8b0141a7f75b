"""HIP force engine against the oracle and the reference's golden vectors (FP64)."""
import sys

import numpy as np
import pytest

from md_sim_amd import abi, systems
from helpers import golden, golden_positions, rel_err, rms_rel

pytestmark = pytest.mark.gpu

TOL = 1e-10  # relative force error, FP64 engine vs FP64 reference path


def _engine(s, **kw):
    from md_sim_amd.engine import ForceEngine
    return ForceEngine(s, **kw)


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _compare(r, ref, tol=TOL, energies=True):
    for a in range(3):
        assert rel_err(r["f_site"][a], ref["f_site"][a]) < tol, ("f_site", a, rel_err(r["f_site"][a], ref["f_site"][a]))
    assert rel_err(r["f_com"], ref["f_com"]) < tol
    assert rel_err(r["torque"], ref["torque"]) < tol
    if energies:
        assert rel_err(r["e_vdw"], ref["e_vdw"]) < tol
        assert rel_err(r["e_coul_real"], ref["e_coul_real"]) < tol
        for k in ("e_recip", "e_recip_ll", "e_recip_wl", "e_self", "e_slab", "e_wall_lr", "e_conf", "press_wall"):
            # e_recip_wl = uewa - ewqWW - ewqLL is a difference: its error scales with uewa
            base = abs(ref["e_recip"]) if k == "e_recip_wl" else abs(ref[k])
            scale = max(base, 1e-6 * abs(ref["e_conf"]), 1e-12)
            assert abs(r[k] - ref[k]) <= tol * scale, (k, r[k], ref[k])
        # f_wall: forces on the two walls and, in the middle column, the total force on the
        # moving sites (zero up to roundoff): absolute tolerance on the force scale
        fscale = max(np.abs(ref["f_wall"]).max(), np.abs(ref["f_site"]).max())
        assert np.abs(r["f_wall"] - ref["f_wall"]).max() <= tol * fscale


@pytest.mark.parametrize("name", ["bulk216", "pore_small"])
def test_engine_matches_reference_golden(oracle_lib, name):
    _say("case", name)
    g = golden(name)
    s = systems.config(name)
    pos = golden_positions(g)
    eng = _engine(s, positions=pos)
    r = eng.force_get(*pos)
    # pair set bit-exact with the reference's lists
    l1, l2, w1, w2 = eng.pairs()
    assert np.array_equal(l1 + 1, g["list1"]) and np.array_equal(l2 + 1, g["list2"])
    assert np.array_equal(w1 + 1, g["listw1"]) and np.array_equal(w2 + 1, g["listw2"])
    for a, k in enumerate("xyz"):
        assert rel_err(r["f_site"][a], g["f_site_" + k]) < TOL, (k, rel_err(r["f_site"][a], g["f_site_" + k]))
    assert rel_err(r["torque"][0], g["tx"]) < TOL
    assert rel_err(r["e_vdw"], g["elj"]) < TOL
    assert rel_err(r["e_coul_real"], g["eqq"]) < TOL
    for k, gk in [("e_recip", "uewa"), ("e_self", "uself"), ("e_slab", "urf"), ("e_wall_lr", "ulr"),
                  ("e_conf", "uconf"), ("press_wall", "press"), ("e_recip_ll", "ewqLL"), ("e_recip_ww", "ewqWW")]:
        ref = g[gk].item()
        assert abs(r[k] - ref) <= TOL * max(abs(ref), 1e-6), (k, r[k], ref)
    kv, ak, sf = eng.kvectors()
    assert np.array_equal(kv.ravel(), g["k_vec"].astype(int))
    assert rel_err(ak, g["k_ak"]) < 1e-14
    assert rel_err(sf.real, g["k_sre"]) < TOL and rel_err(sf.imag, g["k_sim"]) < TOL
    # and against the oracle on everything the oracle reports
    o = oracle_lib.Oracle(s)
    ro = o.compute(*pos)
    _compare(r, ro)


def test_engine_matches_oracle_slab2048(oracle_lib):
    _say("case slab2048")
    s = systems.config("slab2048")
    eng = _engine(s)
    r = eng.force_get(*__import__("md_sim_amd.abi", fromlist=["positions"]).positions(s))
    o = oracle_lib.Oracle(s)
    ro = o.compute_system()
    l1, l2, w1, w2 = eng.pairs()
    m1, m2, v1, v2 = o.pairs()
    assert np.array_equal(l1, m1) and np.array_equal(l2, m2)
    assert np.array_equal(w1, v1) and np.array_equal(w2, v2)
    _compare(r, ro)
    g = golden("slab2048")
    for a, k in enumerate("xyz"):
        assert rel_err(r["f_site"][a][g["sel"]], g["f_site_" + k]) < TOL


def test_engine_matches_reference_sample():
    """The reference's own sample system (surface OH groups: constraints and angles on
    rutile) through the HIP engine, at the FP64 bar; the pair lists are the reference's."""
    from md_sim_amd import refio
    from md_sim_amd.engine import ForceEngine
    from golden.make_golden import pair_digest
    print("case sample_rutile", flush=True)
    g = golden("sample_rutile")
    # driven from the reference's own input files, read by refio
    txt = lambda k: bytes(g["file_" + k]).decode()
    td = refio.load_reference_inputs(txt("y_fld"), txt("y_top"), txt("y_cfg"), txt("y_inp"), name="sample_rutile")
    e = ForceEngine(td)
    # force_intra's angle histogram ngang (src/forcem.f90:617-625), from the group types
    # refio read from the topology file
    assert np.array_equal(td.angle_type, g["angle_type"])
    e.angle_histogram_enable(td.angle_type, g["ngang"].shape[1])
    r = e.force_get(*td.positions())
    assert np.array_equal(e.angle_histogram(), g["ngang"])
    e.force_get(*td.positions())
    assert np.array_equal(e.angle_histogram(reset=True), 2 * g["ngang"])
    assert e.angle_histogram().sum() == 0
    for a, k in enumerate("xyz"):
        assert rel_err(r["f_site"][a], g["f_site_" + k]) < 1e-10
    assert rel_err(r["torque"][1], g["ty"]) < 1e-10 and rel_err(r["f_com"][0], g["fcx"]) < 1e-10
    for k, gk in [("e_recip", "uewa"), ("e_slab", "urf"), ("e_wall_lr", "ulr"), ("e_conf", "uconf"),
                  ("press_wall", "press")]:
        assert abs(r[k] - g[gk].item()) <= 1e-10 * max(abs(g[gk].item()), 1.0), k
    l1, l2, w1, w2 = e.pairs()
    assert pair_digest(l1 + 1, l2 + 1) == str(g["list_digest"])
    assert pair_digest(w1 + 1, w2 + 1) == str(g["wlist_digest"])


def test_engine_thin_crystal_wall_pairs(oracle_lib):
    """A two-layer crystal thinner than the list radius: every lead-frozen pair within the
    radius is listed (the frozen-site cells are at least that high)."""
    print("case thin crystal", flush=True)
    s = systems.replica_set(1, seed=4200)[0]
    assert s.wall_pos[:, 2].max() - s.wall_pos[:, 2].min() < s.rcut
    pos = abi.positions(s)
    e = _engine(s)
    r = e.force_get(*pos)
    o = oracle_lib.Oracle(s)
    ro = o.compute(*pos)
    assert r["n_wall_pairs_listed"] == ro["n_wall_pairs_listed"]
    l1, l2, w1, w2 = e.pairs()
    m1, m2, u1, u2 = o.pairs()
    assert np.array_equal(w1, u1) and np.array_equal(w2, u2)
    assert rel_err(r["f_site"], ro["f_site"]) < TOL


@pytest.mark.parametrize("name", ["pore_small", "bulk216"])
def test_observables_match_reference_golden(name):
    """ngc and efipos, the reference's hot-loop observables (src/forcem.f90:347-354,388-395,
    486-495,930-934), against its own dump for the golden configuration: the distance
    histogram bit for bit, the ions' z-force profile at 1e-10; two steps accumulate twice."""
    from md_sim_amd.engine import ForceEngine
    _say("case observables", name)
    g = golden(name)
    s = systems.config(name)
    e = ForceEngine(s)
    e.observables_enable(float(g["rdel"]))
    pos = golden_positions(g)
    e.force_get(*pos)
    ngc, efp = e.observables()
    assert ngc.shape == g["ngc"].shape and efp.shape == g["efipos"].shape
    assert np.array_equal(ngc, g["ngc"])
    scale = max(np.abs(g["efipos"]).max(), 1e-30)
    assert np.abs(efp - g["efipos"]).max() <= 1e-10 * scale
    if name == "pore_small":
        assert g["ngc"].sum() > 0 and np.abs(g["efipos"]).max() > 0
    e.force_get(*pos)
    ngc2, efp2 = e.observables(reset=True)
    assert np.array_equal(ngc2, 2 * g["ngc"])
    assert np.abs(efp2 - 2 * g["efipos"]).max() <= 2e-10 * scale
    ngc3, _ = e.observables()
    assert ngc3.sum() == 0
    e.close()
