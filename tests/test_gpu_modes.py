"""HIP engine against the oracle outside the reference's default operating point: exact
constants and accurate erfc, a given alpha, the fully periodic form, and the surface terms
(constraints, angles) - each also pinned on the CPU by closed forms
(tests/test_oracle_closed_forms.py)."""
import sys

import numpy as np
import pytest

from md_sim_amd import abi, systems
from helpers import rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-10


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _compare(s, oracle_lib, fscale=None, **kw):
    """Engine vs oracle; ``fscale`` gives the force scale for configurations whose forces
    vanish by symmetry (then the error is measured against it)."""
    from md_sim_amd.engine import ForceEngine
    pos = abi.positions(s)
    r = ForceEngine(s, **kw).force_get(*pos)
    o = oracle_lib.Oracle(s, **kw)
    ro = o.compute(*pos)
    for a in range(3):
        if fscale is None:
            assert rel_err(r["f_site"][a], ro["f_site"][a]) < TOL, a
        else:
            assert np.abs(r["f_site"][a] - ro["f_site"][a]).max() < TOL * fscale, a
    if s.nrig:
        assert rel_err(r["torque"], ro["torque"]) < TOL
        assert rel_err(r["f_com"], ro["f_com"]) < TOL
    for k in ("e_conf", "e_recip", "e_self", "e_slab", "e_wall_lr"):
        assert abs(r[k] - ro[k]) <= TOL * max(abs(ro[k]), abs(ro["e_recip"]), 1.0), k
    assert rel_err(r["e_vdw"], ro["e_vdw"]) < TOL
    assert rel_err(r["e_coul_real"], ro["e_coul_real"]) < TOL
    assert r["n_pairs_listed"] == ro["n_pairs_listed"]
    assert r["n_wall_pairs_listed"] == ro["n_wall_pairs_listed"]
    return r, ro


def test_madelung_crystal_periodic(oracle_lib):
    _say("case nacl4 z-periodic exact")
    s = systems.config("nacl4")
    kw = dict(z_mode=abi.Z_PERIODIC, alpha=0.35, const_mode=abi.CONST_EXACT, erfc_mode=abi.ERFC_ACCURATE)
    facq = 8.987551784952e9 * 1.6021892e-19 ** 2 * 6.022045e23 / 1e-10 / 1e3
    r0 = s.box[0] / 8
    # forces vanish on the perfect crystal: compare against the scale of one ion pair's force
    r, _ = _compare(s, oracle_lib, fscale=facq / r0 ** 2, **kw)
    e_mad = -(s.nmove / 2) * 1.747564594633182 * facq / r0
    assert abs((r["e_conf"] + r["e_self"]) / e_mad - 1) < 1e-7
    _say("case nacl4 shaken")
    _compare(systems.rocksalt_bulk(4, shake=0.25), oracle_lib, **kw)


@pytest.mark.parametrize("name,kw", [
    ("slab2048", dict(const_mode=abi.CONST_EXACT, erfc_mode=abi.ERFC_ACCURATE)),
    ("slab2048", dict(alpha=0.32)),
    ("bulk216", dict(z_mode=abi.Z_PERIODIC)),
    ("bulk216", dict(z_mode=abi.Z_PERIODIC, const_mode=abi.CONST_EXACT, erfc_mode=abi.ERFC_ACCURATE,
                     alpha=0.35)),
])
def test_modes_match_oracle(oracle_lib, name, kw):
    _say("case", name, kw)
    _compare(systems.config(name), oracle_lib, **kw)


def test_surface_terms_match_oracle(oracle_lib):
    """Constrained surface pairs and angle terms, on the slab with mobile ions."""
    _say("case pore_small constraints+angles")
    s = systems.config("pore_small")
    nl = s.nmove - 4
    o0 = oracle_lib.Oracle(s, n_liquid_sites=nl)
    o0.compute_system()
    l1, l2, w1, w2 = o0.pairs()
    sp = [(i, j) for i, j in zip(l1.tolist(), l2.tolist()) if i >= nl and j >= nl]
    sw = [(i, j) for i, j in zip(w1.tolist(), w2.tolist()) if i >= nl]
    cons = [sp[0], sw[0]]
    ions = np.arange(s.nriga, s.nmove)
    tri = [(s.nmove + 3, ions[0], ions[1]), (ions[2], ions[3], 10), (0, 3, ions[4]), (s.nmove, ions[5], ions[6])]
    kw = dict(n_liquid_sites=nl, constraints=([c[0] for c in cons], [c[1] for c in cons], [1.7, 2.3]),
              angles=([t[0] for t in tri], [t[1] for t in tri], [t[2] for t in tri], [300.0, 150.0, 80.0, 500.0],
                      [1.9, 1.2, 2.2, 1.6]))
    _compare(s, oracle_lib, **kw)
    from md_sim_amd.engine import ForceEngine
    e = ForceEngine(s, **kw)
    e.force_get(*abi.positions(s))
    o = oracle_lib.Oracle(s, **kw)
    o.compute_system()
    m1, m2, u1, u2 = e.pairs()
    n1, n2, v1, v2 = o.pairs()
    assert np.array_equal(m1, n1) and np.array_equal(m2, n2)
    assert np.array_equal(u1, v1) and np.array_equal(u2, v2)
    assert cons[0] not in set(zip(m1.tolist(), m2.tolist()))
    assert cons[1] not in set(zip(u1.tolist(), u2.tolist()))
