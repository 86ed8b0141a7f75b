"""The ngc oracle against the histograms the reference's own force loop wrote (CPU)."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), ".."))
from md_sim_amd import systems  # noqa: E402
from oracle import observables  # noqa: E402
from helpers import golden  # noqa: E402


def _ngc(name, z_periodic=None, rdel_scale=1.0):
    s = systems.config(name)
    g = golden(name)
    zp = s.z_periodic if z_periodic is None else z_periodic
    return observables.ngc(g["x0"], g["y0"], g["z0"], g["itype"], s.nsi(), s.vdw_tables()[0], s.box,
                           s.rcut, float(g["rdel"]) * rdel_scale, g["ngc"].shape[0], s.nijt,
                           g["list1"], g["list2"], g["listw1"], g["listw2"], zp), g["ngc"]


@pytest.mark.parametrize("name", ["pore_small", "bulk216"])
def test_ngc_matches_reference_dump(name):
    got, ref = _ngc(name)
    assert ref.sum() > 0
    assert np.array_equal(got, ref), np.abs(got - ref).sum()


def test_ngc_form_matters():
    # the moving-moving counts exist only in the 3-D form: the slab form's loops give nothing
    # for bulk water, and the 3-D loops add counts to the slab pore
    got, ref = _ngc("bulk216", z_periodic=False)
    assert got.sum() == 0 and ref.sum() > 0
    got, ref = _ngc("pore_small", z_periodic=True)
    assert got.sum() > ref.sum()


def test_ngc_bin_width_matters():
    # a bin index rounded differently (bins shifted by 5%) breaks the match
    got, ref = _ngc("bulk216", rdel_scale=1.05)
    assert not np.array_equal(got, ref)
    assert got.sum() <= ref.sum()


def _sample_angles():
    g = golden("sample_rutile")
    a = np.asarray(g["angles"], np.float64)
    n = len(a) // 5
    a = a.reshape(5, n)
    geo = g["geometry"]
    ijk = [a[r].astype(np.int64) - 1 for r in range(3)]
    return g, ijk, (float(geo[0]), float(geo[1]), float(geo[2]))


def test_ngang_matches_reference_dump():
    # force_intra's angle histogram on the reference's own sample (54 Ti-O-H angles of the
    # surface OH groups): the restatement equals the array the reference's loop wrote
    g, (ai, aj, ak), box = _sample_angles()
    ref = g["ngang"]
    got = observables.ngang(g["x0"], g["y0"], g["z0"], box, ai, aj, ak, g["angle_type"], ref.shape[1])
    assert ref.sum() == len(ai) == 54
    assert np.array_equal(got, ref)


def test_ngang_vertex_and_sign_matter():
    # the vertex at the end site, or the cosine without its sign, moves the counts
    g, (ai, aj, ak), box = _sample_angles()
    ref = g["ngang"]
    got = observables.ngang(g["x0"], g["y0"], g["z0"], box, aj, ai, ak, g["angle_type"], ref.shape[1])
    assert not np.array_equal(got, ref)
    flip = observables.ngang(-g["x0"], -g["y0"], -g["z0"], box, ai, aj, ak, g["angle_type"], ref.shape[1])
    assert np.array_equal(flip, ref)   # inversion keeps every angle
    mir = observables.ngang(g["x0"], g["y0"], g["z0"], box, ai, ak, aj, g["angle_type"], ref.shape[1])
    assert not np.array_equal(mir, ref)
