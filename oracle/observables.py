"""CPU statement of the pair-loop observables - TEST INFRASTRUCTURE ONLY.

The checker for the engine's on-device ngc histogram (md_sim_amd/csrc/epilogue.hip,
k_obs_ngc): tests call it, the product path never does.  Plain numpy in float64 with the
reference's operation order (numpy never fuses a multiply into an add), so bins match the
reference's bit for bit.

ngc(l, ijt) counts site pairs at distance r with l = int(r / rdel + 0.5) per type pair ijt
(1-based in the reference, column ijt-1 here), over:
  * moving-moving pairs - the 3-D form only (src/forcem_back.f90:288-298,348-349,353-382):
    every site pair of a listed molecule pair i < j whose lead distance is within rcut
    (min image in x, y, z), at (xio + rs(ix)) - rs(jx), rs = site offset from its lead.
    The slab form has these counts commented out (src/forcem.f90:357,399);
  * moving-frozen pairs - both forms (src/forcem.f90:418-425,459-466,493-495;
    src/forcem_back.f90:405-412,443-449,472-474): every site of a listed moving molecule
    against the listed frozen site, lead distance within rcut (z imaged in the 3-D form only),
    at xi + rs(ix).

Pinned against the ngc arrays the reference's own force loop writes (tests/golden, dumped by
oracle/ref_harness.f90 after one force_real call): pore_small covers the moving-frozen loop of
the slab form, bulk216 the moving-moving loop of the 3-D form.
"""
import numpy as np


def _anint(v):
    # Fortran anint: round half away from zero
    return np.where(v >= 0.0, np.floor(v + 0.5), -np.floor(-v + 0.5))


def _image(d, L):
    return d - L * _anint(d * (1.0 / L))


def ngc(x, y, z, itype, nsi, ijtype, box, rcut, rdel, nbins, nijt, list1, list2,
        listw1, listw2, z_periodic):
    """Site-pair distance histogram of one force call; int64 [nbins, nijt].

    x, y, z, itype: per site (0-based arrays); nsi[t-1], ijtype[t-1, u-1] for 1-based
    types t, u; list1/list2 and listw1/listw2: the reference's 1-based pair lists.
    """
    lx, ly, lz = box
    rcutsq = rcut * rcut
    rdeli = 1.0 / rdel
    out = np.zeros((nbins, nijt), np.int64)

    def count(xd, yd, zd, ta, tb):
        r = np.sqrt(xd * xd + yd * yd + zd * zd)
        l = int(r * rdeli + 0.5)
        if 0 <= l < nbins:
            out[l, ijtype[ta - 1, tb - 1] - 1] += 1

    if z_periodic:
        for i, j in zip(np.asarray(list1) - 1, np.asarray(list2) - 1):
            dx = _image(x[i] - x[j], lx)
            dy = _image(y[i] - y[j], ly)
            dz = _image(z[i] - z[j], lz)
            if dx * dx + dy * dy + dz * dz > rcutsq:
                continue
            for a in range(nsi[itype[i] - 1]):
                xa = dx + (x[i + a] - x[i])
                ya = dy + (y[i + a] - y[i])
                za = dz + (z[i + a] - z[i])
                for b in range(nsi[itype[j] - 1]):
                    count(xa - (x[j + b] - x[j]), ya - (y[j + b] - y[j]), za - (z[j + b] - z[j]),
                          itype[i + a], itype[j + b])
    for i, j in zip(np.asarray(listw1) - 1, np.asarray(listw2) - 1):
        dx = _image(x[i] - x[j], lx)
        dy = _image(y[i] - y[j], ly)
        dz = z[i] - z[j]
        if z_periodic:
            dz = _image(dz, lz)
        if dx * dx + dy * dy + dz * dz > rcutsq:
            continue
        for a in range(nsi[itype[i] - 1]):
            count(dx + (x[i + a] - x[i]), dy + (y[i + a] - y[i]), dz + (z[i + a] - z[i]),
                  itype[i + a], itype[j])
    return out


NGANG_BINS = 202   # the reference's ngang(:, 0:langmax), langmax = 201 (src/measure.f90:96-99)


def ngang(x, y, z, box, ang_i, ang_j, ang_k, ang_type, n_atypes, z_periodic=False):
    """force_intra's angle histogram of one call (src/forcem.f90:570-625); int64
    [NGANG_BINS, n_atypes], the memory of ngang(n_atypes, 0:201).

    Per angle term (0-based sites i, j, k; ang_type = its group's 1-based type iatype):
    a = r_k - r_j, b = r_j - r_i with the x, y minimum image (z too in the 3-D form),
    cos = -(a.b) / |a| / |b| with 1/|a| = 1/sqrt(a.a), bin l = int((cos + 1) 100) + 1.
    """
    lx, ly, lz = box
    lxi, lyi, lzi = 1.0 / lx, 1.0 / ly, 1.0 / lz
    out = np.zeros((NGANG_BINS, n_atypes), np.int64)
    for i, j, k, it in zip(ang_i, ang_j, ang_k, ang_type):
        ax = x[k] - x[j]
        ax = ax - lx * _anint(ax * lxi)
        ay = y[k] - y[j]
        ay = ay - ly * _anint(ay * lyi)
        az = z[k] - z[j]
        bx = x[j] - x[i]
        bx = bx - lx * _anint(bx * lxi)
        by = y[j] - y[i]
        by = by - ly * _anint(by * lyi)
        bz = z[j] - z[i]
        if z_periodic:
            az = az - lz * _anint(az * lzi)
            bz = bz - lz * _anint(bz * lzi)
        ai = 1.0 / np.sqrt(ax * ax + ay * ay + az * az)
        bi = 1.0 / np.sqrt(bx * bx + by * by + bz * bz)
        ab = ax * bx + ay * by + az * bz
        cang = -(ab * ai * bi)
        l = int((cang + 1.0) * 100.0) + 1
        out[l, it - 1] += 1
    return out
