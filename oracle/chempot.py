"""CPU statement of the reference's Widom test-particle insertion - TEST INFRASTRUCTURE ONLY.

The checker for the engine's on-device insertions (md_sim_amd/csrc/chempot.hip): tests call
it, the product path never does.  It follows `chempot` (src/forcem.f90:1226-1507, the same
routine as src/forcem_back.f90:1192-1473) step by step in float64 with the reference's
operation order, and the generator `rnd` of src/rndmod.f90:28-35.

One call makes up to `ichmax` insertions of an SPC/E water into the configuration of the
step:
  * oxygen at (x, y, z) = ((r-0.5) lx, (r-0.5) ly, z) with z = (r-0.5) 2 lwmax redrawn until
    |z| > lwmin (center distances, src/measure.f90:151-154), bin il = int((|z| - lwmin)
    facbin) + 1, irot = 20 orientations if il > ilmid else 1; inave(il) += irot;
  * a scan of every site except type 2 within rcut of the oxygen (|dz| first, then x, y
    minimum image): an oxygen closer than 2.35 A or another site closer than 2 A is an
    overlap, which ends the whole call (`exit`, src/forcem.f90:1336); otherwise the oxygen's
    vdW and erfc-Coulomb energy, and the partner list;
  * per orientation (h1 uniform in a ball, h2 = h1 x z normalised, m = h2 x h1, angle 2 pi r):
    hydrogens at O + h1 and O + aco h1 + asi (cos a h2 + sin a m); the hydrogen-partner and
    oxygen-partner-hydrogen erfc-Coulomb terms; 0.5 A_k Re(rho_ins(k) conj(S(k))) over the
    k-set; - crf sum q z; the wall tail terms of type 1 for both walls; pave(il) +=
    exp(-beta u), beta = 1000 / (R T).

`crf` is a local of the reference's chempot (src/forcem.f90:1236) that nothing assigns, so
the reference reads an undefined value there; the caller passes the value to use (the tests
pin which value the reference build produced).

Pinned against the reference's own pave/inave/seed dumps (oracle/ref_harness.f90 with
REF_CHEMPOT, tests/golden/chempot.npz): a dense pore where insertions overlap, and a sparse
one with and without overlaps.
"""
import math

import numpy as np

MODULUS = 67099547.0
MULT = 8192.0
SCALE = 1.4903230270690e-8
REF_SEED = 14135621769.0 + 3.0      # rnd_init on rank 0 (src/measure.f90:150)


def f32(v):
    return float(np.float32(v))


def constants(reference=True):
    """chempot's literals: default-real (single precision) in the reference's build."""
    w = f32 if reference else float
    return dict(aco=w(-0.333333), asi=w(0.942809), rinsmin=4.0,
                rwinsmin=float(np.float32(np.float32(2.35) * np.float32(2.35))) if reference else 2.35 ** 2,
                qsw=(w(-0.8476), w(0.4238), w(0.4238)), ph_r=w(8.31441), lo=w(0.001),
                pi=w(3.1415926535898) if reference else math.pi)


class Lcg:
    """seed = mod(8192 seed, 67099547); r = seed * 1.4903230270690e-8 (src/rndmod.f90:28-35).
    The product stays below 2^53, so fmod is exact."""

    def __init__(self, seed=REF_SEED):
        self.seed = float(seed)

    def __call__(self):
        self.seed = math.fmod(MULT * self.seed, MODULUS)
        return self.seed * SCALE


def _anint(v):
    return math.floor(v + 0.5) if v >= 0.0 else -math.floor(-v + 0.5)


def _erfc_term(rij, alfa, c):
    alfar = alfa * rij
    t = 1.0 / (1.0 + c["p"] * alfar)
    exp2a = math.exp(-alfar * alfar)
    return ((((c["a5"] * t + c["a4"]) * t + c["a3"]) * t + c["a2"]) * t + c["a1"]) * t * exp2a


def wall_energy(z, tab, lwh, rcut):
    """Wall tail energy of one site at height z from both walls with its type's tables
    (ulr_w, ulr_l, flr_lja, flr_ljc, flr_bua, flr_buc, flr_bur): the per-site term of the
    force path's ulr (src/forcem.f90:512-531) and chempot's (src/forcem.f90:1481-1492)."""
    u = 0.0
    for jw in (-1.0, 1.0):
        dw = lwh - jw * z
        if dw < rcut:
            u += tab["ulr_w"]
        else:
            u += tab["ulr_l"]
            u += (tab["flr_lja"] / (9.0 * dw ** 9) - tab["flr_ljc"] / (3.0 * dw ** 3))
            t = tab["flr_bur"]
            exp2a = tab["flr_bua"] * math.exp(-dw / t)
            u += (t * (2.0 * t + dw) * exp2a - tab["flr_buc"] / (3.0 * dw ** 3))
    return u


def chempot(st, par, rng, pave, inave, crf=0.0, reference=True, trace=None):
    """One call of chempot; pave (float64) and inave (int64) [lchmax] accumulate in place.

    st: dict of the step's state - x, y, z, q, itype (1-based types), nsi[t-1],
        ijtype[t-1, u-1] (1-based pair numbers), vkind/eps/sig/aa/rhoi/cc indexed by pair
        number (index 0 unused; eps = 4 epsilon, sig = sigma^6 as the reference stores them),
        box, rcut, alfa, facq, erfc constants a1..a5, p, kvec [nk, 3], ak [nk], qsum [nk]
        complex, wall (ulr_w, ulr_l, flr_lja, flr_ljc, flr_bua, flr_buc, flr_bur of type 1),
        lwh.
    par: ichmax, lwmin, lwmax (center distances), facbin, lchmax, ilmid, temper.
    trace: a list that receives, per orientation, dict(ic, sites [3][3], u_ox, u_h, u_k,
    u_rf, u_wall, u) - the oxygen scan's energy, the hydrogen terms, k-space, slab
    correction and wall parts of the insertion energy u.
    Returns the number of insertions made (the overlapping one included).
    """
    c = constants(reference)
    x, y, z, q, itype = st["x"], st["y"], st["z"], st["q"], st["itype"]
    lx, ly, lz = st["box"]
    lxi, lyi = 1.0 / lx, 1.0 / ly
    rcut = st["rcut"]
    rcutsq = rcut * rcut
    alfa, facq = st["alfa"], st["facq"]
    qsw = c["qsw"]
    beta = 1000.0 / (c["ph_r"] * par["temper"])
    wall = st["wall"]
    kvec = np.asarray(st["kvec"], np.int64)
    ak = np.asarray(st["ak"], np.float64)
    qs_conj = np.conj(np.asarray(st["qsum"], np.complex128))
    pipi = (2.0 * c["pi"] / lx, 2.0 * c["pi"] / ly, 2.0 * c["pi"] / lz)
    klmax = int(kvec[:, 0].max()) if len(kvec) else 0
    kmmax = int(np.abs(kvec[:, 1]).max()) if len(kvec) else 0
    knmax = int(np.abs(kvec[:, 2]).max()) if len(kvec) else 0
    lwmin, lwmax = par["lwmin"], par["lwmax"]
    made = 0
    for ic in range(par["ichmax"]):
        oxr = (rng() - 0.5) * lx
        oyr = (rng() - 0.5) * ly
        while True:
            ozr = ((rng() - 0.5) * lwmax) * 2.0
            if abs(ozr) > lwmin:
                break
        il = int((abs(ozr) - lwmin) * par["facbin"]) + 1
        irot = 20 if il > par["ilmid"] else 1
        inave[il - 1] += irot
        made += 1

        # oxygen scan (src/forcem.f90:1282-1335)
        uchem = 0.0
        overlap = False
        nbr = []
        for j in range(len(x)):
            jt = int(itype[j])
            if jt == 2:
                continue
            zio = ozr - z[j]
            if abs(zio) > rcut:
                continue
            xio = oxr - x[j]
            xio = xio - lx * _anint(xio * lxi)
            yio = oyr - y[j]
            yio = yio - ly * _anint(yio * lyi)
            rsq = xio * xio + yio * yio + zio * zio
            if rsq > rcutsq:
                continue
            ijt = int(st["ijtype"][0, jt - 1])
            if rsq < c["rwinsmin"]:
                if jt == 1 or rsq < c["rinsmin"]:
                    overlap = True
                    break
            rsqi = 1.0 / rsq
            rij = math.sqrt(rsq)
            vk = int(st["vkind"][ijt])
            if vk == 1:
                r6 = st["sig"][ijt] * (rsqi * rsqi * rsqi)
                uchem += st["eps"][ijt] * r6 * (r6 - 1.0)
            elif vk == 2:
                r6 = st["cc"][ijt] * (rsqi * rsqi * rsqi)
                aux = st["aa"][ijt] * math.exp(-rij * st["rhoi"][ijt])
                uchem += aux - r6
            aux = facq * qsw[0] * q[j] / rij
            uchem += _erfc_term(rij, alfa, st) * aux
            nbr.append((j, xio, yio, zio))
        if overlap:
            break
        uchemx = uchem

        o = (oxr, oyr, ozr)
        for ir in range(irot):
            while True:
                h1 = [rng() - 0.5, rng() - 0.5, rng() - 0.5]
                ri = h1[0] * h1[0] + h1[1] * h1[1] + h1[2] * h1[2]
                if ri < 0.25 and ri > c["lo"]:
                    break
            ri = 1.0 / math.sqrt(ri)
            h1 = [v * ri for v in h1]
            ri = 1.0 / math.sqrt(h1[0] * h1[0] + h1[1] * h1[1])
            h2 = [-h1[1] * ri, h1[0] * ri, 0.0]
            mr = [h2[1] * h1[2] - h2[2] * h1[1], h2[2] * h1[0] - h2[0] * h1[2], h2[0] * h1[1] - h2[1] * h1[0]]
            alph = 2.0 * c["pi"] * rng()
            cosa, sina = math.cos(alph), math.sin(alph)
            rins = [list(o),
                    [(o[d] + c["aco"] * h1[d]) + c["asi"] * (cosa * h2[d] + sina * mr[d]) for d in range(3)],
                    [o[d] + h1[d] for d in range(3)]]

            # real-space hydrogen terms (src/forcem.f90:1381-1433)
            uchem = 0.0
            for (j, xio, yio, zio) in nbr:
                jt = int(itype[j])
                if jt == 1:
                    for ix in range(3):
                        xi = (xio + rins[ix][0]) - oxr
                        yi = (yio + rins[ix][1]) - oyr
                        zi = (zio + rins[ix][2]) - ozr
                        qx = facq * qsw[ix]
                        for jx in range(j + (1 if ix == 0 else 0), j + int(st["nsi"][jt - 1])):
                            xd = xi - (x[jx] - x[j])
                            yd = yi - (y[jx] - y[j])
                            zd = zi - (z[jx] - z[j])
                            rij = math.sqrt(xd * xd + yd * yd + zd * zd)
                            uchem += _erfc_term(rij, alfa, st) * (qx * q[jx] / rij)
                else:
                    for ix in (1, 2):
                        xi = (xio + rins[ix][0]) - oxr
                        yi = (yio + rins[ix][1]) - oyr
                        zi = (zio + rins[ix][2]) - ozr
                        qx = facq * qsw[ix]
                        rij = math.sqrt(xi * xi + yi * yi + zi * zi)
                        uchem += _erfc_term(rij, alfa, st) * (qx * q[j] / rij)
            u_h = uchem
            uchem = uchem + uchemx

            # k-space (src/forcem.f90:1439-1475): el/em/en by the reference's recursions
            ecomp = []
            for ax, kmx, neg in ((0, klmax, False), (1, kmmax, True), (2, knmax, True)):
                t = np.zeros((3, 2 * kmx + 1), np.complex128)
                for iq in range(3):
                    e1 = complex(math.cos(pipi[ax] * rins[iq][ax]), math.sin(pipi[ax] * rins[iq][ax]))
                    t[iq, kmx] = 1.0
                    if kmx >= 1:
                        t[iq, kmx + 1] = e1
                    for k in range(2, kmx + 1):
                        t[iq, kmx + k] = t[iq, kmx + k - 1] * e1
                    for k in range(1, kmx + 1):
                        t[iq, kmx - k] = np.conj(t[iq, kmx + k])
                ecomp.append((t, kmx))
            (el, klo), (em, kmo), (en, kno) = ecomp
            rho = np.zeros(len(kvec), np.complex128)
            for iq in range(3):
                rho += (qsw[iq] * el[iq, klo + kvec[:, 0]] * em[iq, kmo + kvec[:, 1]]) * en[iq, kno + kvec[:, 2]]
            terms = 0.5 * ak * (rho * qs_conj).real
            u_k = 0.0
            for v in terms:
                u_k += v
            uchem += u_k
            u_rf = -crf * (qsw[0] * rins[0][2] + qsw[1] * rins[1][2] + qsw[2] * rins[2][2])
            uchem = uchem + u_rf

            # wall tail terms of type 1 for both walls (src/forcem.f90:1481-1492)
            u_wall = wall_energy(rins[0][2], wall, st["lwh"], rcut)
            uchem += u_wall
            pave[il - 1] += math.exp(-beta * uchem)
            if trace is not None:
                trace.append(dict(ic=ic, sites=[list(v) for v in rins], u_ox=uchemx, u_h=u_h, u_k=u_k,
                                  u_rf=u_rf, u_wall=u_wall, u=uchem))
    return made


def state_from_dump(d, box, reference=True):
    """The oracle's state from a harness dump (oracle/ref_harness.f90) of the same step."""
    nit = int(d["types"][0])
    par = np.asarray(d["chem_par"], np.float64)
    pad = lambda a: np.r_[0.0, np.asarray(a, np.float64)]
    c = dict(a1=f32(0.254829592), a2=f32(-0.284496736), a3=f32(1.421413741), a4=f32(-1.453152027),
             a5=f32(1.061405429), p=f32(0.3275911)) if reference else \
        dict(a1=0.254829592, a2=-0.284496736, a3=1.421413741, a4=-1.453152027, a5=1.061405429, p=0.3275911)
    if "chem_par" not in d:
        raise KeyError("not a chempot dump")
    st = dict(x=np.asarray(d["x0"]), y=np.asarray(d["y0"]), z=np.asarray(d["z0"]), q=np.asarray(d["qs"]),
              itype=np.asarray(d["itype"]), nsi=np.asarray(d["nsi"]),
              ijtype=np.asarray(d["ijtype"]).reshape(nit, nit),
              vkind=np.r_[0, np.asarray(d["vdwnum"])], eps=pad(d["eps"]), sig=pad(d["sig"]), aa=pad(d["aa"]),
              rhoi=pad(d["rhoi"]), cc=pad(d["cc"]), box=tuple(box), rcut=float(par[12]),
              alfa=float(np.ravel(d["alfa"])[0]), facq=float(np.ravel(d["facq"])[0]),
              kvec=np.asarray(d["k_vec"]).reshape(-1, 3).astype(np.int64), ak=np.asarray(d["k_ak"]),
              qsum=np.asarray(d["k_sre"]) + 1j * np.asarray(d["k_sim"]), lwh=float(par[4]),
              wall=dict(ulr_w=par[5], ulr_l=par[6], flr_lja=par[7], flr_ljc=par[8], flr_bua=par[9],
                        flr_buc=par[10], flr_bur=par[11]))
    st.update(c)
    return st
