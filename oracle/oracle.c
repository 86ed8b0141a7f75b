/*
 * CPU oracle for the md_sim force path - TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C FP64 restatement of the reference's force evaluation, following its loops
 * and arithmetic order.  Only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline may load it, as the checker; the product path never does.
 *
 *   orc_build_lists   nnlist_init + nnlist_make   src/nnlist.f90:14-164
 *   orc_real          force_real loops A and B    src/forcem.f90:233-552
 *   orc_wall_tables   force_real_init             src/forcem.f90:157-231
 *   orc_kset          force_fouri_init            src/forcem.f90:657-825
 *   orc_recip         force_fouri                 src/forcem.f90:827-937
 *   orc_intra         force_intra                 src/forcem.f90:565-652
 *   orc_sum           force_sum                   src/forcem.f90:969-1081
 *   orc_compute       force_get                   src/forcem.f90:93-129
 *
 * Pinning: golden vectors from the reference itself (tests/golden), brute-force pair
 * sets, closed forms (Madelung constant, two-charge Ewald, LJ/Buckingham), Newton's
 * third law and finite differences - see tests/test_oracle_*.py.
 *
 * Build: gcc -O2 -fPIC -shared -ffp-contract=off (no FMA contraction, so the arithmetic
 * is the reference's non-fast-math FP64 path).
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../include/mdforce.h"

/* ---- constants (src/global.f90:18-24, src/forcem.f90:11-13,58-60) ---------------- */
typedef struct {
    double pi, e, na, eps0, k4pi, a2m;
    double a1, a2, a3, a4, a5, p;
    double c309, c201f_unused;
    double press_na_scale, press_bar;   /* 1e-30, 1e-5 literals of the pressure formula */
} orc_consts;

static void consts_init(orc_consts *c, int mode) {
    if (mode == MDF_CONST_REFERENCE) {
        /* default-real literals widened to REAL*8 */
        c->pi = (double)3.1415926535898f;
        c->e = (double)1.6021892E-19f;
        c->na = (double)6.022045E23f;
        c->eps0 = (double)8.85418782E-12f;
        c->k4pi = (double)8.987551784952e+9f;
        c->a2m = (double)1.0E-10f;
        c->a1 = (double)0.254829592f; c->a2 = (double)-0.284496736f; c->a3 = (double)1.421413741f;
        c->a4 = (double)-1.453152027f; c->a5 = (double)1.061405429f; c->p = (double)0.3275911f;
        c->c309 = (double)3.09f;
        c->press_na_scale = (double)1e-30f;
        c->press_bar = (double)1e-5f;
    } else {
        c->pi = 3.14159265358979323846;
        c->e = 1.6021892E-19;
        c->na = 6.022045E23;
        c->eps0 = 1.0 / (4.0 * 3.14159265358979323846 * 8.987551784952e+9);
        c->k4pi = 8.987551784952e+9;
        c->a2m = 1.0E-10;
        c->a1 = 0.254829592; c->a2 = -0.284496736; c->a3 = 1.421413741;
        c->a4 = -1.453152027; c->a5 = 1.061405429; c->p = 0.3275911;
        c->c309 = 3.09;
        c->press_na_scale = 1e-30;
        c->press_bar = 1e-5;
    }
}

typedef struct orc_system {
    mdf_topology t;          /* copy; arrays below are owned copies */
    mdf_params p;
    orc_consts c;
    int nijt;
    /* owned topology arrays */
    int32_t *itype, *nsi, *ijtype, *vkind, *con_a, *con_b, *ang_i, *ang_j, *ang_k;
    double *q, *eps4, *sig6, *aa, *rhoi, *cc, *con_len, *ang_kt, *ang_th;
    /* state */
    double *x, *y, *z;       /* [n_total] */
    double *cx, *cy, *cz;    /* [n_rig]   */
    double *rsx, *rsy, *rsz; /* [n_move] site offset from lead (src/move.f90:157-161) */
    double *ox, *oy, *oz, *ocx, *ocy, *ocz;  /* list snapshots (src/nnlist.f90:99-119) */
    /* constants */
    double lx, ly, lz, lxi, lyi, lzi, lwh;
    double alfa, spii, facq, rcutsq, rskinsq;
    /* lists */
    int32_t *l1, *l2, *lw1, *lw2;
    int64_t ipack, wpack, cap1, capw;
    int rebuilt;
    /* wall tail tables, per type (1-based index) */
    double flr_l[16], flr_w[16], flr_lja[16], flr_ljc[16], flr_bua[16], flr_buc[16], flr_bur[16];
    double ulr_l[16], ulr_w[16];
    /* reciprocal */
    int klmax, kmmax, knmax, nk;
    int32_t *kv;             /* [3*nk] */
    double *ak, *sw_re, *sw_im, *s_re, *s_im;  /* [nk] */
    double ewqWW, uself, pipix, pipiy, pipiz;
    /* outputs */
    double *fx, *fy, *fz;    /* [n_move] */
    double *elj, *eqq;       /* [nijt] */
    double fws[3][3];        /* [xyz][sgn+1] */
    double uewa, ewqLL, ewqWL, urf, ulr, uavex, vir[9];
} orc_system;

static char g_err[512];
const char *orc_last_error(void) { return g_err; }

#define DUP(dst, src, n, T) do { dst = (T *)malloc(sizeof(T) * ((n) > 0 ? (n) : 1)); \
    if ((n) > 0 && src) memcpy(dst, src, sizeof(T) * (n)); } while (0)

/* Fortran anint: round half away from zero */
static inline double anint(double v) { return round(v); }

/* ---- force_real_init: wall tail tables (src/forcem.f90:157-231) ------------------- */
static void orc_wall_tables(orc_system *s) {
    const mdf_topology *t = &s->t;
    int nit = t->n_types;
    int nliq_types = nit - 2;  /* nit - nst(ntype): the crystal has two site types */
    double lw = t->wall_lw, rcut = s->p.rcut;
    double rholiq = (double)t->n_move / (s->lx * s->ly * lw);
    double rhowall = (double)(t->wall_nss[0] + t->wall_nss[1]) /
                     (2.0 * t->wall_cell[0] * t->wall_cell[1] * t->wall_cell[2]);
    double xdens[16];
    for (int i = 0; i < 16; i++) xdens[i] = 0.0;
    for (int i = 0; i < t->n_move; i++) xdens[s->itype[i]] += 1.0;
    for (int it = 1; it <= nliq_types; it++) xdens[it] = xdens[it] / (double)(float)t->n_move * rholiq;
    float tot = (float)(t->wall_nss[0] + t->wall_nss[1]);
    xdens[nit - 1] = (double)((float)t->wall_nss[0] / tot) * rhowall;
    xdens[nit] = (double)((float)t->wall_nss[1] / tot) * rhowall;
    for (int it = 0; it < 16; it++) {
        s->flr_l[it] = s->flr_w[it] = s->flr_lja[it] = s->flr_ljc[it] = 0.0;
        s->flr_bua[it] = s->flr_buc[it] = 0.0; s->flr_bur[it] = 1.0;
        s->ulr_l[it] = s->ulr_w[it] = 0.0;
    }
    for (int it = 1; it <= nliq_types; it++) {
        for (int jt = 1; jt <= nit; jt++) {
            int ijt = s->ijtype[(it - 1) * nit + (jt - 1)] - 1;
            double sgn = (jt <= nliq_types) ? -1.0 : 1.0;
            double flr = 0.0, ul = 0.0;
            if (s->vkind[ijt] == 1) {
                double e = s->eps4[ijt], g = s->sig6[ijt];
                flr = e * g * (g / (10.0 * pow(rcut, 10)) - 1.0 / (4.0 * pow(rcut, 4)));
                ul = e * g * (g / (90.0 * pow(rcut, 9)) - 1.0 / (12.0 * pow(rcut, 3)));
                s->flr_lja[it] += sgn * xdens[jt] * e * g * g / 10.0;
                s->flr_ljc[it] += sgn * xdens[jt] * e * g / 4.0;
            } else if (s->vkind[ijt] == 2) {
                double ro = 1.0 / s->rhoi[ijt];
                flr = ro * (ro + rcut) * s->aa[ijt] * exp(-rcut / ro) - s->cc[ijt] / (4.0 * pow(rcut, 4));
                ul = ro * ro * (2.0 * ro + rcut) * s->aa[ijt] * exp(-rcut / ro) - s->cc[ijt] / (12.0 * pow(rcut, 3));
                s->flr_bua[it] += sgn * xdens[jt] * s->aa[ijt] * ro;
                s->flr_buc[it] += sgn * xdens[jt] * s->cc[ijt] / 4.0;
                s->flr_bur[it] = ro;
            }
            if (sgn < 0.0) { s->flr_l[it] += flr * xdens[jt]; s->ulr_l[it] += ul * xdens[jt]; }
            else           { s->flr_w[it] += flr * xdens[jt]; s->ulr_w[it] += ul * xdens[jt]; }
        }
    }
    double tp = 2.0 * s->c.pi;
    for (int it = 0; it < 16; it++) {
        s->flr_lja[it] *= tp; s->flr_ljc[it] *= tp; s->flr_bua[it] *= tp; s->flr_buc[it] *= tp;
        s->flr_l[it] *= tp; s->flr_w[it] *= tp; s->ulr_l[it] *= tp; s->ulr_w[it] *= tp;
    }
}

/* ---- nnlist (src/nnlist.f90) ------------------------------------------------------- */
static int is_constrained(const orc_system *s, int i, int j) {
    for (int k = 0; k < s->t.n_constraints; k++)
        if ((s->con_a[k] == i && s->con_b[k] == j) || (s->con_a[k] == j && s->con_b[k] == i)) return 1;
    return 0;
}

static int zwrap_list(const orc_system *s) { return s->p.z_mode == MDF_Z_PERIODIC; }
static int zwrap_force(const orc_system *s) { return s->p.z_mode != MDF_Z_OPEN; }

/* candidate enumeration order i<j row-major (nnlist.f90:37-55,66-79) filtered by
   dx^2+dy^2+dz^2 < rskinsq with x,y minimum image (nnlist.f90:123-156)                */
static int orc_build_lists(orc_system *s) {
    const mdf_topology *t = &s->t;
    int nm = t->n_move, nt = t->n_total;
    for (int i = 0; i < t->n_rig; i++) { s->ocx[i] = s->cx[i]; s->ocy[i] = s->cy[i]; s->ocz[i] = s->cz[i]; }
    for (int i = t->n_rigid_sites; i < nm; i++) { s->ox[i] = s->x[i]; s->oy[i] = s->y[i]; s->oz[i] = s->z[i]; }
    int64_t left = 0;
    for (int i = 0; i < nm - 1; i++) {
        if (s->nsi[s->itype[i] - 1] == 0) continue;
        for (int j = i + 1; j < nm; j++) {
            if (s->nsi[s->itype[j] - 1] == 0) continue;
            if (i >= t->n_liquid_sites && j >= t->n_liquid_sites && is_constrained(s, i, j)) continue;
            double dx = s->x[i] - s->x[j]; dx = dx - s->lx * anint(dx * s->lxi);
            double dy = s->y[i] - s->y[j]; dy = dy - s->ly * anint(dy * s->lyi);
            double dz = s->z[i] - s->z[j];
            if (zwrap_list(s)) dz = dz - s->lz * anint(dz * s->lzi);
            if (dx * dx + dy * dy + dz * dz < s->rskinsq) {
                if (left >= s->cap1) {
                    s->cap1 = s->cap1 * 2 + 1024;
                    s->l1 = (int32_t *)realloc(s->l1, sizeof(int32_t) * s->cap1);
                    s->l2 = (int32_t *)realloc(s->l2, sizeof(int32_t) * s->cap1);
                }
                s->l1[left] = i; s->l2[left] = j; left++;
            }
        }
    }
    s->ipack = left;
    left = 0;
    for (int i = 0; i < nm; i++) {
        if (s->nsi[s->itype[i] - 1] == 0) continue;
        for (int j = nm; j < nt; j++) {
            if (i >= t->n_liquid_sites && is_constrained(s, i, j)) continue;
            double dx = s->x[i] - s->x[j]; dx = dx - s->lx * anint(dx * s->lxi);
            double dy = s->y[i] - s->y[j]; dy = dy - s->ly * anint(dy * s->lyi);
            double dz = s->z[i] - s->z[j];
            if (zwrap_list(s)) dz = dz - s->lz * anint(dz * s->lzi);
            if (dx * dx + dy * dy + dz * dz < s->rskinsq) {
                if (left >= s->capw) {
                    s->capw = s->capw * 2 + 1024;
                    s->lw1 = (int32_t *)realloc(s->lw1, sizeof(int32_t) * s->capw);
                    s->lw2 = (int32_t *)realloc(s->lw2, sizeof(int32_t) * s->capw);
                }
                s->lw1[left] = i; s->lw2[left] = j; left++;
            }
        }
    }
    s->wpack = left;
    s->rebuilt = 1;
    return MDF_OK;
}

/* erfc(alfar)*exp-form pieces (forcem.f90:335-340) */
static inline void coul_terms(const orc_system *s, double alfar, double *erfx, double *exp2a) {
    *exp2a = exp(-alfar * alfar);
    if (s->p.erfc_mode == MDF_ERFC_AS7126) {
        double t = 1.0 / (1.0 + s->c.p * alfar);
        *erfx = ((((s->c.a5 * t + s->c.a4) * t + s->c.a3) * t + s->c.a2) * t + s->c.a1) * t * *exp2a;
    } else {
        *erfx = erfc(alfar);
    }
}

static inline void vir_add(orc_system *s, double dx, double dy, double dz, double f) {
    s->vir[0] += dx * dx * f; s->vir[1] += dx * dy * f; s->vir[2] += dx * dz * f;
    s->vir[3] += dy * dx * f; s->vir[4] += dy * dy * f; s->vir[5] += dy * dz * f;
    s->vir[6] += dz * dx * f; s->vir[7] += dz * dy * f; s->vir[8] += dz * dz * f;
}

/* ---- force_real (src/forcem.f90:233-552) --------------------------------------------- */
static int orc_real(orc_system *s) {
    const mdf_topology *t = &s->t;
    int nit = t->n_types, nm = t->n_move;
    /* skin test (forcem.f90:245-250) */
    double aux = 0.0;
    for (int i = 0; i < t->n_rig; i++) {
        double d = (s->ocx[i] - s->cx[i]) * (s->ocx[i] - s->cx[i]) + (s->ocy[i] - s->cy[i]) * (s->ocy[i] - s->cy[i]) +
                   (s->ocz[i] - s->cz[i]) * (s->ocz[i] - s->cz[i]);
        if (d > aux) aux = d;
    }
    for (int i = t->n_rigid_sites; i < nm; i++) {
        double d = (s->ox[i] - s->x[i]) * (s->ox[i] - s->x[i]) + (s->oy[i] - s->y[i]) * (s->oy[i] - s->y[i]) +
                   (s->oz[i] - s->z[i]) * (s->oz[i] - s->z[i]);
        if (d > aux) aux = d;
    }
    s->rebuilt = 0;
    if (4.0 * aux > s->p.skin * s->p.skin) orc_build_lists(s);

    double pe = 0.0;  /* uavex partial: every pair energy (luin is always true) */
    int zf = zwrap_force(s);
    /* loop A: moving x moving (forcem.f90:288-403) */
    for (int64_t k = 0; k < s->ipack; k++) {
        int i = s->l1[k], j = s->l2[k];
        double xio = s->x[i] - s->x[j]; xio = xio - s->lx * anint(xio * s->lxi);
        double yio = s->y[i] - s->y[j]; yio = yio - s->ly * anint(yio * s->lyi);
        double zio = s->z[i] - s->z[j];
        if (zf) zio = zio - s->lz * anint(zio * s->lzi);
        double rsq = xio * xio + yio * yio + zio * zio;
        if (rsq > s->rcutsq) continue;
        int it = s->itype[i], jt = s->itype[j];
        int ijt = s->ijtype[(it - 1) * nit + (jt - 1)] - 1;
        double rsqi = 1.0 / rsq, rij = sqrt(rsq), fij, r6;
        switch (s->vkind[ijt]) {
            case 0: fij = 0.0; break;
            case 1:
                r6 = s->sig6[ijt] * (rsqi * rsqi * rsqi);
                s->elj[ijt] += s->eps4[ijt] * r6 * (r6 - 1.0);
                pe += s->eps4[ijt] * r6 * (r6 - 1.0);
                fij = 6.0 * rsqi * s->eps4[ijt] * r6 * (2.0 * r6 - 1.0);
                break;
            case 2:
                r6 = s->cc[ijt] * (rsqi * rsqi * rsqi);
                aux = s->aa[ijt] * exp(-rij * s->rhoi[ijt]);
                s->elj[ijt] += aux - r6;
                pe += aux - r6;
                fij = aux * s->rhoi[ijt] / rij - 6.0 * rsqi * r6;
                break;
            default:
                snprintf(g_err, sizeof g_err, "force_real(): unknown interaction type (f-f)! pair %d %d", i, j);
                return MDF_ERR_VDW_TYPE;
        }
        vir_add(s, xio, yio, zio, fij);
        aux = s->facq * s->q[i] * s->q[j] / rij;
        double alfar = s->alfa * rij, erfx, exp2a;
        coul_terms(s, alfar, &erfx, &exp2a);
        s->eqq[ijt] += erfx * aux;
        pe += erfx * aux;
        double fc = (s->spii * alfar * exp2a + erfx) * aux * rsqi;
        vir_add(s, xio, yio, zio, fc);
        fij = fij + fc;
        s->fx[i] += xio * fij; s->fy[i] += yio * fij; s->fz[i] += zio * fij;
        s->fx[j] -= xio * fij; s->fy[j] -= yio * fij; s->fz[j] -= zio * fij;
        int jmin = j + 1;
        for (int ix = i; ix <= i + s->nsi[it - 1] - 1; ix++) {
            double xi = xio + s->rsx[ix], yi = yio + s->rsy[ix], zi = zio + s->rsz[ix];
            double qx = s->facq * s->q[ix];
            for (int jx = jmin; jx <= j + s->nsi[jt - 1] - 1; jx++) {
                double xd = xi - s->rsx[jx], yd = yi - s->rsy[jx], zd = zi - s->rsz[jx];
                double r2 = xd * xd + yd * yd + zd * zd;
                double r2i = 1.0 / r2, r = sqrt(r2);
                double a2 = qx * s->q[jx] / r;
                double ar = s->alfa * r, ex, e2;
                coul_terms(s, ar, &ex, &e2);
                s->eqq[ijt] += ex * a2;
                pe += ex * a2;
                double f = (s->spii * ar * e2 + ex) * a2 * r2i;
                vir_add(s, xd, yd, zd, f);
                s->fx[ix] += xd * f; s->fy[ix] += yd * f; s->fz[ix] += zd * f;
                s->fx[jx] -= xd * f; s->fy[jx] -= yd * f; s->fz[jx] -= zd * f;
            }
            jmin = j;
        }
    }
    /* surface-site force bookkeeping (forcem.f90:408-413) */
    for (int i = t->n_liquid_sites; i < nm; i++) {
        int sg = (signbit(s->z[i]) ? -1 : 1);
        s->fws[2][sg + 1] += s->fz[i];
    }
    /* loop B: moving x frozen (forcem.f90:418-498) */
    for (int64_t k = 0; k < s->wpack; k++) {
        int i = s->lw1[k], j = s->lw2[k];
        double xi = s->x[i] - s->x[j]; xi = xi - s->lx * anint(xi * s->lxi);
        double yi = s->y[i] - s->y[j]; yi = yi - s->ly * anint(yi * s->lyi);
        double zi = s->z[i] - s->z[j];
        if (zf) zi = zi - s->lz * anint(zi * s->lzi);
        double rsq = xi * xi + yi * yi + zi * zi;
        if (rsq > s->rcutsq) continue;
        int sg = (signbit(s->z[j]) ? -1 : 1);
        int it = s->itype[i];
        int ijt = s->ijtype[(it - 1) * nit + (s->itype[j] - 1)] - 1;
        double fij, rsqi, rij, r6;
        switch (s->vkind[ijt]) {
            case 0: fij = 0.0; break;
            case 1:
                rsqi = 1.0 / rsq;
                r6 = s->sig6[ijt] * (rsqi * rsqi * rsqi);
                s->elj[ijt] += s->eps4[ijt] * r6 * (r6 - 1.0);
                pe += s->eps4[ijt] * r6 * (r6 - 1.0);
                fij = 6.0 * rsqi * s->eps4[ijt] * r6 * (2.0 * r6 - 1.0);
                break;
            case 2:
                rsqi = 1.0 / rsq; rij = sqrt(rsq);
                r6 = s->cc[ijt] * (rsqi * rsqi * rsqi);
                aux = s->aa[ijt] * exp(-rij * s->rhoi[ijt]);
                s->elj[ijt] += aux - r6;
                pe += aux - r6;
                fij = aux * s->rhoi[ijt] / rij - 6.0 * rsqi * r6;
                break;
            default:
                snprintf(g_err, sizeof g_err, "force_real(): unknown interaction type (f-s)! pair %d %d", i, j);
                return MDF_ERR_VDW_TYPE;
        }
        double qx = s->facq * s->q[j];
        for (int ix = i; ix <= i + s->nsi[it - 1] - 1; ix++) {
            double xd = xi + s->rsx[ix], yd = yi + s->rsy[ix], zd = zi + s->rsz[ix];
            double r2 = xd * xd + yd * yd + zd * zd;
            double r = sqrt(r2), r2i = 1.0 / r2;
            double a2 = qx * s->q[ix] / r;
            double ar = s->alfa * r, ex, e2;
            coul_terms(s, ar, &ex, &e2);
            s->eqq[ijt] += ex * a2;
            pe += ex * a2;
            double fc = (s->spii * ar * e2 + ex) * a2 * r2i;
            vir_add(s, xd, yd, zd, fij + fc);
            fij = fij + fc;
            s->fx[ix] += xd * fij; s->fy[ix] += yd * fij; s->fz[ix] += zd * fij;
            if (it <= t->n_ltype) {
                s->fws[0][sg + 1] -= xd * fij; s->fws[1][sg + 1] -= yd * fij; s->fws[2][sg + 1] -= zd * fij;
            }
            fij = 0.0;
        }
    }
    /* wall long-range tail correction (forcem.f90:509-543) */
    s->ulr = 0.0;
    if (s->p.wall_lr) {
        double rcut = s->p.rcut;
        for (int i = 0; i < nm; i++) {
            int it = s->itype[i];
            if (fabs(s->q[i]) < 0.7) continue;
            for (int jj = -1; jj <= 1; jj += 2) {
                double a = 0.0, sgn = (double)jj;
                double dw = s->lwh - sgn * s->z[i];
                if (dw < rcut) {
                    a = a + sgn * s->flr_w[it];
                    s->ulr += s->ulr_w[it];
                } else {
                    a = a + sgn * s->flr_l[it];
                    s->ulr += s->ulr_l[it];
                    double dwa = 1.0 / (dw * dw);
                    a = a + sgn * (s->flr_lja[it] * pow(dwa, 5) - s->flr_ljc[it] * (dwa * dwa));
                    s->ulr += (s->flr_lja[it] / (9.0 * pow(dw, 9)) - s->flr_ljc[it] / (3.0 * (dw * dw * dw)));
                    double tt = s->flr_bur[it];
                    double e2 = s->flr_bua[it] * exp(-dw / tt);
                    a = a + sgn * ((tt + dw) * e2 - s->flr_buc[it] * (dwa * dwa));
                    s->ulr += (tt * (2.0 * tt + dw) * e2 - s->flr_buc[it] / (3.0 * (dw * dw * dw)));
                }
                s->fz[i] -= a;
                if (it <= t->n_ltype) s->fws[2][jj + 1] += a;
            }
        }
    }
    s->uavex = pe + s->ulr;
    return MDF_OK;
}

/* ---- force_fouri_init (src/forcem.f90:657-825) -------------------------------------- */
static int orc_kset(orc_system *s) {
    const mdf_topology *t = &s->t;
    double lmin = fmin(s->lx, fmin(s->ly, s->lz));
    int kmax = s->p.kmax;
    s->pipix = 2.0 * s->c.pi / s->lx;
    s->pipiy = 2.0 * s->c.pi / s->ly;
    s->pipiz = 2.0 * s->c.pi / s->lz;
    double alf4i = -1.0 / (4.0 * s->alfa * s->alfa);
    double fac = 2.0 * s->c.e * s->c.e * s->c.na / (s->c.eps0 * s->c.a2m * 1000.0 * s->lx * s->ly * s->lz);
    s->klmax = (int)lround((double)(float)kmax * s->lx / lmin);
    s->kmmax = (int)lround((double)(float)kmax * s->ly / lmin);
    s->knmax = (int)lround((double)(float)kmax * s->lz / lmin);
    int klmax = s->klmax, kmmax = s->kmmax, knmax = s->knmax;
    /* count then fill, in the enumeration order of forcem.f90:743-776 */
    int cap = (klmax + 1) * (2 * kmmax + 1) * (2 * knmax + 1);
    s->kv = (int32_t *)malloc(sizeof(int32_t) * 3 * cap);
    s->ak = (double *)malloc(sizeof(double) * cap);
    int n = 0, kmmin = 0, knmin = 1;
    for (int kl = 0; kl <= klmax; kl++) {
        for (int km = kmmin; km <= kmmax; km++) {
            for (int kn = knmin; kn <= knmax; kn++) {
                float ksq = (float)(kl * kl + km * km) + (float)((int64_t)kn * kn * kmax * kmax) / (float)(knmax * knmax);
                if (ksq < (float)(kmax * kmax) + 2.01f) {
                    double rl = s->pipix * (double)(float)kl, rm = s->pipiy * (double)(float)km, rn = s->pipiz * (double)(float)kn;
                    double rksq = rl * rl + rm * rm + rn * rn;
                    s->kv[3 * n] = kl; s->kv[3 * n + 1] = km; s->kv[3 * n + 2] = kn;
                    s->ak[n] = fac * exp(rksq * alf4i) / rksq;
                    n++;
                }
            }
            knmin = -knmax;
        }
        kmmin = -kmmax;
    }
    s->nk = n;
    s->sw_re = (double *)calloc(n > 0 ? n : 1, sizeof(double));
    s->sw_im = (double *)calloc(n > 0 ? n : 1, sizeof(double));
    s->s_re = (double *)calloc(n > 0 ? n : 1, sizeof(double));
    s->s_im = (double *)calloc(n > 0 ? n : 1, sizeof(double));
    /* static wall structure factor (forcem.f90:709-770), by the same recursion */
    int nw = t->n_total - t->n_move;
    s->ewqWW = 0.0;
    if (nw > 0) {
        int L = klmax + 1, M = 2 * kmmax + 1, N = 2 * knmax + 1;
        double *elr = malloc(sizeof(double) * nw * L), *eli = malloc(sizeof(double) * nw * L);
        double *emr = malloc(sizeof(double) * nw * M), *emi = malloc(sizeof(double) * nw * M);
        double *enr = malloc(sizeof(double) * nw * N), *eni = malloc(sizeof(double) * nw * N);
        for (int w = 0; w < nw; w++) {
            int i = t->n_move + w;
            double *Er = elr + (size_t)w * L, *Ei = eli + (size_t)w * L;
            double *Mr = emr + (size_t)w * M + kmmax, *Mi = emi + (size_t)w * M + kmmax;
            double *Nr = enr + (size_t)w * N + knmax, *Ni = eni + (size_t)w * N + knmax;
            Er[0] = 1.0; Ei[0] = 0.0; Mr[0] = 1.0; Mi[0] = 0.0; Nr[0] = 1.0; Ni[0] = 0.0;
            if (klmax >= 1) { Er[1] = cos(s->pipix * s->x[i]); Ei[1] = sin(s->pipix * s->x[i]); }
            if (kmmax >= 1) { Mr[1] = cos(s->pipiy * s->y[i]); Mi[1] = sin(s->pipiy * s->y[i]); }
            if (knmax >= 1) { Nr[1] = cos(s->pipiz * s->z[i]); Ni[1] = sin(s->pipiz * s->z[i]); }
            for (int k = 2; k <= klmax; k++) { Er[k] = Er[k-1]*Er[1] - Ei[k-1]*Ei[1]; Ei[k] = Er[k-1]*Ei[1] + Ei[k-1]*Er[1]; }
            for (int k = 2; k <= kmmax; k++) { Mr[k] = Mr[k-1]*Mr[1] - Mi[k-1]*Mi[1]; Mi[k] = Mr[k-1]*Mi[1] + Mi[k-1]*Mr[1]; }
            for (int k = 1; k <= kmmax; k++) { Mr[-k] = Mr[k]; Mi[-k] = -Mi[k]; }
            for (int k = 2; k <= knmax; k++) { Nr[k] = Nr[k-1]*Nr[1] - Ni[k-1]*Ni[1]; Ni[k] = Nr[k-1]*Ni[1] + Ni[k-1]*Nr[1]; }
            for (int k = 1; k <= knmax; k++) { Nr[-k] = Nr[k]; Ni[-k] = -Ni[k]; }
        }
        int half = nw / 2;
        for (int kk = 0; kk < n; kk++) {
            int kl = s->kv[3 * kk], km = s->kv[3 * kk + 1], kn = s->kv[3 * kk + 2];
            double r1 = 0, i1 = 0, r2 = 0, i2 = 0;
            for (int w = 0; w < nw; w++) {
                double q = s->q[t->n_move + w];
                double ar = q * elr[(size_t)w * L + kl], ai = q * eli[(size_t)w * L + kl];
                double br = emr[(size_t)w * M + kmmax + km], bi = emi[(size_t)w * M + kmmax + km];
                double mr = ar * br - ai * bi, mi = ar * bi + ai * br;
                double cr = enr[(size_t)w * N + knmax + kn], ci = eni[(size_t)w * N + knmax + kn];
                double er = mr * cr - mi * ci, ei = mr * ci + mi * cr;
                if (w < half) { r1 += er; i1 += ei; } else { r2 += er; i2 += ei; }
            }
            s->sw_re[kk] = r1 + r2; s->sw_im[kk] = i1 + i2;
            s->ewqWW += s->ak[kk] * (s->sw_re[kk] * s->sw_re[kk] + s->sw_im[kk] * s->sw_im[kk]);
        }
        free(elr); free(eli); free(emr); free(emi); free(enr); free(eni);
    }
    s->ewqWW *= 0.5;
    /* self energy (forcem.f90:794-822): atoms, first rigid molecule's pairs, constraints */
    double q2 = 0.0;
    for (int i = 0; i < t->n_total; i++) q2 += s->q[i] * s->q[i];
    s->uself = -0.5 * s->spii * s->facq * s->alfa * q2;
    if (t->n_rig > 0) {
        for (int a = 0; a < 3; a++)
            for (int b = a + 1; b < 3; b++) {
                double rij = sqrt((s->x[a] - s->x[b]) * (s->x[a] - s->x[b]) + (s->y[a] - s->y[b]) * (s->y[a] - s->y[b]) +
                                  (s->z[a] - s->z[b]) * (s->z[a] - s->z[b]));
                double ex, e2;
                coul_terms(s, s->alfa * rij, &ex, &e2);
                s->uself = s->uself - (double)t->n_rig * s->facq * s->q[a] * s->q[b] * (1.0 - ex) / rij;
            }
    }
    for (int k = 0; k < t->n_constraints; k++) {
        double rij = s->con_len[k], ex, e2;
        coul_terms(s, s->alfa * rij, &ex, &e2);
        s->uself = s->uself - s->facq * s->q[s->con_a[k]] * s->q[s->con_b[k]] * (1.0 - ex) / rij;
    }
    return MDF_OK;
}

/* ---- force_fouri (src/forcem.f90:827-937) ------------------------------------------ */
static void orc_recip(orc_system *s) {
    const mdf_topology *t = &s->t;
    int nm = t->n_move;
    int klmax = s->klmax, kmmax = s->kmmax, knmax = s->knmax;
    int L = klmax + 1, M = 2 * kmmax + 1, N = 2 * knmax + 1;
    double *elr = malloc(sizeof(double) * nm * L), *eli = malloc(sizeof(double) * nm * L);
    double *emr = malloc(sizeof(double) * nm * M), *emi = malloc(sizeof(double) * nm * M);
    double *enr = malloc(sizeof(double) * nm * N), *eni = malloc(sizeof(double) * nm * N);
    /* site-major per k index; recursion as forcem.f90:836-862 */
    for (int i = 0; i < nm; i++) {
        double *Er = elr + (size_t)i * L, *Ei = eli + (size_t)i * L;
        double *Mr = emr + (size_t)i * M + kmmax, *Mi = emi + (size_t)i * M + kmmax;
        double *Nr = enr + (size_t)i * N + knmax, *Ni = eni + (size_t)i * N + knmax;
        Er[0] = 1.0; Ei[0] = 0.0; Mr[0] = 1.0; Mi[0] = 0.0; Nr[0] = 1.0; Ni[0] = 0.0;
        if (klmax >= 1) { Er[1] = cos(s->pipix * s->x[i]); Ei[1] = sin(s->pipix * s->x[i]); }
        if (kmmax >= 1) { Mr[1] = cos(s->pipiy * s->y[i]); Mi[1] = sin(s->pipiy * s->y[i]); }
        if (knmax >= 1) { Nr[1] = cos(s->pipiz * s->z[i]); Ni[1] = sin(s->pipiz * s->z[i]); }
        for (int k = 2; k <= klmax; k++) { Er[k] = Er[k-1]*Er[1] - Ei[k-1]*Ei[1]; Ei[k] = Er[k-1]*Ei[1] + Ei[k-1]*Er[1]; }
        for (int k = 2; k <= kmmax; k++) { Mr[k] = Mr[k-1]*Mr[1] - Mi[k-1]*Mi[1]; Mi[k] = Mr[k-1]*Mi[1] + Mi[k-1]*Mr[1]; }
        for (int k = 1; k <= kmmax; k++) { Mr[-k] = Mr[k]; Mi[-k] = -Mi[k]; }
        for (int k = 2; k <= knmax; k++) { Nr[k] = Nr[k-1]*Nr[1] - Ni[k-1]*Ni[1]; Ni[k] = Nr[k-1]*Ni[1] + Ni[k-1]*Nr[1]; }
        for (int k = 1; k <= knmax; k++) { Nr[-k] = Nr[k]; Ni[-k] = -Ni[k]; }
    }
    double *er_ = malloc(sizeof(double) * nm), *ei_ = malloc(sizeof(double) * nm);
    s->ewqLL = 0.0;
    s->uewa = 0.0;
    double lamb = 1.0 / (4.0 * s->alfa * s->alfa);
    for (int kk = 0; kk < s->nk; kk++) {
        int kl = s->kv[3 * kk], km = s->kv[3 * kk + 1], kn = s->kv[3 * kk + 2];
        double rl = s->pipix * (double)(float)kl, rm = s->pipiy * (double)(float)km, rn = s->pipiz * (double)(float)kn;
        double sr = 0.0, si = 0.0;
        for (int i = 0; i < nm; i++) {
            double q = s->q[i];
            double ar = q * elr[(size_t)i * L + kl], ai = q * eli[(size_t)i * L + kl];
            double br = emr[(size_t)i * M + kmmax + km], bi = emi[(size_t)i * M + kmmax + km];
            double mr = ar * br - ai * bi, mi = ar * bi + ai * br;
            double cr = enr[(size_t)i * N + knmax + kn], ci = eni[(size_t)i * N + knmax + kn];
            er_[i] = mr * cr - mi * ci; ei_[i] = mr * ci + mi * cr;
            sr += er_[i]; si += ei_[i];
        }
        double qr = sr + s->sw_re[kk], qi = si + s->sw_im[kk];
        s->s_re[kk] = qr; s->s_im[kk] = qi;
        double akx = s->ak[kk];
        s->ewqLL += akx * (sr * sr + si * si);
        double ek = akx * (qr * qr + qi * qi);
        s->uewa += ek;
        for (int i = 0; i < nm; i++) {
            double ft = akx * (qr * ei_[i] - qi * er_[i]);
            s->fx[i] += ft * rl; s->fy[i] += ft * rm; s->fz[i] += ft * rn;
        }
        /* reciprocal virial of the full charge set, half-space sum */
        double ksq = rl * rl + rm * rm + rn * rn, kv3[3] = {rl, rm, rn};
        double g = 2.0 * (1.0 / ksq + lamb);
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < 3; b++)
                s->vir[a * 3 + b] += 0.5 * ek * ((a == b ? 1.0 : 0.0) - g * kv3[a] * kv3[b]);
    }
    s->ewqLL *= 0.5;
    s->uewa *= 0.5;
    s->ewqWL = s->uewa - s->ewqWW - s->ewqLL;
    s->uavex += s->uewa;
    /* slab correction (forcem.f90:923-928) */
    s->urf = 0.0;
    if (s->p.ew3dc) {
        double dipz = 0.0;
        for (int i = 0; i < nm; i++) dipz += s->q[i] * s->z[i];
        double crf = -2.0 * s->c.pi * s->lzi * s->lxi * s->lyi;
        s->urf = -crf * s->facq * (dipz * dipz);
        crf = 2.0 * crf * s->facq * dipz;
        for (int i = 0; i < nm; i++) s->fz[i] += crf * s->q[i];
    }
    free(elr); free(eli); free(emr); free(emi); free(enr); free(eni); free(er_); free(ei_);
}

/* ---- force_intra (src/forcem.f90:565-652) ------------------------------------------ */
static void orc_intra(orc_system *s) {
    int zf = zwrap_force(s);
    for (int a = 0; a < s->t.n_angles; a++) {
        int i = s->ang_i[a], j = s->ang_j[a], k = s->ang_k[a];
        double ax = s->x[k] - s->x[j]; ax = ax - s->lx * anint(ax * s->lxi);
        double ay = s->y[k] - s->y[j]; ay = ay - s->ly * anint(ay * s->lyi);
        double az = s->z[k] - s->z[j]; if (zf) az = az - s->lz * anint(az * s->lzi);
        double bx = s->x[j] - s->x[i]; bx = bx - s->lx * anint(bx * s->lxi);
        double by = s->y[j] - s->y[i]; by = by - s->ly * anint(by * s->lyi);
        double bz = s->z[j] - s->z[i]; if (zf) bz = bz - s->lz * anint(bz * s->lzi);
        double ai = 1.0 / sqrt(ax * ax + ay * ay + az * az);
        double bi = 1.0 / sqrt(bx * bx + by * by + bz * bz);
        double ab = ax * bx + ay * by + az * bz;
        double cang = -ab * ai * bi;
        double fij = s->ang_kt[a] * (acos(cang) - s->ang_th[a]) / sqrt(1.0 - cang * cang);
        fij = fij * ai * bi;
        double bb = ab * bi * bi, aa = ab * ai * ai, tmp;
        tmp = aa * ax - bx; s->fx[k] += fij * tmp; s->fx[j] += fij * (bb * bx - ax - tmp);
        tmp = aa * ay - by; s->fy[k] += fij * tmp; s->fy[j] += fij * (bb * by - ay - tmp);
        tmp = aa * az - bz; s->fz[k] += fij * tmp; s->fz[j] += fij * (bb * bz - az - tmp);
    }
}

/* ---- force_sum (src/forcem.f90:969-1081), one rank --------------------------------- */
static void orc_sum(orc_system *s, mdf_outputs *o) {
    const mdf_topology *t = &s->t;
    int nm = t->n_move, nijt = s->nijt;
    if (o->f_site) {
        memcpy(o->f_site, s->fx, sizeof(double) * nm);
        memcpy(o->f_site + nm, s->fy, sizeof(double) * nm);
        memcpy(o->f_site + 2 * nm, s->fz, sizeof(double) * nm);
    }
    for (int m = 0; m < t->n_rig; m++) {
        int j = 3 * m;
        double tx = 0, ty = 0, tz = 0, fx = 0, fy = 0, fz = 0;
        for (int a = 0; a < 3; a++) {
            double rx = s->x[j + a] - s->cx[m], ry = s->y[j + a] - s->cy[m], rz = s->z[j + a] - s->cz[m];
            tx += ry * s->fz[j + a] - rz * s->fy[j + a];
            ty += rz * s->fx[j + a] - rx * s->fz[j + a];
            tz += rx * s->fy[j + a] - ry * s->fx[j + a];
            fx += s->fx[j + a]; fy += s->fy[j + a]; fz += s->fz[j + a];
        }
        if (o->torque) { o->torque[m] = tx; o->torque[t->n_rig + m] = ty; o->torque[2 * t->n_rig + m] = tz; }
        if (o->f_com) { o->f_com[m] = fx; o->f_com[t->n_rig + m] = fy; o->f_com[2 * t->n_rig + m] = fz; }
    }
    double sum = 0.0;
    for (int k = 0; k < nijt; k++) sum += s->elj[k];
    for (int k = 0; k < nijt; k++) sum += s->eqq[k];
    sum += s->uewa;
    double uconf = sum + s->uself + s->urf;
    uconf = uconf - s->uself;
    if (o->e_vdw) memcpy(o->e_vdw, s->elj, sizeof(double) * nijt);
    if (o->e_coul_real) memcpy(o->e_coul_real, s->eqq, sizeof(double) * nijt);
    o->e_recip = s->uewa; o->e_recip_ll = s->ewqLL; o->e_recip_wl = s->ewqWL; o->e_recip_ww = s->ewqWW;
    o->e_self = s->uself; o->e_slab = s->urf; o->e_wall_lr = s->ulr; o->e_conf = uconf;
    o->e_avex = s->uavex + s->urf;
    o->press_wall = (s->fws[2][2] - s->fws[2][0]) / (2.0 * s->lx * s->ly) * 1000.0 /
                    (s->c.na * s->c.press_na_scale) * s->c.press_bar;
    for (int a = 0; a < 3; a++) for (int b = 0; b < 3; b++) o->f_wall[a * 3 + b] = s->fws[a][b];
    /* fws(0) = total force (forcem.f90:990-992) */
    double sx = 0, sy = 0, sz = 0;
    for (int i = 0; i < nm; i++) { sx += s->fx[i]; sy += s->fy[i]; sz += s->fz[i]; }
    o->f_wall[1] = sx; o->f_wall[4] = sy; o->f_wall[7] = sz;
    memcpy(o->virial, s->vir, sizeof(double) * 9);
    o->list_rebuilt = s->rebuilt;
    o->n_pairs_listed = (int32_t)s->ipack;
    o->n_wall_pairs_listed = (int32_t)s->wpack;
    o->n_kvec = s->nk;
}

/* ---- public API -------------------------------------------------------------------- */
static void set_positions(orc_system *s, const double *x, const double *y, const double *z,
                          const double *cx, const double *cy, const double *cz, int all) {
    const mdf_topology *t = &s->t;
    int n = all ? t->n_total : t->n_move;
    memcpy(s->x, x, sizeof(double) * n); memcpy(s->y, y, sizeof(double) * n); memcpy(s->z, z, sizeof(double) * n);
    if (t->n_rig > 0) {
        memcpy(s->cx, cx, sizeof(double) * t->n_rig); memcpy(s->cy, cy, sizeof(double) * t->n_rig);
        memcpy(s->cz, cz, sizeof(double) * t->n_rig);
    }
    /* lead offsets (move_predict, src/move.f90:157-161): rigid sites 2,3 relative to site 1;
       leads and single-site molecules 0 */
    memset(s->rsx, 0, sizeof(double) * t->n_move);
    memset(s->rsy, 0, sizeof(double) * t->n_move);
    memset(s->rsz, 0, sizeof(double) * t->n_move);
    for (int m = 0; m < t->n_rig; m++)
        for (int a = 1; a < 3; a++) {
            s->rsx[3 * m + a] = s->x[3 * m + a] - s->x[3 * m];
            s->rsy[3 * m + a] = s->y[3 * m + a] - s->y[3 * m];
            s->rsz[3 * m + a] = s->z[3 * m + a] - s->z[3 * m];
        }
}

int orc_create(const mdf_topology *top, const mdf_params *par, const double *x, const double *y,
               const double *z, const double *cx, const double *cy, const double *cz, orc_system **out) {
    *out = NULL;
    if (top->n_types > 14 || top->n_types < 3) {
        snprintf(g_err, sizeof g_err, "n_types out of range");
        return MDF_ERR_ARG;
    }
    orc_system *s = (orc_system *)calloc(1, sizeof(orc_system));
    s->t = *top;
    s->p = *par;
    consts_init(&s->c, par->const_mode);
    int nit = top->n_types, nt = top->n_total, nm = top->n_move;
    s->nijt = nit * (nit + 1) / 2;
    DUP(s->itype, top->site_type, nt, int32_t);
    DUP(s->q, top->charge, nt, double);
    DUP(s->nsi, top->nsi, nit, int32_t);
    DUP(s->ijtype, top->ijtype, nit * nit, int32_t);
    DUP(s->vkind, top->vdw_kind, s->nijt, int32_t);
    DUP(s->eps4, top->vdw_eps4, s->nijt, double);
    DUP(s->sig6, top->vdw_sig6, s->nijt, double);
    DUP(s->aa, top->vdw_a, s->nijt, double);
    DUP(s->rhoi, top->vdw_rhoi, s->nijt, double);
    DUP(s->cc, top->vdw_c, s->nijt, double);
    DUP(s->con_a, top->con_a, top->n_constraints, int32_t);
    DUP(s->con_b, top->con_b, top->n_constraints, int32_t);
    DUP(s->con_len, top->con_len, top->n_constraints, double);
    DUP(s->ang_i, top->ang_i, top->n_angles, int32_t);
    DUP(s->ang_j, top->ang_j, top->n_angles, int32_t);
    DUP(s->ang_k, top->ang_k, top->n_angles, int32_t);
    DUP(s->ang_kt, top->ang_kt, top->n_angles, double);
    DUP(s->ang_th, top->ang_theta, top->n_angles, double);
    s->x = malloc(sizeof(double) * nt); s->y = malloc(sizeof(double) * nt); s->z = malloc(sizeof(double) * nt);
    int nr = top->n_rig > 0 ? top->n_rig : 1;
    s->cx = calloc(nr, sizeof(double)); s->cy = calloc(nr, sizeof(double)); s->cz = calloc(nr, sizeof(double));
    s->ocx = calloc(nr, sizeof(double)); s->ocy = calloc(nr, sizeof(double)); s->ocz = calloc(nr, sizeof(double));
    s->ox = calloc(nm, sizeof(double)); s->oy = calloc(nm, sizeof(double)); s->oz = calloc(nm, sizeof(double));
    s->rsx = calloc(nm, sizeof(double)); s->rsy = calloc(nm, sizeof(double)); s->rsz = calloc(nm, sizeof(double));
    s->fx = calloc(nm, sizeof(double)); s->fy = calloc(nm, sizeof(double)); s->fz = calloc(nm, sizeof(double));
    s->elj = calloc(s->nijt, sizeof(double)); s->eqq = calloc(s->nijt, sizeof(double));
    set_positions(s, x, y, z, cx, cy, cz, 1);
    s->lx = par->box[0]; s->ly = par->box[1]; s->lz = par->box[2];
    s->lxi = 1.0 / s->lx; s->lyi = 1.0 / s->ly; s->lzi = 1.0 / s->lz;
    s->lwh = 0.5 * top->wall_lw;
    /* force_init constants (forcem.f90:58-60) */
    s->spii = 2.0 / sqrt(s->c.pi);
    s->alfa = par->alpha_mode == MDF_ALPHA_REFERENCE ? s->c.c309 / par->rcut : par->alpha;
    s->facq = s->c.k4pi * s->c.e * s->c.e * s->c.na / (1000.0 * s->c.a2m);
    s->rcutsq = par->rcut * par->rcut;
    /* rskinsq with molecule 1's lead-COM distance (nnlist.f90:23-26) */
    double off = 0.0;
    if (top->n_rig > 0)
        off = sqrt((s->x[0] - s->cx[0]) * (s->x[0] - s->cx[0]) + (s->y[0] - s->cy[0]) * (s->y[0] - s->cy[0]) +
                   (s->z[0] - s->cz[0]) * (s->z[0] - s->cz[0]));
    double rs = par->rcut + par->skin;
    rs = rs + 2.0 * off;
    s->rskinsq = rs * rs;
    orc_kset(s);
    orc_wall_tables(s);
    orc_build_lists(s);
    *out = s;
    return MDF_OK;
}

int orc_compute(orc_system *s, const double *x, const double *y, const double *z, const double *cx,
                const double *cy, const double *cz, mdf_outputs *o) {
    const mdf_topology *t = &s->t;
    int nm = t->n_move;
    set_positions(s, x, y, z, cx, cy, cz, 0);
    memset(s->fx, 0, sizeof(double) * nm); memset(s->fy, 0, sizeof(double) * nm); memset(s->fz, 0, sizeof(double) * nm);
    memset(s->elj, 0, sizeof(double) * s->nijt); memset(s->eqq, 0, sizeof(double) * s->nijt);
    memset(s->fws, 0, sizeof(s->fws));
    memset(s->vir, 0, sizeof(s->vir));
    s->uewa = 0.0;
    int rc = orc_real(s);
    if (rc) return rc;
    orc_recip(s);
    orc_intra(s);
    orc_sum(s, o);
    return MDF_OK;
}

int orc_get_pairs(orc_system *s, int32_t *i_mm, int32_t *j_mm, int32_t cap_mm, int32_t *n_mm, int32_t *i_mw,
                  int32_t *j_mw, int32_t cap_mw, int32_t *n_mw) {
    *n_mm = (int32_t)s->ipack; *n_mw = (int32_t)s->wpack;
    if (s->ipack > cap_mm || s->wpack > cap_mw) return MDF_ERR_ARG;
    memcpy(i_mm, s->l1, sizeof(int32_t) * s->ipack); memcpy(j_mm, s->l2, sizeof(int32_t) * s->ipack);
    memcpy(i_mw, s->lw1, sizeof(int32_t) * s->wpack); memcpy(j_mw, s->lw2, sizeof(int32_t) * s->wpack);
    return MDF_OK;
}

int orc_get_kvectors(orc_system *s, int32_t *kvec, double *ak, int32_t cap, int32_t *n) {
    *n = s->nk;
    if (s->nk > cap) return MDF_ERR_ARG;
    memcpy(kvec, s->kv, sizeof(int32_t) * 3 * s->nk);
    memcpy(ak, s->ak, sizeof(double) * s->nk);
    return MDF_OK;
}

int orc_get_structure_factor(orc_system *s, double *sf, int32_t cap) {
    if (s->nk > cap) return MDF_ERR_ARG;
    for (int k = 0; k < s->nk; k++) { sf[2 * k] = s->s_re[k]; sf[2 * k + 1] = s->s_im[k]; }
    return MDF_OK;
}

/* constants the oracle derived, for the closed-form tests */
void orc_constants(orc_system *s, double *out) {
    out[0] = s->alfa; out[1] = s->facq; out[2] = s->spii; out[3] = s->rskinsq;
    out[4] = s->c.pi; out[5] = s->klmax; out[6] = s->kmmax; out[7] = s->knmax;
}

void orc_destroy(orc_system *s) {
    if (!s) return;
    void *ptrs[] = {s->itype, s->nsi, s->ijtype, s->vkind, s->con_a, s->con_b, s->ang_i, s->ang_j, s->ang_k,
                    s->q, s->eps4, s->sig6, s->aa, s->rhoi, s->cc, s->con_len, s->ang_kt, s->ang_th,
                    s->x, s->y, s->z, s->cx, s->cy, s->cz, s->rsx, s->rsy, s->rsz, s->ox, s->oy, s->oz,
                    s->ocx, s->ocy, s->ocz, s->l1, s->l2, s->lw1, s->lw2, s->kv, s->ak, s->sw_re, s->sw_im,
                    s->s_re, s->s_im, s->fx, s->fy, s->fz, s->elj, s->eqq};
    for (size_t k = 0; k < sizeof(ptrs) / sizeof(ptrs[0]); k++) free(ptrs[k]);
    free(s);
}
