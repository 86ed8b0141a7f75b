// Real-space pair kernel: Ewald erfc Coulomb + Lennard-Jones/Buckingham over the pair
// lists, moving-moving (reference loop A, src/forcem.f90:288-403) and moving-frozen
// (loop B, src/forcem.f90:418-498).
//
// Group cut-off on the lead sites, with every site pair of the two molecules evaluated
// when the leads are inside it (src/forcem.f90:299,360-402,460-497); vdW acts on the lead
// pair only; the erfc is the A&S 7.1.26 form with the exact-derivative force
// (src/forcem.f90:333-340).  A group of G lanes owns one lead: lanes stride over its list
// row, accumulate the forces of the lead's own sites (the full list visits every pair
// from both sides, so no atomics and no write conflicts), then reduce with lane shuffles.
// Energies and virial of moving pairs are halved for the double visit; frozen pairs are
// visited once.  Per-lead sums go to column-major partial arrays for a deterministic
// reduction.  Arithmetic type R is double (FP64) or float (FP32 with FP64 accumulation);
// minimum image and the cut-off test are FP64 without contraction in both, so the set of
// interacting pairs is the reference's.
#include "mdf_internal.h"

namespace mdf {

template <typename R>
struct Coul {
    R alfa, spii, p, a1, a2, a3, a4, a5;
    int erfc_mode;
    // energy (erfc(ar) q/r) and force/r  for a site pair at squared distance r2,
    // with qq = facq q_i q_j
    __device__ __forceinline__ void eval(R r2, R qq, R &e, R &f) const {
        R r = sqrt(r2);
        R r2i = R(1) / r2;
        R aux = qq / r;
        R ar = alfa * r;
        R ex = exp(-ar * ar);
        R erfx;
        if (erfc_mode == 0) {
            R t = R(1) / (R(1) + p * ar);
            erfx = ((((a5 * t + a4) * t + a3) * t + a2) * t + a1) * t * ex;
        } else {
            erfx = erfc(ar);
        }
        e = erfx * aux;
        f = (spii * ar * ex + erfx) * aux * r2i;
    }
};

template <typename R>
__device__ __forceinline__ void vdw_eval(int kind, R rsq, R eps4, R sig6, R A, R rhoi, R Cc, R &e, R &f) {
    R rsqi = R(1) / rsq;
    if (kind == 1) {  // Lennard-Jones, eps4 = 4 eps, sig6 = sigma^6 (src/forcem.f90:315-319)
        R r6 = sig6 * (rsqi * rsqi * rsqi);
        e = eps4 * r6 * (r6 - R(1));
        f = R(6) * rsqi * eps4 * r6 * (R(2) * r6 - R(1));
    } else if (kind == 2) {  // Buckingham A exp(-r/rho) - C/r^6 (src/forcem.f90:320-325)
        R rij = sqrt(rsq);
        R r6 = Cc * (rsqi * rsqi * rsqi);
        R aux = A * exp(-rij * rhoi);
        e = aux - r6;
        f = aux * rhoi / rij - R(6) * rsqi * r6;
    } else {
        e = R(0);
        f = R(0);
    }
}

__device__ __forceinline__ double shfl_xor_d(double v, int m) { return __shfl_xor(v, m, 64); }

template <typename R, int G>
__global__ __launch_bounds__(256) void k_pair(DevSys S) {
    const int lanes_per_block = blockDim.x;
    const int gid = blockIdx.x * (lanes_per_block / G) + threadIdx.x / G;
    const int sub = threadIdx.x % G;
    const bool active = gid < S.n_lead;

    double f[3][3];
    double ev[NCLS], ec[NCLS], vir[9], fws[6];
#pragma unroll
    for (int a = 0; a < 3; a++) f[a][0] = f[a][1] = f[a][2] = 0.0;
#pragma unroll
    for (int c = 0; c < NCLS; c++) ev[c] = ec[c] = 0.0;
#pragma unroll
    for (int c = 0; c < 9; c++) vir[c] = 0.0;
#pragma unroll
    for (int c = 0; c < 6; c++) fws[c] = 0.0;

    Coul<R> cl;
    cl.alfa = (R)S.alfa; cl.spii = (R)S.spii; cl.p = (R)S.p;
    cl.a1 = (R)S.a1; cl.a2 = (R)S.a2; cl.a3 = (R)S.a3; cl.a4 = (R)S.a4; cl.a5 = (R)S.a5;
    cl.erfc_mode = S.erfc_mode;

    int i = 0, ti = 0, nsi_i = 0;
    if (active) {
        i = S.lead[gid];
        ti = S.itype[i];
        nsi_i = S.nsi[ti];
        const double xi = S.x[i], yi = S.y[i], zi = S.z[i];
        // site offsets from the lead (the reference's rsx, src/move.f90:157-161) and charges
        double rxi[3] = {0, 0, 0}, ryi[3] = {0, 0, 0}, rzi[3] = {0, 0, 0}, qi[3] = {0, 0, 0};
#pragma unroll
        for (int a = 0; a < 3; a++) {
            if (a < nsi_i) {
                if (a > 0) { rxi[a] = S.x[i + a] - xi; ryi[a] = S.y[i + a] - yi; rzi[a] = S.z[i + a] - zi; }
                qi[a] = S.facq * S.q[i + a];
            }
        }
        const int nt1 = S.n_types + 1;
        // ---- moving partners (loop A) ----
        const int n = min(S.cnt_mm[gid], S.cap_mm);
        const int *row = S.nbr_mm + (size_t)gid * S.cap_mm;
        for (int k = sub; k < n; k += G) {
            const int j = row[k];
            const double xj = S.x[j], yj = S.y[j], zj = S.z[j];
            // lead-lead minimum image as the reference (x,y always; z per variant), no FMA
            double dx = __dsub_rn(xi, xj); dx = __dsub_rn(dx, __dmul_rn(S.lx, anint_d(__dmul_rn(dx, S.lxi))));
            double dy = __dsub_rn(yi, yj); dy = __dsub_rn(dy, __dmul_rn(S.ly, anint_d(__dmul_rn(dy, S.lyi))));
            double dz = __dsub_rn(zi, zj);
            if (S.zper_force) dz = __dsub_rn(dz, __dmul_rn(S.lz, anint_d(__dmul_rn(dz, S.lzi))));
            const double rsq = __dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), __dmul_rn(dz, dz));
            if (rsq > S.rcutsq) continue;
            const int tj = S.itype[j];
            const int nsi_j = S.nsi[tj];
            const int ijt = S.ijtype[ti * nt1 + tj];
            const int cj = S.cls[tj];
            R evd, fvd;
            vdw_eval<R>(S.vkind[ijt], (R)rsq, (R)S.eps4[ijt], (R)S.sig6[ijt], (R)S.aa[ijt], (R)S.rhoi[ijt],
                        (R)S.cc[ijt], evd, fvd);
            double e_v = 0.5 * (double)evd, e_c = 0.0;
            // partner site offsets
            double rxj[3] = {0, 0, 0}, ryj[3] = {0, 0, 0}, rzj[3] = {0, 0, 0}, qj[3] = {0, 0, 0};
#pragma unroll
            for (int b = 0; b < 3; b++) {
                if (b < nsi_j) {
                    if (b > 0) { rxj[b] = S.x[j + b] - xj; ryj[b] = S.y[j + b] - yj; rzj[b] = S.z[j + b] - zj; }
                    qj[b] = S.q[j + b];
                }
            }
            // lead pair: vdW + Coulomb
            {
                R ecl, fcl;
                cl.eval((R)rsq, (R)(qi[0] * qj[0]), ecl, fcl);
                e_c += (double)ecl;
                const double fij = (double)fvd + (double)fcl;
                f[0][0] += dx * fij; f[0][1] += dy * fij; f[0][2] += dz * fij;
                const double hf = 0.5 * fij;
                vir[0] += dx * dx * hf; vir[1] += dx * dy * hf; vir[2] += dx * dz * hf;
                vir[4] += dy * dy * hf; vir[5] += dy * dz * hf; vir[8] += dz * dz * hf;
            }
            // the other site pairs of the two molecules, uncut (src/forcem.f90:360-402)
#pragma unroll
            for (int a = 0; a < 3; a++) {
                if (a >= nsi_i) break;
                const double xa = dx + rxi[a], ya = dy + ryi[a], za = dz + rzi[a];
#pragma unroll
                for (int b = 0; b < 3; b++) {
                    if (b >= nsi_j) break;
                    if (a == 0 && b == 0) continue;
                    const double xd = xa - rxj[b], yd = ya - ryj[b], zd = za - rzj[b];
                    const double r2 = xd * xd + yd * yd + zd * zd;
                    R es, fs;
                    cl.eval((R)r2, (R)(qi[a] * qj[b]), es, fs);
                    e_c += (double)es;
                    const double fd = (double)fs;
                    f[a][0] += xd * fd; f[a][1] += yd * fd; f[a][2] += zd * fd;
                    const double hf = 0.5 * fd;
                    vir[0] += xd * xd * hf; vir[1] += xd * yd * hf; vir[2] += xd * zd * hf;
                    vir[4] += yd * yd * hf; vir[5] += yd * zd * hf; vir[8] += zd * zd * hf;
                }
            }
            e_c *= 0.5;
#pragma unroll
            for (int c = 0; c < NCLS; c++) {
                if (c == cj) { ev[c] += e_v; ec[c] += e_c; }
            }
        }
        // surface-site bookkeeping between the two loops (src/forcem.f90:408-413)
        if (i >= S.n_liquid_sites) {
            int sg = signbit(zi) ? 0 : 1;
            fws[sg * 3 + 2] += f[0][2];
        }
        // ---- frozen partners (loop B) ----
        const int nw = min(S.cnt_mw[gid], S.cap_mw);
        const int *wrow = S.nbr_mw + (size_t)gid * S.cap_mw;
        const bool wall_bk = ti <= S.n_ltype;
        for (int k = sub; k < nw; k += G) {
            const int j = wrow[k];
            const double xj = S.x[j], yj = S.y[j], zj = S.z[j];
            double dx = __dsub_rn(xi, xj); dx = __dsub_rn(dx, __dmul_rn(S.lx, anint_d(__dmul_rn(dx, S.lxi))));
            double dy = __dsub_rn(yi, yj); dy = __dsub_rn(dy, __dmul_rn(S.ly, anint_d(__dmul_rn(dy, S.lyi))));
            double dz = __dsub_rn(zi, zj);
            if (S.zper_force) dz = __dsub_rn(dz, __dmul_rn(S.lz, anint_d(__dmul_rn(dz, S.lzi))));
            const double rsq = __dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), __dmul_rn(dz, dz));
            if (rsq > S.rcutsq) continue;
            const int sg = signbit(zj) ? 0 : 1;
            const int tj = S.itype[j];
            const int ijt = S.ijtype[ti * nt1 + tj];
            const int cj = S.cls[tj];
            R evd, fvd;
            vdw_eval<R>(S.vkind[ijt], (R)rsq, (R)S.eps4[ijt], (R)S.sig6[ijt], (R)S.aa[ijt], (R)S.rhoi[ijt],
                        (R)S.cc[ijt], evd, fvd);
            double e_v = (double)evd, e_c = 0.0;
            const double qw = S.facq * S.q[j];
            double fij = (double)fvd;
#pragma unroll
            for (int a = 0; a < 3; a++) {
                if (a >= nsi_i) break;
                const double xd = dx + rxi[a], yd = dy + ryi[a], zd = dz + rzi[a];
                const double r2 = (a == 0) ? rsq : xd * xd + yd * yd + zd * zd;
                R es, fs;
                cl.eval((R)r2, (R)(qw * S.q[i + a]), es, fs);
                e_c += (double)es;
                fij += (double)fs;
                f[a][0] += xd * fij; f[a][1] += yd * fij; f[a][2] += zd * fij;
                if (wall_bk) { fws[sg * 3 + 0] -= xd * fij; fws[sg * 3 + 1] -= yd * fij; fws[sg * 3 + 2] -= zd * fij; }
                vir[0] += xd * xd * fij; vir[1] += xd * yd * fij; vir[2] += xd * zd * fij;
                vir[4] += yd * yd * fij; vir[5] += yd * zd * fij; vir[8] += zd * zd * fij;
                fij = 0.0;
            }
#pragma unroll
            for (int c = 0; c < NCLS; c++) {
                if (c == cj) { ev[c] += e_v; ec[c] += e_c; }
            }
        }
    }
    // ---- reduce over the G lanes of the group ----
#pragma unroll
    for (int m = G / 2; m > 0; m >>= 1) {
#pragma unroll
        for (int a = 0; a < 3; a++) {
            f[a][0] += shfl_xor_d(f[a][0], m); f[a][1] += shfl_xor_d(f[a][1], m); f[a][2] += shfl_xor_d(f[a][2], m);
        }
#pragma unroll
        for (int c = 0; c < NCLS; c++) { ev[c] += shfl_xor_d(ev[c], m); ec[c] += shfl_xor_d(ec[c], m); }
        vir[0] += shfl_xor_d(vir[0], m); vir[1] += shfl_xor_d(vir[1], m); vir[2] += shfl_xor_d(vir[2], m);
        vir[4] += shfl_xor_d(vir[4], m); vir[5] += shfl_xor_d(vir[5], m); vir[8] += shfl_xor_d(vir[8], m);
#pragma unroll
        for (int c = 0; c < 6; c++) fws[c] += shfl_xor_d(fws[c], m);
    }
    if (active && sub == 0) {
        for (int a = 0; a < nsi_i && a < 3; a++) {
            S.fx[i + a] = f[a][0]; S.fy[i + a] = f[a][1]; S.fz[i + a] = f[a][2];
        }
        const size_t L = (size_t)S.n_lead;
        double *P = S.part;
#pragma unroll
        for (int c = 0; c < NCLS; c++) { P[(COL_EV + c) * L + gid] = ev[c]; P[(COL_EC + c) * L + gid] = ec[c]; }
        vir[3] = vir[1]; vir[6] = vir[2]; vir[7] = vir[5];
#pragma unroll
        for (int c = 0; c < 9; c++) P[(COL_VIR + c) * L + gid] = vir[c];
#pragma unroll
        for (int c = 0; c < 6; c++) P[(COL_FWS + c) * L + gid] = fws[c];
    }
}

template __global__ void k_pair<double, 16>(DevSys);
template __global__ void k_pair<float, 16>(DevSys);
template __global__ void k_pair<double, 8>(DevSys);
template __global__ void k_pair<float, 8>(DevSys);

// forces of moving sites that are neither leads nor lead-carried are zero before the
// reciprocal part: clear every moving-site force the pair kernel does not write
__global__ void k_zero_forces(DevSys S) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < S.n_move) { S.fx[t] = 0.0; S.fy[t] = 0.0; S.fz[t] = 0.0; }
}

}  // namespace mdf
