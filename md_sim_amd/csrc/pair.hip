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

// 1/sqrt(x) and 1/x from the hardware estimates (v_rsq_f64, v_rcp_f64) refined by
// Newton-Raphson iterations (each doubles the correct bits: ~2^-22 -> 2^-44 -> 2^-88), which
// is within an ulp of the IEEE result at a fraction of the cost of the IEEE sequences.
__device__ __forceinline__ double rsqrt_nr(double x) {
    double y = __builtin_amdgcn_rsq(x);
    double h = 0.5 * x;
#pragma unroll
    for (int it = 0; it < 3; it++) {
        double r = __builtin_fma(-h * y, y, 0.5);   // 0.5 - x y^2 / 2
        y = __builtin_fma(y, r, y);                 // y (1.5 - x y^2 / 2)
    }
    return y;
}
__device__ __forceinline__ double rcp_nr(double x) {
    double y = __builtin_amdgcn_rcp(x);
#pragma unroll
    for (int it = 0; it < 2; it++) {
        double e = __builtin_fma(-x, y, 1.0);
        y = __builtin_fma(y, e, y);
    }
    return y;
}
__device__ __forceinline__ float rsqrt_nr(float x) {
    float y = __builtin_amdgcn_rsqf(x);
    float r = __builtin_fmaf(-0.5f * x * y, y, 0.5f);
    return __builtin_fmaf(y, r, y);
}
__device__ __forceinline__ float rcp_nr(float x) {
    float y = __builtin_amdgcn_rcpf(x);
    float e = __builtin_fmaf(-x, y, 1.0f);
    return __builtin_fmaf(y, e, y);
}

template <typename R>
struct Coul {
    R alfa, spii, p, a1, a2, a3, a4, a5;
    int erfc_mode;
    // energy (erfc(ar) q/r) and force/r  for a site pair at squared distance r2,
    // with qq = facq q_i q_j
    __device__ __forceinline__ void eval(R r2, R qq, R &e, R &f) const {
        const R ri = rsqrt_nr(r2);      // 1/r
        const R r = r2 * ri;
        const R r2i = ri * ri;
        const R aux = qq * ri;
        const R ar = alfa * r;
        const R ex = exp(-ar * ar);
        R erfx;
        if (erfc_mode == 0) {
            const R t = rcp_nr(R(1) + p * ar);
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
    R rsqi = rcp_nr(rsq);
    if (kind == 1) {  // Lennard-Jones, eps4 = 4 eps, sig6 = sigma^6 (src/forcem.f90:315-319)
        R r6 = sig6 * (rsqi * rsqi * rsqi);
        e = eps4 * r6 * (r6 - R(1));
        f = R(6) * rsqi * eps4 * r6 * (R(2) * r6 - R(1));
    } else if (kind == 2) {  // Buckingham A exp(-r/rho) - C/r^6 (src/forcem.f90:320-325)
        R rinv = rsqrt_nr(rsq);
        R rij = rsq * rinv;
        R r6 = Cc * (rsqi * rsqi * rsqi);
        R aux = A * exp(-rij * rhoi);
        e = aux - r6;
        f = aux * rhoi * rinv - R(6) * rsqi * r6;
    } else {
        e = R(0);
        f = R(0);
    }
}

__device__ __forceinline__ double shfl_xor_d(double v, int m) { return __shfl_xor(v, m, 64); }

template <typename R, int G, bool VIR, bool ROLLED>
__device__ __forceinline__ void pair_body(const DevSys &S, int bid, int nblk) {
    // per-lane energy-by-class and wall-force accumulators live in LDS slots owned by the
    // lane (no sharing, no atomics); only the site forces and the virial stay in registers
    __shared__ double s_e[2 * NCLS][256];
    __shared__ double s_w[6][256];
    const int tid = threadIdx.x;
    const int gid = bid * (blockDim.x / G) + tid / G;
    const int sub = tid % G;
    const bool active = gid < S.n_lead;
#pragma unroll
    for (int c = 0; c < 2 * NCLS; c++) s_e[c][tid] = 0.0;
#pragma unroll
    for (int c = 0; c < 6; c++) s_w[c][tid] = 0.0;

    double f[3][3];
    double vir[6];
#pragma unroll
    for (int a = 0; a < 3; a++) f[a][0] = f[a][1] = f[a][2] = 0.0;
#pragma unroll
    for (int c = 0; c < 6; c++) vir[c] = 0.0;

    Coul<R> cl;
    cl.alfa = (R)S.alfa; cl.spii = (R)S.spii; cl.p = (R)S.p;
    cl.a1 = (R)S.a1; cl.a2 = (R)S.a2; cl.a3 = (R)S.a3; cl.a4 = (R)S.a4; cl.a5 = (R)S.a5;
    cl.erfc_mode = S.erfc_mode;

    int i = 0, nsi_i = 0;
    if (active) {
        i = S.lead[gid];
        const int ti = S.itype[i];
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
#pragma unroll 1
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
            R evd, fvd;
            vdw_eval<R>(S.vkind[ijt], (R)rsq, (R)S.eps4[ijt], (R)S.sig6[ijt], (R)S.aa[ijt], (R)S.rhoi[ijt],
                        (R)S.cc[ijt], evd, fvd);
            double e_c = 0.0;
            // lead pair: vdW + Coulomb
            {
                R ecl, fcl;
                cl.eval((R)rsq, (R)(qi[0] * S.q[j]), ecl, fcl);
                e_c += (double)ecl;
                const double fij = (double)fvd + (double)fcl;
                f[0][0] += dx * fij; f[0][1] += dy * fij; f[0][2] += dz * fij;
                if (VIR) {
                    const double hf = 0.5 * fij;
                    vir[0] += dx * dx * hf; vir[1] += dx * dy * hf; vir[2] += dx * dz * hf;
                    vir[3] += dy * dy * hf; vir[4] += dy * dz * hf; vir[5] += dz * dz * hf;
                }
            }
            // the other site pairs of the two molecules, uncut (src/forcem.f90:360-402)
            if constexpr (ROLLED) {
              if (nsi_i > 1 || nsi_j > 1) {
                // one site pair per iteration, accumulator chosen by selects: few live registers
                const int npair = nsi_i * nsi_j;
#pragma unroll 1
                for (int pq = 1; pq < npair; pq++) {
                    const int b = pq / nsi_i, a = pq - b * nsi_i;
                    const double xb = (b > 0) ? S.x[j + b] - xj : 0.0;
                    const double yb = (b > 0) ? S.y[j + b] - yj : 0.0;
                    const double zb = (b > 0) ? S.z[j + b] - zj : 0.0;
                    const double qb = S.q[j + b];
                    const double ra_x = a == 0 ? 0.0 : (a == 1 ? rxi[1] : rxi[2]);
                    const double ra_y = a == 0 ? 0.0 : (a == 1 ? ryi[1] : ryi[2]);
                    const double ra_z = a == 0 ? 0.0 : (a == 1 ? rzi[1] : rzi[2]);
                    const double qa = a == 0 ? qi[0] : (a == 1 ? qi[1] : qi[2]);
                    const double xd = (dx + ra_x) - xb, yd = (dy + ra_y) - yb, zd = (dz + ra_z) - zb;
                    const double r2 = xd * xd + yd * yd + zd * zd;
                    R es, fs;
                    cl.eval((R)r2, (R)(qa * qb), es, fs);
                    e_c += (double)es;
                    const double fd = (double)fs;
                    const double gx = xd * fd, gy = yd * fd, gz = zd * fd;
                    f[0][0] += a == 0 ? gx : 0.0; f[0][1] += a == 0 ? gy : 0.0; f[0][2] += a == 0 ? gz : 0.0;
                    f[1][0] += a == 1 ? gx : 0.0; f[1][1] += a == 1 ? gy : 0.0; f[1][2] += a == 1 ? gz : 0.0;
                    f[2][0] += a == 2 ? gx : 0.0; f[2][1] += a == 2 ? gy : 0.0; f[2][2] += a == 2 ? gz : 0.0;
                    if (VIR) {
                        const double hf = 0.5 * fd;
                        vir[0] += xd * xd * hf; vir[1] += xd * yd * hf; vir[2] += xd * zd * hf;
                        vir[3] += yd * yd * hf; vir[4] += yd * zd * hf; vir[5] += zd * zd * hf;
                    }
                }
              }
            } else if (nsi_i > 1 || nsi_j > 1) {
#pragma unroll
                for (int b = 0; b < 3; b++) {
                    if (b >= nsi_j) break;
                    const double xb = (b > 0) ? S.x[j + b] - xj : 0.0;
                    const double yb = (b > 0) ? S.y[j + b] - yj : 0.0;
                    const double zb = (b > 0) ? S.z[j + b] - zj : 0.0;
                    const double qb = S.q[j + b];
#pragma unroll
                    for (int a = 0; a < 3; a++) {
                        if (a >= nsi_i) break;
                        if (a == 0 && b == 0) continue;
                        const double xd = (dx + rxi[a]) - xb, yd = (dy + ryi[a]) - yb, zd = (dz + rzi[a]) - zb;
                        const double r2 = xd * xd + yd * yd + zd * zd;
                        R es, fs;
                        cl.eval((R)r2, (R)(qi[a] * qb), es, fs);
                        e_c += (double)es;
                        const double fd = (double)fs;
                        f[a][0] += xd * fd; f[a][1] += yd * fd; f[a][2] += zd * fd;
                        if (VIR) {
                            const double hf = 0.5 * fd;
                            vir[0] += xd * xd * hf; vir[1] += xd * yd * hf; vir[2] += xd * zd * hf;
                            vir[3] += yd * yd * hf; vir[4] += yd * zd * hf; vir[5] += zd * zd * hf;
                        }
                    }
                }
            }
            const int cj = S.cls[tj];
            s_e[cj][tid] += 0.5 * (double)evd;
            s_e[NCLS + cj][tid] += 0.5 * e_c;
        }
        // surface-site bookkeeping between the two loops (src/forcem.f90:408-413)
        if (i >= S.n_liquid_sites) s_w[(signbit(zi) ? 0 : 3) + 2][tid] += f[0][2];
        // ---- frozen partners (loop B) ----
        const int nw = min(S.cnt_mw[gid], S.cap_mw);
        const int *wrow = S.nbr_mw + (size_t)gid * S.cap_mw;
        const bool wall_bk = ti <= S.n_ltype;
#pragma unroll 1
        for (int k = sub; k < nw; k += G) {
            const int j = wrow[k];
            const double xj = S.x[j], yj = S.y[j], zj = S.z[j];
            double dx = __dsub_rn(xi, xj); dx = __dsub_rn(dx, __dmul_rn(S.lx, anint_d(__dmul_rn(dx, S.lxi))));
            double dy = __dsub_rn(yi, yj); dy = __dsub_rn(dy, __dmul_rn(S.ly, anint_d(__dmul_rn(dy, S.lyi))));
            double dz = __dsub_rn(zi, zj);
            if (S.zper_force) dz = __dsub_rn(dz, __dmul_rn(S.lz, anint_d(__dmul_rn(dz, S.lzi))));
            const double rsq = __dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), __dmul_rn(dz, dz));
            if (rsq > S.rcutsq) continue;
            const int sgo = signbit(zj) ? 0 : 3;
            const int tj = S.itype[j];
            const int ijt = S.ijtype[ti * nt1 + tj];
            R evd, fvd;
            vdw_eval<R>(S.vkind[ijt], (R)rsq, (R)S.eps4[ijt], (R)S.sig6[ijt], (R)S.aa[ijt], (R)S.rhoi[ijt],
                        (R)S.cc[ijt], evd, fvd);
            double e_c = 0.0;
            const double qw = S.facq * S.q[j];
            double fij = (double)fvd;
            double wx = 0.0, wy = 0.0, wz = 0.0;
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
                wx -= xd * fij; wy -= yd * fij; wz -= zd * fij;
                if (VIR) {
                    vir[0] += xd * xd * fij; vir[1] += xd * yd * fij; vir[2] += xd * zd * fij;
                    vir[3] += yd * yd * fij; vir[4] += yd * zd * fij; vir[5] += zd * zd * fij;
                }
                fij = 0.0;
            }
            if (wall_bk) { s_w[sgo + 0][tid] += wx; s_w[sgo + 1][tid] += wy; s_w[sgo + 2][tid] += wz; }
            const int cj = S.cls[tj];
            s_e[cj][tid] += (double)evd;
            s_e[NCLS + cj][tid] += e_c;
        }
    }
    // ---- reduce over the G lanes of the group ----
#pragma unroll
    for (int m = G / 2; m > 0; m >>= 1) {
#pragma unroll
        for (int a = 0; a < 3; a++) {
            f[a][0] += shfl_xor_d(f[a][0], m); f[a][1] += shfl_xor_d(f[a][1], m); f[a][2] += shfl_xor_d(f[a][2], m);
        }
        if (VIR) {
#pragma unroll
            for (int c = 0; c < 6; c++) vir[c] += shfl_xor_d(vir[c], m);
        }
    }
    __syncthreads();
    if (active && sub == 0) {
        for (int a = 0; a < nsi_i && a < 3; a++) {
            S.fx[i + a] = f[a][0]; S.fy[i + a] = f[a][1]; S.fz[i + a] = f[a][2];
        }
        const size_t L = (size_t)S.n_lead;
        double *P = S.part;
        for (int c = 0; c < 2 * NCLS; c++) {
            double acc = 0.0;
            for (int l = 0; l < G; l++) acc += s_e[c][tid + l];
            P[(COL_EV + c) * L + gid] = acc;   // COL_EC = COL_EV + NCLS
        }
        for (int c = 0; c < 6; c++) {
            double acc = 0.0;
            for (int l = 0; l < G; l++) acc += s_w[c][tid + l];
            P[(COL_FWS + c) * L + gid] = acc;
        }
        const double v9[9] = {vir[0], vir[1], vir[2], vir[1], vir[3], vir[4], vir[2], vir[4], vir[5]};
        for (int c = 0; c < 9; c++) P[(COL_VIR + c) * L + gid] = VIR ? v9[c] : 0.0;
    }
}

template <typename R, int G, bool VIR, bool ROLLED>
__global__ __launch_bounds__(256) void k_pair(DevSys S) {
    pair_body<R, G, VIR, ROLLED>(S, blockIdx.x, gridDim.x);
}

#define MDF_PAIR_INST(G)                                            \
    template __global__ void k_pair<double, G, true, false>(DevSys);  \
    template __global__ void k_pair<float, G, true, false>(DevSys);   \
    template __global__ void k_pair<double, G, false, false>(DevSys); \
    template __global__ void k_pair<float, G, false, false>(DevSys);  \
    template __global__ void k_pair<double, G, true, true>(DevSys);   \
    template __global__ void k_pair<float, G, true, true>(DevSys);    \
    template __global__ void k_pair<double, G, false, true>(DevSys);  \
    template __global__ void k_pair<float, G, false, true>(DevSys);
MDF_PAIR_INST(8)
MDF_PAIR_INST(16)
MDF_PAIR_INST(32)
#undef MDF_PAIR_INST

// forces of moving sites that are neither leads nor lead-carried are zero before the
// reciprocal part: clear every moving-site force the pair kernel does not write
__global__ void k_zero_forces(DevSys S) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < S.n_move) { S.fx[t] = 0.0; S.fy[t] = 0.0; S.fz[t] = 0.0; }
}

}  // namespace mdf

namespace mdf {
// batched form (one lanes-per-lead choice for the whole batch; unrolled site-pair loop)
template <typename R, int G, bool VIR>
__global__ __launch_bounds__(256) void k_pair_b(const RepArgs *__restrict__ RA, const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    pair_body<R, G, VIR, false>(RA[m.x].S, m.y, 0);
}
#define MDF_PAIR_B_INST(G)                                                                   \
    template __global__ void k_pair_b<double, G, true>(const RepArgs *, const int2 *);    \
    template __global__ void k_pair_b<double, G, false>(const RepArgs *, const int2 *);   \
    template __global__ void k_pair_b<float, G, true>(const RepArgs *, const int2 *);     \
    template __global__ void k_pair_b<float, G, false>(const RepArgs *, const int2 *);
MDF_PAIR_B_INST(8)
MDF_PAIR_B_INST(16)
MDF_PAIR_B_INST(32)
#undef MDF_PAIR_B_INST
}  // namespace mdf
