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
// reduction; the lanes' energy classes are summed in a fixed butterfly.  Arithmetic type R
// is double (FP64) or float (FP32 with FP64 accumulation);
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
    for (int it = 0; it < 2; it++) {
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

// e^x for x <= 0 (the erfc and Buckingham arguments here), FP64, by Cody and Waite's
// argument reduction (Software Manual for the Elementary Functions, 1980): x = n ln2 + r with
// n = rint(x / ln2), |r| <= ln2 / 2, ln2 split into a 21-bit head (n ln2_hi exact) and a tail;
// e^r by its Taylor series to degree 13, whose truncation error at |r| = ln2 / 2 is 4.2e-18
// relative; e^x = 2^n e^r by exponent scaling (underflowing to 0 below x = -745).  About 19
// operations and within 2 ulp, without the range checks a general exp needs.
__device__ __forceinline__ double exp_neg(double x) {
    x = x < -800.0 ? -800.0 : x;   // e^-800 is 0 in FP64; NaN passes through
    const double n = rint(x * 1.4426950408889634);
    double r = __builtin_fma(-n, 0x1.62e42p-1, x);
    r = __builtin_fma(-n, 0x1.fdf473de6af28p-22, r);
    double p = 1.0 / 6227020800.0;            // 1/13!
    p = __builtin_fma(p, r, 1.0 / 479001600.0);
    p = __builtin_fma(p, r, 1.0 / 39916800.0);
    p = __builtin_fma(p, r, 1.0 / 3628800.0);
    p = __builtin_fma(p, r, 1.0 / 362880.0);
    p = __builtin_fma(p, r, 1.0 / 40320.0);
    p = __builtin_fma(p, r, 1.0 / 5040.0);
    p = __builtin_fma(p, r, 1.0 / 720.0);
    p = __builtin_fma(p, r, 1.0 / 120.0);
    p = __builtin_fma(p, r, 1.0 / 24.0);
    p = __builtin_fma(p, r, 1.0 / 6.0);
    p = __builtin_fma(p, r, 0.5);
    p = __builtin_fma(p, r, 1.0);
    p = __builtin_fma(p, r, 1.0);
    return ldexp(p, (int)n);
}
__device__ __forceinline__ double exp_r(double x) { return exp_neg(x); }
__device__ __forceinline__ float exp_r(float x) { return expf(x); }

// ACC = false: the reference's A&S 7.1.26 erfc (src/forcem.f90:11-13,335-337); true: libm erfc
template <typename R, bool ACC>
struct Coul {
    R alfa, spii, p, a1, a2, a3, a4, a5;
    // energy (erfc(ar) q/r) and force/r  for a site pair at squared distance r2,
    // with qq = facq q_i q_j
    __device__ __forceinline__ void eval(R r2, R qq, R &e, R &f) const {
        const R ri = rsqrt_nr(r2);      // 1/r
        const R r = r2 * ri;
        const R r2i = ri * ri;
        const R aux = qq * ri;
        const R ar = alfa * r;
        const R ex = exp_r(-ar * ar);
        R erfx;
        if constexpr (!ACC) {
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
        R aux = A * exp_r(-rij * rhoi);
        e = aux - r6;
        f = aux * rhoi * rinv - R(6) * rsqi * r6;
    } else {
        e = R(0);
        f = R(0);
    }
}

__device__ __forceinline__ double shfl_xor_d(double v, int m) { return __shfl_xor(v, m, 64); }

// one partner molecule as the kernel keeps it: lead position, the other sites as offsets
struct MolRec {
    double x, y, z, q[3];
    double rx[3], ry[3], rz[3];
};
__device__ __forceinline__ MolRec load_rec(const double *__restrict__ prec, int j) {
    const double4 *p = reinterpret_cast<const double4 *>(prec + (size_t)j * PREC);
    const double4 a = p[0], b = p[1], c = p[2];
    MolRec m;
    m.x = a.x; m.y = a.y; m.z = a.z; m.q[0] = a.w;
    m.rx[0] = 0.0; m.ry[0] = 0.0; m.rz[0] = 0.0;
    m.rx[1] = b.x; m.ry[1] = b.y; m.rz[1] = b.z; m.q[1] = b.w;
    m.rx[2] = c.x; m.ry[2] = c.y; m.rz[2] = c.z; m.q[2] = c.w;
    return m;
}

// c-th of n accumulators += v, with the index known only at run time (no dynamic register
// indexing: one select per accumulator)
template <int N>
__device__ __forceinline__ void acc_at(double (&a)[N], int c, double v) {
#pragma unroll
    for (int k = 0; k < N; k++) a[k] += (c == k) ? v : 0.0;
}

// the lead-lead minimum image and squared distance of the reference's loops (x, y always;
// z per variant), FP64 without contraction
__device__ __forceinline__ double lead_d2(const DevSys &S, double xi, double yi, double zi, double xj, double yj,
                                          double zj, double &dx, double &dy, double &dz) {
    dx = __dsub_rn(xi, xj); dx = __dsub_rn(dx, __dmul_rn(S.lx, anint_d(__dmul_rn(dx, S.lxi))));
    dy = __dsub_rn(yi, yj); dy = __dsub_rn(dy, __dmul_rn(S.ly, anint_d(__dmul_rn(dy, S.lyi))));
    dz = __dsub_rn(zi, zj);
    if (S.zper_force) dz = __dsub_rn(dz, __dmul_rn(S.lz, anint_d(__dmul_rn(dz, S.lzi))));
    return __dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), __dmul_rn(dz, dz));
}

constexpr int PAIR_SEG = 128;   // row entries per compaction segment

// G lanes share one lead.  The lead's moving row goes in segments of PAIR_SEG entries: first
// the G lanes test the lead-lead cut-off of the segment's partners (one 32-byte record load
// each) and append the partners inside it to the group's queue in LDS (a ballot prefix
// count, so the queue keeps the row's order); then the G lanes evaluate the queue, every
// lane busy, each partner's record loaded while the one before is evaluated.  STRIDED
// skips the compaction: lane sub takes row entries sub, sub + G, ... directly, and a lane
// whose partner is outside the cut-off idles through that pair.  All nine site
// pairs of the two molecules are evaluated without branches: records carry zero charges
// (and zero offsets) past a molecule's sites, which add exact zeros.  Partner vdW parameters
// come from a small LDS table indexed by (lead slot, partner class).  Energies per partner
// class, the virial and the wall forces stay in registers; the G lanes reduce them with
// shuffles at the end.  The cut-off test and its arithmetic order are the reference's
// (lead-pair minimum image and cut-off test in FP64 without contraction).
template <typename R, int G, bool VIR, bool ACC, bool STRIDED>
__device__ __forceinline__ void pair_body(const DevSys &S, int bid, int nblk) {
    __shared__ double s_tab[MAX_LEAD_TYPES * NCLS * PTAB_W];
    __shared__ int s_q[256 / G][PAIR_SEG];
    for (int k = threadIdx.x; k < MAX_LEAD_TYPES * NCLS * PTAB_W; k += blockDim.x) s_tab[k] = S.ptab[k];
    __syncthreads();
    const int tid = threadIdx.x;
    const int grp = tid / G;
    const int gid = bid * (blockDim.x / G) + grp;
    const int sub = tid % G;
    const int gbase = (tid & 63) - sub;   // the group's first lane in the wave
    const bool active = gid < S.n_lead;
    int *queue = s_q[grp];

    double f[3][3];
    double vir[6];
    double ev[NCLS], ec[NCLS], w[6];
#pragma unroll
    for (int a = 0; a < 3; a++) f[a][0] = f[a][1] = f[a][2] = 0.0;
#pragma unroll
    for (int c = 0; c < 6; c++) { vir[c] = 0.0; w[c] = 0.0; }
#pragma unroll
    for (int c = 0; c < NCLS; c++) { ev[c] = 0.0; ec[c] = 0.0; }

    Coul<R, ACC> cl;
    cl.alfa = (R)S.alfa; cl.spii = (R)S.spii; cl.p = (R)S.p;
    cl.a1 = (R)S.a1; cl.a2 = (R)S.a2; cl.a3 = (R)S.a3; cl.a4 = (R)S.a4; cl.a5 = (R)S.a5;

    int i = 0, nsi_i = 0;
    if (active) {
        i = S.lead[gid];
        const int mi = S.pmeta[i];
        const int slot_i = mi & 15;
        nsi_i = mi >> 4;
        const MolRec me = load_rec(S.prec, i);
        const double xi = me.x, yi = me.y, zi = me.z;
        double qi[3];
#pragma unroll
        for (int a = 0; a < 3; a++) qi[a] = S.facq * me.q[a];   // zero past nsi_i
        const double *tab_i = s_tab + slot_i * NCLS * PTAB_W;
        // ---- moving partners (loop A, src/forcem.f90:288-403) ----
        const int n = min(S.cnt_mm[gid], S.cap_mm);
        const int *row = S.nbr_mm + (size_t)gid * S.cap_mm;
        const double4 *prec4 = reinterpret_cast<const double4 *>(S.prec);
        // one molecule pair: lead pair (vdW + Coulomb) and the other eight site pairs
        auto mol_pair = [&](int mj, const MolRec &pj) {
            double dx, dy, dz;
            const double rsq = lead_d2(S, xi, yi, zi, pj.x, pj.y, pj.z, dx, dy, dz);
            if (STRIDED && rsq > S.rcutsq) return;
            const int cj = mj & 15;
            const double *pt = tab_i + cj * PTAB_W;
            R evd, fvd;
            vdw_eval<R>((int)pt[0], (R)rsq, (R)pt[1], (R)pt[2], (R)pt[3], (R)pt[4], (R)pt[5], evd, fvd);
            double e_c = 0.0;
#pragma unroll
            for (int b = 0; b < 3; b++) {
#pragma unroll
                for (int a = 0; a < 3; a++) {
                    double xd = dx, yd = dy, zd = dz, r2 = rsq;
                    if (a != 0 || b != 0) {   // (0, 0) is the lead pair, at the cut-off test's rsq
                        xd = (dx + me.rx[a]) - pj.rx[b];
                        yd = (dy + me.ry[a]) - pj.ry[b];
                        zd = (dz + me.rz[a]) - pj.rz[b];
                        r2 = xd * xd + yd * yd + zd * zd;
                    }
                    R es, fs;
                    cl.eval((R)r2, (R)(qi[a] * pj.q[b]), es, fs);
                    e_c += (double)es;
                    const double fd = (a == 0 && b == 0) ? (double)fvd + (double)fs : (double)fs;
                    f[a][0] += xd * fd; f[a][1] += yd * fd; f[a][2] += zd * fd;
                    if (VIR) {
                        const double hf = 0.5 * fd;
                        vir[0] += xd * xd * hf; vir[1] += xd * yd * hf; vir[2] += xd * zd * hf;
                        vir[3] += yd * yd * hf; vir[4] += yd * zd * hf; vir[5] += zd * zd * hf;
                    }
                }
            }
            // moving partners are classes 0 .. MAX_LEAD_TYPES - 1 (their lead slot); the full
            // list visits each pair from both sides
            double evh[MAX_LEAD_TYPES], ech[MAX_LEAD_TYPES];
#pragma unroll
            for (int c = 0; c < MAX_LEAD_TYPES; c++) { evh[c] = ev[c]; ech[c] = ec[c]; }
            acc_at<MAX_LEAD_TYPES>(evh, cj, 0.5 * (double)evd);
            acc_at<MAX_LEAD_TYPES>(ech, cj, 0.5 * e_c);
#pragma unroll
            for (int c = 0; c < MAX_LEAD_TYPES; c++) { ev[c] = evh[c]; ec[c] = ech[c]; }
        };
        if constexpr (STRIDED) {
            // lane sub takes row entries sub, sub + G, ...; the next partner's record in flight
            int k = sub;
            int jn = k < n ? row[k] : i;
            int mn = S.pmeta[jn];
            MolRec rn = load_rec(S.prec, jn);
#pragma unroll 1
            for (; k < n; k += G) {
                const int mj = mn;
                const MolRec pj = rn;
                if (k + G < n) {
                    jn = row[k + G];
                    mn = S.pmeta[jn];
                    rn = load_rec(S.prec, jn);
                }
                mol_pair(mj, pj);
            }
        } else {
#pragma unroll 1
            for (int s0 = 0; s0 < n; s0 += PAIR_SEG) {
                const int se = min(n, s0 + PAIR_SEG);
                // compaction: the segment's partners inside the cut-off, in row order
                // every load of the segment is issued before the tests: the lane's row entries, then
                // their partners' lead positions
                int qn = 0;
                constexpr int PER = PAIR_SEG / G;
                int jj[PER];
                double4 hh[PER];
#pragma unroll
                for (int u = 0; u < PER; u++) {
                    const int k = s0 + u * G + sub;
                    jj[u] = k < se ? row[k] : -1;
                }
#pragma unroll
                for (int u = 0; u < PER; u++)
                    hh[u] = jj[u] >= 0 ? prec4[(size_t)jj[u] * (PREC / 4)] : make_double4(0.0, 0.0, 0.0, 0.0);
#pragma unroll
                for (int u = 0; u < PER; u++) {
                    if (s0 + u * G >= se) break;   // uniform over the group
                    double dx, dy, dz;
                    const bool in = jj[u] >= 0 &&
                                    (STRIDED || !(lead_d2(S, xi, yi, zi, hh[u].x, hh[u].y, hh[u].z, dx, dy, dz) > S.rcutsq));
                    const unsigned long long m = (__ballot(in) >> gbase) & ((G == 64) ? ~0ull : ((1ull << G) - 1ull));
                    if (in) queue[qn + __popcll(m & ((1ull << sub) - 1ull))] = jj[u];
                    qn += __popcll(m);
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                // evaluation of the queue, the next partner's record in flight
                int q = sub;
                int jn = q < qn ? queue[q] : i;
                int mn = S.pmeta[jn];
                MolRec rn = load_rec(S.prec, jn);
#pragma unroll 1
                for (; q < qn; q += G) {
                    const int mj = mn;
                    const MolRec pj = rn;
                    if (q + G < qn) {
                        jn = queue[q + G];
                        mn = S.pmeta[jn];
                        rn = load_rec(S.prec, jn);
                    }
                    mol_pair(mj, pj);
                }
                // the queue is rewritten by the next segment
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
        }
        // surface-site bookkeeping between the two loops (src/forcem.f90:408-413)
        if (i >= S.n_liquid_sites) w[(signbit(zi) ? 0 : 3) + 2] += f[0][2];
        // ---- frozen partners (loop B, src/forcem.f90:418-498) ----
        const int nw = min(S.cnt_mw[gid], S.cap_mw);
        const int *wrow = S.nbr_mw + (size_t)gid * S.cap_mw;
        const bool wall_bk = S.itype[i] <= S.n_ltype;
#pragma unroll 1
        for (int kw = sub; kw < nw; kw += G) {
            const int wj = wrow[kw] - S.n_move;
            const double4 pw = S.wrec[wj];
            double dx, dy, dz;
            const double rsq = lead_d2(S, xi, yi, zi, pw.x, pw.y, pw.z, dx, dy, dz);
            if (rsq > S.rcutsq) continue;
            const int sgo = signbit(pw.z) ? 0 : 3;
            const int cw = S.wcls[wj];   // 0 / 1: the two crystal types, classes MAX_LEAD_TYPES + cw
            const double *pt = tab_i + (MAX_LEAD_TYPES + cw) * PTAB_W;
            R evd, fvd;
            vdw_eval<R>((int)pt[0], (R)rsq, (R)pt[1], (R)pt[2], (R)pt[3], (R)pt[4], (R)pt[5], evd, fvd);
            double e_c = 0.0;
            const double qw = S.facq * pw.w;
            double fij = (double)fvd;
            double wx = 0.0, wy = 0.0, wz = 0.0;
#pragma unroll
            for (int a = 0; a < 3; a++) {
                const double xd = dx + me.rx[a], yd = dy + me.ry[a], zd = dz + me.rz[a];
                const double r2 = (a == 0) ? rsq : xd * xd + yd * yd + zd * zd;
                R es, fs;
                cl.eval((R)r2, (R)(qw * me.q[a]), es, fs);
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
            if (wall_bk) {
                if (sgo == 0) { w[0] += wx; w[1] += wy; w[2] += wz; }
                else { w[3] += wx; w[4] += wy; w[5] += wz; }
            }
            double evw[2] = {ev[MAX_LEAD_TYPES], ev[MAX_LEAD_TYPES + 1]};
            double ecw[2] = {ec[MAX_LEAD_TYPES], ec[MAX_LEAD_TYPES + 1]};
            acc_at<2>(evw, cw, (double)evd);
            acc_at<2>(ecw, cw, e_c);
            ev[MAX_LEAD_TYPES] = evw[0]; ev[MAX_LEAD_TYPES + 1] = evw[1];
            ec[MAX_LEAD_TYPES] = ecw[0]; ec[MAX_LEAD_TYPES + 1] = ecw[1];
        }
    }
    // ---- reduce over the G lanes of the group (fixed butterfly) ----
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
#pragma unroll
        for (int c = 0; c < NCLS; c++) { ev[c] += shfl_xor_d(ev[c], m); ec[c] += shfl_xor_d(ec[c], m); }
#pragma unroll
        for (int c = 0; c < 6; c++) w[c] += shfl_xor_d(w[c], m);
    }
    if (active && sub == 0) {
        for (int a = 0; a < nsi_i && a < 3; a++) {
            S.fx[i + a] = f[a][0]; S.fy[i + a] = f[a][1]; S.fz[i + a] = f[a][2];
        }
        const size_t L = (size_t)S.n_lead;
        double *P = S.part;
#pragma unroll
        for (int c = 0; c < NCLS; c++) {
            P[(COL_EV + c) * L + gid] = ev[c];
            P[(COL_EC + c) * L + gid] = ec[c];
        }
#pragma unroll
        for (int c = 0; c < 6; c++) P[(COL_FWS + c) * L + gid] = w[c];
        const double v9[9] = {vir[0], vir[1], vir[2], vir[1], vir[3], vir[4], vir[2], vir[4], vir[5]};
#pragma unroll
        for (int c = 0; c < 9; c++) P[(COL_VIR + c) * L + gid] = VIR ? v9[c] : 0.0;
    }
}

template <typename R, int G, bool VIR, bool ACC, bool STRIDED>
__global__ __launch_bounds__(256) void k_pair(DevSys S) {
    pair_body<R, G, VIR, ACC, STRIDED>(S, blockIdx.x, gridDim.x);
}

#define MDF_PAIR_INST2(G, ST)                                           \
    template __global__ void k_pair<double, G, true, false, ST>(DevSys);  \
    template __global__ void k_pair<float, G, true, false, ST>(DevSys);   \
    template __global__ void k_pair<double, G, false, false, ST>(DevSys); \
    template __global__ void k_pair<float, G, false, false, ST>(DevSys);  \
    template __global__ void k_pair<double, G, true, true, ST>(DevSys);   \
    template __global__ void k_pair<float, G, true, true, ST>(DevSys);    \
    template __global__ void k_pair<double, G, false, true, ST>(DevSys);  \
    template __global__ void k_pair<float, G, false, true, ST>(DevSys);
#define MDF_PAIR_INST(G) MDF_PAIR_INST2(G, false) MDF_PAIR_INST2(G, true)
MDF_PAIR_INST(8)
MDF_PAIR_INST(16)
MDF_PAIR_INST(32)
#undef MDF_PAIR_INST2
#undef MDF_PAIR_INST

// forces of moving sites that are neither leads nor lead-carried are zero before the
// reciprocal part: clear every moving-site force the pair kernel does not write
__global__ void k_zero_forces(DevSys S) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < S.n_move) { S.fx[t] = 0.0; S.fy[t] = 0.0; S.fz[t] = 0.0; }
}

}  // namespace mdf

namespace mdf {
// batched form (one lanes-per-lead choice, loop form and erfc form for the whole batch)
template <typename R, int G, bool VIR, bool ACC, bool STRIDED>
__global__ __launch_bounds__(256) void k_pair_b(const RepArgs *__restrict__ RA, const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    pair_body<R, G, VIR, ACC, STRIDED>(RA[m.x].S, m.y, 0);
}
#define MDF_PAIR_B_INST2(G, ST)                                                                        \
    template __global__ void k_pair_b<double, G, true, false, ST>(const RepArgs *, const int2 *);  \
    template __global__ void k_pair_b<double, G, false, false, ST>(const RepArgs *, const int2 *); \
    template __global__ void k_pair_b<float, G, true, false, ST>(const RepArgs *, const int2 *);   \
    template __global__ void k_pair_b<float, G, false, false, ST>(const RepArgs *, const int2 *);  \
    template __global__ void k_pair_b<double, G, true, true, ST>(const RepArgs *, const int2 *);   \
    template __global__ void k_pair_b<double, G, false, true, ST>(const RepArgs *, const int2 *);  \
    template __global__ void k_pair_b<float, G, true, true, ST>(const RepArgs *, const int2 *);    \
    template __global__ void k_pair_b<float, G, false, true, ST>(const RepArgs *, const int2 *);
#define MDF_PAIR_B_INST(G) MDF_PAIR_B_INST2(G, false) MDF_PAIR_B_INST2(G, true)
MDF_PAIR_B_INST(8)
MDF_PAIR_B_INST(16)
MDF_PAIR_B_INST(32)
#undef MDF_PAIR_B_INST2
#undef MDF_PAIR_B_INST
}  // namespace mdf
