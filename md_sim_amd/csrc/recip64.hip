// FP64 reciprocal-space Ewald sum on the matrix cores (v_mfma_f64_16x16x4_f64), the FP64
// form of the reference's k-sum (src/forcem.f90:827-918).
//
// The phase of site j at k = (kl, km, kn) factorises: e^{ik.r_j} = P_j(row) E_j(kn) with
// P_j = q_j e^{i pi (kl u_j + km v_j)} per k-set row (kl, km) and E_j = e^{i pi kn w_j}.  The
// rows' kn ranges are symmetric about 0 (up to the half-space row (0, 0), which keeps kn > 0),
// and E_j(-n) = conj E_j(n), so with c = cos(n pi w_j), s = sin(n pi w_j) for n >= 0
//   S(row, +n) = sum_j P_j (c + i s),   S(row, -n) = sum_j P_j (c - i s):
// both come from the four real sums sum Pr c, sum Pi c, sum Pr s, sum Pi s.  Over a block of
// sites these are real matrix products [Pr; Pi] (rows x sites) . [c_0 c_1 s_1 c_2 s_2 ...]
// (sites x columns) - 2 FMAs per k-vector and site instead of the 4 of a complex product.
// The back-projection factorises the same way: with T = A conj S, U = T(+n) + T(-n),
// V = T(+n) - T(-n),
//   G_j(row)  = sum_kn T E_j     = sum_n (U c + i V s),
//   Gz_j(row) = sum_kn kn T E_j  = sum_n n (V c + i U s),
//   F_j = q_j sum_rows (2 pi / L) (kl, km, kn) Im(P_j G) (Im(P_j Gz) for z),
// so per row four real outputs (Re and Im of G and Gz) are a product of a (4 x columns)
// coefficient matrix with the (columns x sites) cos/sin matrix: 4 FMAs per k-vector and site.
//
// Operand layout of the instruction (one wave): lane l supplies A[l % 16][l / 16] and
// B[l / 16][l % 16] and receives D[4 e + l / 16][l % 16], e = 0..3 (checked on the device).
#include "mdf_internal.h"

namespace mdf {

typedef double d4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ d4v mfma64(double a, double b, d4v c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// e^m from e, e^2, e^4, ... (m below 2^SF64_POWS)
__device__ __forceinline__ double2 cpow_bits(const double2 *pw, int m) {
    double2 r = make_double2(1.0, 0.0);
#pragma unroll
    for (int b = 0; b < SF64_POWS; b++)
        if (m & (1 << b)) r = cmul(r, pw[b]);
    return r;
}

// ---- structure factor -------------------------------------------------------------------
// Grid (site blocks, passes); a block of SF64_WAVES waves takes a run of sb_sites sites in
// stages of SF64_SUB and, in one pass, a set of work slots: a slot is one 16-column tile of
// one 16-row tile of the k-set rows sorted by kn extent, and a wave holds up to SF64_SLOTS
// slots (consecutive in row-tile order, so at most three row tiles), dealt so that the four
// SIMDs of a compute unit get about equal numbers.  Per stage the block builds, in LDS, the
// base phases e^{i pi u}, e^{i pi v}, e^{i pi w} of the stage's sites and their powers 2, 4,
// 8; then q e^{i pi kl u} (kl = 0..klmax), e^{i pi km v} (km = 0..kmmax; negative km by
// conjugation) and the cos/sin columns the pass reaches: once per site and pass (at slab16k
// one pass holds every row).  Each wave forms the A operand P = (q X)(Y) of its slots' rows
// per lane and runs two products per slot and K-step of 4 sites.  Partial sums per site block
// go to partial[sb][pcoff[row] + col] (columns of a row: 2 E + 1), summed by k_sf64_reduce.
struct Sf64Args {
    const double *x, *y, *z, *q;
    int n;
    double cux, cuy, cuz;
    const int2 *rowkk;
    const int4 *waves;      // [pass][SF64_WAVES] slots (rt << 8 | ct) and count | same1 << 4 | same2 << 5
    const int *pass_ncol;   // cos/sin columns a pass reaches
    int n_rows, klmax, kmmax;
    const int *pcoff;
    int pcols;
    double2 *partial;
    int sb_sites;
    int bx, by;   // the block: site block, pass
};

// the stage loop of one wave with N slots; S1 (S2): slot 1 (2) is in the row tile of slot 0
// (1), so its A operand is the same; NT sizes the LDS column rows of the launch.  Every wave
// of the block builds tables and meets the barriers.  The waves of a block run different
// instantiations, so all LDS they share is the kernel's: sQ is passed in.
template <int N, bool S1, bool S2, int NT>
__device__ __forceinline__ void sf64_run(const Sf64Args &A, const int4 wk, double *sQ) {
    extern __shared__ __align__(16) double sf64_smem[];
    const int klp = A.klmax + 1, kmp = A.kmmax + 1;
    constexpr int ep = sf64_ep(NT);
    double2 *sPow = reinterpret_cast<double2 *>(sf64_smem);    // [SUB][3][SF64_POWS]
    double2 *sQX = sPow + SF64_SUB * 3 * SF64_POWS;             // [SUB][klp]
    double2 *sY = sQX + SF64_SUB * klp;                         // [SUB][kmp]
    double *sE = reinterpret_cast<double *>(sY + SF64_SUB * kmp);   // [SUB][ep]
    const int lane = threadIdx.x & 63;
    const int g = lane >> 4, c = lane & 15;
    const int ncol = A.pass_ncol[A.by];
    constexpr int NA = N > 0 ? N : 1;
    const int sv[3] = {wk.x, wk.y, wk.z};
    int rt[NA], ct[NA], my_kl[NA], my_km[NA];
    double ysg[NA];
#pragma unroll
    for (int s = 0; s < NA; s++) {
        rt[s] = N > 0 ? sv[s] >> 8 : 0;
        ct[s] = N > 0 ? sv[s] & 255 : 0;
        const int2 kk = A.rowkk[16 * rt[s] + c];   // padded past n_rows with (0, 0)
        my_kl[s] = kk.x;
        my_km[s] = abs(kk.y);
        ysg[s] = kk.y < 0 ? -1.0 : 1.0;   // e^{-i pi km v} = conj e^{i pi km v}
    }
    const int s_begin = A.bx * A.sb_sites, s_end = min(A.n, s_begin + A.sb_sites);
    d4v accR[NA], accI[NA];
#pragma unroll
    for (int s = 0; s < NA; s++) {
        accR[s] = d4v{0.0, 0.0, 0.0, 0.0};
        accI[s] = d4v{0.0, 0.0, 0.0, 0.0};
    }
    // thread (site j, axis ax) of a stage's base phases holds the next stage's coordinate and
    // charge in registers: loaded during the products of the stage before
    const int pj = threadIdx.x / 3, pax = threadIdx.x - 3 * pj;
    const bool pthr = threadIdx.x < 3 * SF64_SUB;
    double pcoord = 0.0, pq = 0.0;
    auto fetch = [&](int s0) {
        const int i = s0 + pj;
        const bool ok = pthr && i < s_end;
        pcoord = ok ? (pax == 0 ? A.x[i] * A.cux : pax == 1 ? A.y[i] * A.cuy : A.z[i] * A.cuz) : 0.0;
        pq = ok ? A.q[i] : 0.0;
    };
    // base phases and their powers of the fetched stage: one (site, axis) per thread
    auto phases = [&]() {
        if (pthr) {
            double sn, cs;
            sincospi(pcoord, &sn, &cs);
            double2 e = make_double2(cs, sn);
            double2 *pw = sPow + (pj * 3 + pax) * SF64_POWS;
#pragma unroll
            for (int b = 0; b < SF64_POWS; b++) {
                pw[b] = e;
                e = cmul(e, e);
            }
            if (pax == 0) sQ[pj] = pq;
        }
    };
    fetch(s_begin);
    phases();
    if (s_begin + SF64_SUB < s_end) fetch(s_begin + SF64_SUB);
    // table entries, by the first 512 threads of the block, each wave on one table: threads
    // 0-127 build q e^{i pi kl u} for kl = h, h + 4, ... (site tid / 4, h = tid % 4), threads
    // 128-255 e^{i pi km v} likewise, threads 256-511 the cos/sin columns n = h, h + 8, ...
    // (site (tid - 256) / 8, h < 8): a short product recursion from e^h (at most three
    // products) in steps of e^4 or e^8
    static_assert(SF64_SUB * 16 <= SF64_WAVES * 64, "16 table threads per site");
    const int ax = threadIdx.x < 4 * SF64_SUB ? 0 : threadIdx.x < 8 * SF64_SUB ? 1 : 2;
    const int tj = ax < 2 ? (threadIdx.x & (4 * SF64_SUB - 1)) >> 2 : (threadIdx.x - 8 * SF64_SUB) >> 3;
    const int ph = ax < 2 ? threadIdx.x & 3 : threadIdx.x & 7;
    const bool tthr = threadIdx.x < 16 * SF64_SUB;
    for (int s0 = s_begin; s0 < s_end; s0 += SF64_SUB) {
        // the tables are rewritten only after every wave has passed the products of the
        // stage before (the barrier), and the phases of this stage are in place
        __syncthreads();
        if (tthr) {
            const double2 *pw = sPow + (tj * 3 + ax) * SF64_POWS;
            double2 e = cpow_bits(pw, ph);
            if (ax == 0) {
                const double2 st = pw[2];   // e^4
                const double qj = sQ[tj];
                for (int kl = ph; kl < klp; kl += 4) {
                    sQX[tj * klp + kl] = make_double2(qj * e.x, qj * e.y);
                    e = cmul(e, st);
                }
            } else if (ax == 1) {
                const double2 st = pw[2];
                for (int m = ph; m < kmp; m += 4) {
                    sY[tj * kmp + m] = e;
                    e = cmul(e, st);
                }
            } else {
                const double2 st = pw[3];   // e^8
                for (int nn = ph; 2 * nn - 1 < ncol; nn += 8) {   // c_n at column 2n - 1, s_n at 2n
                    if (nn == 0) sE[tj * ep] = 1.0;
                    else {
                        sE[tj * ep + 2 * nn - 1] = e.x;
                        if (2 * nn < ncol) sE[tj * ep + 2 * nn] = e.y;
                    }
                    e = cmul(e, st);
                }
            }
        }
        __syncthreads();
        // the next stage's phases (their readers, the table build above, are done): the
        // threads that make them join the products afterwards, the other waves start at once
        if (s0 + SF64_SUB < s_end) {
            phases();
            if (s0 + 2 * SF64_SUB < s_end) fetch(s0 + 2 * SF64_SUB);
        }
        if constexpr (N > 0) {
            // one K-step at a time: the registers hold three slots' accumulators, not the
            // operands of several K-steps; the other waves of the SIMD cover the LDS latency
#pragma unroll 1
            for (int ks = 0; ks < SF64_SUB / 4; ks++) {
                const int js = 4 * ks + g;
                double pr[NA], pi_[NA];
#pragma unroll
                for (int s = 0; s < NA; s++) {
                    const bool same = (s == 1 && S1) || (s == 2 && S2);
                    if (same) {
                        pr[s] = pr[s - 1];
                        pi_[s] = pi_[s - 1];
                    } else {
                        const double2 X = sQX[js * klp + my_kl[s]], Y0 = sY[js * kmp + my_km[s]];
                        const double yr = Y0.x, yi = ysg[s] * Y0.y;
                        pr[s] = X.x * yr - X.y * yi;
                        pi_[s] = X.x * yi + X.y * yr;
                    }
                }
#pragma unroll
                for (int s = 0; s < NA; s++) {
                    const double b = sE[js * ep + 16 * ct[s] + c];
                    accR[s] = mfma64(pr[s], b, accR[s]);
                    accI[s] = mfma64(pi_[s], b, accI[s]);
                }
            }
        }
    }
    if constexpr (N > 0) {
        // lane holds rows 4 e + g of the slot's row tile at column 16 ct + c
#pragma unroll
        for (int s = 0; s < NA; s++) {
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int i = 16 * rt[s] + 4 * e + g;
                if (i >= A.n_rows) continue;
                const int off = A.pcoff[i], nc = A.pcoff[i + 1] - off;
                const int col = 16 * ct[s] + c;
                if (col < nc) A.partial[(size_t)A.bx * A.pcols + off + col] = make_double2(accR[s][e], accI[s][e]);
            }
        }
    }
}

template <int NT>
__device__ __forceinline__ void sf64_body(const Sf64Args &A) {
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   // uniform: slots in scalar registers
    const int4 wk = A.waves[A.by * SF64_WAVES + wave];
    __shared__ double sQ[SF64_SUB];
    switch (wk.w) {   // wave-uniform: one specialised stage loop per slot pattern
        case 1: sf64_run<1, false, false, NT>(A, wk, sQ); break;
        case 2: sf64_run<2, false, false, NT>(A, wk, sQ); break;
        case 2 | 16: sf64_run<2, true, false, NT>(A, wk, sQ); break;
        case 3: sf64_run<3, false, false, NT>(A, wk, sQ); break;
        case 3 | 16: sf64_run<3, true, false, NT>(A, wk, sQ); break;
        case 3 | 32: sf64_run<3, false, true, NT>(A, wk, sQ); break;
        case 3 | 48: sf64_run<3, true, true, NT>(A, wk, sQ); break;
        default: sf64_run<0, false, false, NT>(A, wk, sQ); break;
    }
}

template <int NT>
__global__ __launch_bounds__(64 * SF64_WAVES) void k_sf64(
    const double *__restrict__ x, const double *__restrict__ y, const double *__restrict__ z,
    const double *__restrict__ q, int n, double cux, double cuy, double cuz, const int2 *__restrict__ rowkk,
    const int4 *__restrict__ waves, const int *__restrict__ pass_ncol, int n_rows, int klmax, int kmmax,
    const int *__restrict__ pcoff, int pcols, double2 *__restrict__ partial, int sb_sites,
    const MtsState *__restrict__ mts) {
    if (mts && !mts->exact) return;   // a surrogate step: no k-sum
    const Sf64Args A{x, y, z, q, n, cux, cuy, cuz, rowkk, waves, pass_ncol, n_rows, klmax, kmmax, pcoff, pcols,
                     partial, sb_sites, (int)blockIdx.x, (int)blockIdx.y};
    sf64_body<NT>(A);
}
// batched: block map entry (system, local block), local block = pass * site blocks + site block
template <int NT>
__global__ __launch_bounds__(64 * SF64_WAVES) void k_sf64_b(const RepArgs *__restrict__ RA,
                                                            const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    const RepArgs &r = RA[m.x];
    const Sf64Args A{r.S.x, r.S.y, r.S.z, r.S.q, r.S.n_move, r.cux, r.cuy, r.cuz, r.s64_rowkk, r.s64_waves,
                     r.s64_pncol, r.nrows, r.klmax, r.kmmax, r.s64_pcoff, r.s64_pcols, r.s64_part, r.s64_sites,
                     m.y % r.s64_nsb, m.y / r.s64_nsb};
    sf64_body<NT>(A);
}

#define SF64_INST(NT)                                                                                            \
    template __global__ void k_sf64<NT>(const double *, const double *, const double *, const double *, int,     \
                                        double, double, double, const int2 *, const int4 *, const int *, int, int, \
                                        int, const int *, int, double2 *, int, const MtsState *);                \
    template __global__ void k_sf64_b<NT>(const RepArgs *, const int2 *);
SF64_INST(1) SF64_INST(2) SF64_INST(3) SF64_INST(4) SF64_INST(5) SF64_INST(6) SF64_INST(7) SF64_INST(8)
#undef SF64_INST

// ---- reduction, one block per k-set row (reference order) ---------------------------------
// Column sums over the site blocks (fixed order for the row's width: the block's threads are
// nsl slices of cw >= ncol columns, cw a multiple of 16; slice sl adds site blocks sl, sl + nsl,
// ..., then the slices are added in order),
// S(row, +-n) + S_wall, T = A conj S, the energy and virial partials of the row, and the
// back-projection's coefficient operand for the row: in M-tile m = row / 4 (4 rows x 4 outputs
// Re G, Im G, Re Gz, Im Gz; output e of row r is matrix row 4 e + r % 4), K-step ks of 4
// columns, lane (col % 4) * 16 + (4 e + r % 4): tpk[(bpoff[m] + ks) * 64 + lane].
constexpr int SF64R_COLS = 128;   // columns of a row: 2 E + 1 <= 127

// row r of the k-set (reference order)
__device__ __forceinline__ void sf64r_body(
    const double2 *__restrict__ partial, int n_sb, int pcols, const int *__restrict__ pcoff,
    const int *__restrict__ sorted_of, const KRow *__restrict__ rows, const double2 *__restrict__ s_wall,
    const double *__restrict__ ak, double pipix, double pipiy, double pipiz, double lamb,
    double2 *__restrict__ s_out, double4 *__restrict__ t_out, double *__restrict__ epart,
    const int2 *__restrict__ bptile, double *__restrict__ tpk, int r) {
    __shared__ double2 red[SFR_THREADS];   // [slice][cw]
    __shared__ double2 sSum[SF64R_COLS];
    __shared__ double2 sTp[64], sTm[64];   // T(row, +n), T(row, -n) (zero outside the row)
    const KRow row = rows[r];
    const int i = sorted_of[r];
    const int off = pcoff[i], ncol = pcoff[i + 1] - off, E = (ncol - 1) / 2;
    const int cw = 16 * ((ncol + 15) / 16), nsl = SFR_THREADS / cw;
    const int col = threadIdx.x % cw, sl = threadIdx.x / cw;
    double sr = 0.0, si = 0.0;
    if (col < ncol && sl < nsl) {
#pragma unroll 4
        for (int b = sl; b < n_sb; b += nsl) {
            const double2 v = partial[(size_t)b * pcols + off + col];
            sr += v.x; si += v.y;
        }
    }
    if (sl < nsl) red[threadIdx.x] = make_double2(sr, si);
    __syncthreads();
    if (threadIdx.x < SF64R_COLS) {
        double ar = 0.0, ai = 0.0;
        if (threadIdx.x < cw)
            for (int k = 0; k < nsl; k++) { ar += red[k * cw + threadIdx.x].x; ai += red[k * cw + threadIdx.x].y; }
        sSum[threadIdx.x] = make_double2(ar, ai);   // (sum Pr X, sum Pi X), X = column col
    }
    __syncthreads();
    const int t = threadIdx.x;
    if (t < 64) {
        double acc[RE_N];
#pragma unroll
        for (int cc = 0; cc < RE_N; cc++) acc[cc] = 0.0;
        double2 tp = make_double2(0.0, 0.0), tm = tp;
        if (t <= E) {
            const double2 C = t == 0 ? sSum[0] : sSum[2 * t - 1];            // (sum Pr c, sum Pi c)
            const double2 Sn = t == 0 ? make_double2(0.0, 0.0) : sSum[2 * t];  // (sum Pr s, sum Pi s)
            for (int sg = 1; sg >= -1; sg -= 2) {
                const int kn = sg * t;
                if (t == 0 && sg < 0) break;
                if (kn < row.kn0 || kn > row.kn1) continue;
                const int k = row.koff + (kn - row.kn0);
                // S_moving(+n) = (Prc - Pis) + i (Prs + Pic); S_moving(-n) = (Prc + Pis) + i (Pic - Prs)
                const double mr = C.x - sg * Sn.y, mi = C.y + sg * Sn.x;
                const double2 w = s_wall[k];
                const double a = ak[k];
                const double qr = mr + w.x, qi = mi + w.y;
                s_out[k] = make_double2(qr, qi);
                const double2 tk = make_double2(a * qr, -a * qi);
                t_out[k] = make_double4(tk.x, tk.y, kn * tk.x, kn * tk.y);
                if (sg > 0) tp = tk; else tm = tk;
                const double ek = a * (qr * qr + qi * qi);
                acc[0] += ek;
                acc[1] += a * (mr * mr + mi * mi);
                const double kx = pipix * row.kl, ky = pipiy * row.km, kz = pipiz * kn;
                const double gg = 2.0 * (1.0 / (kx * kx + ky * ky + kz * kz) + lamb);
                const double h = 0.5 * ek;
                acc[2] += h * (1.0 - gg * kx * kx);
                acc[3] += h * (1.0 - gg * ky * ky);
                acc[4] += h * (1.0 - gg * kz * kz);
                acc[5] += -h * gg * kx * ky;
                acc[6] += -h * gg * kx * kz;
                acc[7] += -h * gg * ky * kz;
            }
        }
        sTp[t] = tp;
        sTm[t] = tm;
#pragma unroll
        for (int cc = 0; cc < RE_N; cc++)
            for (int o = 32; o > 0; o >>= 1) acc[cc] += __shfl_xor(acc[cc], o, 64);
        if (t == 0)
#pragma unroll
            for (int cc = 0; cc < RE_N; cc++) epart[r * RE_N + cc] = acc[cc];
    }
    __syncthreads();
    // the row's coefficient operand: columns 0 .. 4 nks - 1 of its M-tile, 4 outputs each
    const int2 bt = bptile[r >> 2];   // (offset in K-steps, K-steps)
    const int ro = r & 3;
    for (int idx = threadIdx.x; idx < 4 * 4 * bt.y; idx += blockDim.x) {
        const int cl = idx >> 2, e = idx & 3;
        double v = 0.0;
        if (cl < ncol) {
            const int nn = (cl + 1) >> 1;
            const bool cosine = cl == 0 || (cl & 1);
            double2 U, V;
            if (nn == 0) { U = sTp[0]; V = make_double2(0.0, 0.0); }
            else {
                const double2 a = sTp[nn], b = sTm[nn];
                U = make_double2(a.x + b.x, a.y + b.y);
                V = make_double2(a.x - b.x, a.y - b.y);
            }
            const double fn = (double)nn;
            if (cosine) v = e == 0 ? U.x : e == 1 ? U.y : e == 2 ? fn * V.x : fn * V.y;
            else v = e == 0 ? -V.y : e == 1 ? V.x : e == 2 ? -fn * U.y : fn * U.x;
        }
        tpk[((size_t)bt.x + (cl >> 2)) * 64 + (cl & 3) * 16 + 4 * e + ro] = v;
    }
}

__global__ __launch_bounds__(SFR_THREADS) void k_sf64_reduce(
    const double2 *__restrict__ partial, int n_sb, int pcols, const int *__restrict__ pcoff,
    const int *__restrict__ sorted_of, const KRow *__restrict__ rows, const double2 *__restrict__ s_wall,
    const double *__restrict__ ak, double pipix, double pipiy, double pipiz, double lamb,
    double2 *__restrict__ s_out, double4 *__restrict__ t_out, double *__restrict__ epart,
    const int2 *__restrict__ bptile, double *__restrict__ tpk, const MtsState *__restrict__ mts) {
    if (mts && !mts->exact) return;
    sf64r_body(partial, n_sb, pcols, pcoff, sorted_of, rows, s_wall, ak, pipix, pipiy, pipiz, lamb, s_out, t_out,
               epart, bptile, tpk, blockIdx.x);
}
// batched: block map entry (system, row)
__global__ __launch_bounds__(SFR_THREADS) void k_sf64_reduce_b(const RepArgs *__restrict__ RA,
                                                               const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    const RepArgs &r = RA[m.x];
    sf64r_body(r.s64_part, r.S.n_move > 0 ? r.s64_nsb : 0, r.s64_pcols, r.s64_pcoff, r.s64_sorted_of, r.rows,
               r.s_wall, r.ak, r.pipix, r.pipiy, r.pipiz, r.lamb, r.s_out, r.t_out, r.kpart, r.b64_tile, r.b64_tpk,
               m.y);
}

// ---- back-projection ------------------------------------------------------------------------
// Grid (site blocks, row groups); a wave owns NG groups of 16 sites and keeps their cos/sin
// operand (column 4 ks + g for K-step ks, lane group g) in registers.  The block stages each
// M-tile's coefficient operand (nks x 64 doubles) in LDS one tile ahead; per M-tile and K-step
// every wave runs NG products.  Lane (g, c) then holds Re G, Im G, Re Gz, Im Gz of row 4 m + g
// at its site and contracts them with P of that row, which walks the rows of its lane group
// (km + 4 at the same kl by one complex product, a fresh sincospi otherwise).  The four lane
// groups of a site add up at the end; row groups write their own partial forces.
constexpr int BP64_WAVES = 4;

// N K-steps of a wave's products (N a compile-time count: no branches between them)
template <int N, int NKS, int NG>
__device__ __forceinline__ void bp64_products(const double *sa, int lane, const double (&Bv)[NG][NKS], d4v (&acc)[NG]) {
#pragma unroll
    for (int ks = 0; ks < N; ks++) {
        const double a = sa[ks * 64 + lane];
#pragma unroll
        for (int ng = 0; ng < NG; ng++) acc[ng] = mfma64(a, Bv[ng][ks], acc[ng]);
    }
}

// block (bx, by) of a (site blocks) x (row groups) grid
template <int NKS, int NG>
__device__ __forceinline__ void bp64_body(const double *__restrict__ x, const double *__restrict__ y,
                                          const double *__restrict__ z, const double *__restrict__ q, int n,
                                          double cux, double cuy, double cuz, const int4 *__restrict__ rinfo,
                                          int n_rows, const int2 *__restrict__ bptile,
                                          const int2 *__restrict__ rgroups, const double *__restrict__ tpk,
                                          double pipix, double pipiy, double pipiz, double *__restrict__ fpart,
                                          int bx, int by) {
    __shared__ double sA[2][NKS][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, c = lane & 15;
    const int2 rg = rgroups[by];   // M-tiles [rg.x, rg.y)
    const int site0 = (bx * BP64_WAVES + wave) * 16 * NG;
    // an M-tile's operand goes through registers: its global loads are issued before the
    // previous tile's products and written to LDS after them, so their latency is hidden
    constexpr int PF = (NKS * 64 + 64 * BP64_WAVES - 1) / (64 * BP64_WAVES);   // doubles per thread
    double pf[PF];
    auto fetch = [&](int m) {
        const int2 bt = bptile[m];
        const double *src = tpk + (size_t)bt.x * 64;
#pragma unroll
        for (int k = 0; k < PF; k++) {
            const int e = threadIdx.x + k * 64 * BP64_WAVES;
            pf[k] = e < bt.y * 64 ? src[e] : 0.0;
        }
    };
    auto put = [&](int b) {
#pragma unroll
        for (int k = 0; k < PF; k++) {
            const int e = threadIdx.x + k * 64 * BP64_WAVES;
            if (e < NKS * 64) sA[b][e >> 6][e & 63] = pf[k];
        }
    };
    if (rg.x < rg.y) { fetch(rg.x); put(0); }
    // cos/sin operands: lane group g takes columns 4 ks + g, i.e. n = 2 ks + n0 with
    // n0 = 0, 1, 1, 2 and cos for g = 1, 3 (and column 0), sin for g = 0, 2
    const int n0 = (g + 1) >> 1;
    const bool cosine = g & 1;
    double Bv[NG][NKS];
    double u[NG], v[NG], qq[NG];
    double2 v4[NG], P[NG];
#pragma unroll
    for (int ng = 0; ng < NG; ng++) {
        const int i = site0 + 16 * ng + c;
        const bool ok = i < n;
        u[ng] = ok ? x[i] * cux : 0.0;
        v[ng] = ok ? y[i] * cuy : 0.0;
        qq[ng] = ok ? q[i] : 0.0;
        const double w = ok ? z[i] * cuz : 0.0;
        double sn, cs;
        sincospi((double)n0 * w, &sn, &cs);
        double2 e = make_double2(cs, sn);
        sincospi(2.0 * w, &sn, &cs);
        const double2 e2 = make_double2(cs, sn);
#pragma unroll
        for (int ks = 0; ks < NKS; ks++) {
            Bv[ng][ks] = (cosine || (ks == 0 && g == 0)) ? e.x : e.y;
            e = cmul(e, e2);
        }
        sincospi(4.0 * v[ng], &sn, &cs);
        v4[ng] = make_double2(cs, sn);
        P[ng] = make_double2(0.0, 0.0);
    }
    double fx[NG], fy[NG], fz[NG];
#pragma unroll
    for (int ng = 0; ng < NG; ng++) { fx[ng] = 0.0; fy[ng] = 0.0; fz[ng] = 0.0; }
    for (int m = rg.x; m < rg.y; m++) {
        __syncthreads();   // tile m staged; the other buffer's readers are done
        const bool more = m + 1 < rg.y;
        if (more) fetch(m + 1);
        const int b = (m - rg.x) & 1;
        const int nks = bptile[m].y;
        const int r = 4 * m + g;
        const int4 ri = rinfo[min(r, n_rows - 1)];   // kl, km, fresh (row r - 4 is not (kl, km - 4))
        d4v acc[NG];
#pragma unroll
        for (int ng = 0; ng < NG; ng++) acc[ng] = d4v{0.0, 0.0, 0.0, 0.0};
        __builtin_amdgcn_s_setprio(1);
        // block-uniform K-step count: one branch-free product sequence per count
        switch (nks) {
#define BP64_CASE(N) \
    case N: if constexpr (N <= NKS) bp64_products<N, NKS, NG>(&sA[b][0][0], lane, Bv, acc); break;
            BP64_CASE(1) BP64_CASE(2) BP64_CASE(3) BP64_CASE(4) BP64_CASE(5) BP64_CASE(6) BP64_CASE(7)
            BP64_CASE(8) BP64_CASE(9) BP64_CASE(10) BP64_CASE(11) BP64_CASE(12) BP64_CASE(13) BP64_CASE(14)
            BP64_CASE(15) BP64_CASE(16) BP64_CASE(17) BP64_CASE(18) BP64_CASE(19) BP64_CASE(20) BP64_CASE(21)
            BP64_CASE(22) BP64_CASE(23) BP64_CASE(24) BP64_CASE(25) BP64_CASE(26) BP64_CASE(27) BP64_CASE(28)
            BP64_CASE(29) BP64_CASE(30) BP64_CASE(31) BP64_CASE(32)
#undef BP64_CASE
            default: break;
        }
        __builtin_amdgcn_s_setprio(0);
        if (more) put((m + 1 - rg.x) & 1);
        if (r < n_rows) {
            const bool fresh = ri.z || m == rg.x;
#pragma unroll
            for (int ng = 0; ng < NG; ng++) {
                if (fresh) {
                    double sn, cs;
                    sincospi((double)ri.x * u[ng] + (double)ri.y * v[ng], &sn, &cs);
                    P[ng] = make_double2(cs, sn);
                } else {
                    P[ng] = cmul(P[ng], v4[ng]);
                }
                const double im = P[ng].x * acc[ng][1] + P[ng].y * acc[ng][0];
                fx[ng] += (double)ri.x * im;
                fy[ng] += (double)ri.y * im;
                fz[ng] += P[ng].x * acc[ng][3] + P[ng].y * acc[ng][2];
            }
        }
    }
#pragma unroll
    for (int ng = 0; ng < NG; ng++) {
        double ax = fx[ng], ay = fy[ng], az = fz[ng];
        ax += __shfl_xor(ax, 16, 64); ax += __shfl_xor(ax, 32, 64);
        ay += __shfl_xor(ay, 16, 64); ay += __shfl_xor(ay, 32, 64);
        az += __shfl_xor(az, 16, 64); az += __shfl_xor(az, 32, 64);
        const int i = site0 + 16 * ng + c;
        if (g == 0 && i < n) {
            double *fp = fpart + (size_t)by * 3 * n;
            fp[i] = qq[ng] * pipix * ax;
            fp[(size_t)n + i] = qq[ng] * pipiy * ay;
            fp[2 * (size_t)n + i] = qq[ng] * pipiz * az;
        }
    }
}

template <int NKS, int NG>
__global__ __launch_bounds__(64 * BP64_WAVES) void k_bp64(const double *__restrict__ x, const double *__restrict__ y,
                                                         const double *__restrict__ z, const double *__restrict__ q,
                                                         int n, double cux, double cuy, double cuz,
                                                         const int4 *__restrict__ rinfo, int n_rows,
                                                         const int2 *__restrict__ bptile, const int2 *__restrict__ rgroups,
                                                         const double *__restrict__ tpk, double pipix, double pipiy,
                                                         double pipiz, double *__restrict__ fpart,
                                                         const MtsState *__restrict__ mts) {
    if (mts && !mts->exact) return;
    bp64_body<NKS, NG>(x, y, z, q, n, cux, cuy, cuz, rinfo, n_rows, bptile, rgroups, tpk, pipix, pipiy, pipiz, fpart,
                       blockIdx.x, blockIdx.y);
}
// batched: block map entry (system, local block), local block = row group * site blocks + site block
template <int NKS, int NG>
__global__ __launch_bounds__(64 * BP64_WAVES) void k_bp64_b(const RepArgs *__restrict__ RA, const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    const RepArgs &r = RA[m.x];
    bp64_body<NKS, NG>(r.S.x, r.S.y, r.S.z, r.S.q, r.S.n_move, r.cux, r.cuy, r.cuz, r.b64_rinfo, r.nrows, r.b64_tile,
                       r.b64_rg, r.b64_tpk, r.pipix, r.pipiy, r.pipiz, r.fpart, m.y % r.b64_nsb, m.y / r.b64_nsb);
}

#define BP64_INST(NKS, NG)                                                                                       \
    template __global__ void k_bp64<NKS, NG>(const double *, const double *, const double *, const double *, int,  \
                                             double, double, double, const int4 *, int, const int2 *, const int2 *, \
                                             const double *, double, double, double, double *, const MtsState *); \
    template __global__ void k_bp64_b<NKS, NG>(const RepArgs *, const int2 *);
BP64_INST(4, 1) BP64_INST(8, 1) BP64_INST(12, 1) BP64_INST(16, 1) BP64_INST(20, 1) BP64_INST(24, 1) BP64_INST(28, 1)
BP64_INST(32, 1) BP64_INST(4, 2) BP64_INST(8, 2) BP64_INST(12, 2) BP64_INST(16, 2) BP64_INST(20, 2) BP64_INST(24, 2)
BP64_INST(28, 2) BP64_INST(32, 2)
// three site groups per wave at exact K-step counts 9..20 (their operands fill the registers)
BP64_INST(9, 3) BP64_INST(10, 3) BP64_INST(11, 3) BP64_INST(12, 3) BP64_INST(13, 3) BP64_INST(14, 3) BP64_INST(15, 3)
BP64_INST(16, 3) BP64_INST(17, 3) BP64_INST(18, 3) BP64_INST(19, 3) BP64_INST(20, 3)
#undef BP64_INST

}  // namespace mdf
