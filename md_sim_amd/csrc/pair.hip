// Real-space pair kernel: Ewald erfc Coulomb + Lennard-Jones/Buckingham over the pair
// lists, moving-moving (reference loop A, src/forcem.f90:288-403) and moving-frozen
// (loop B, src/forcem.f90:418-498).
//
// Group cut-off on the lead sites, with every site pair of the two molecules evaluated
// when the leads are inside it (src/forcem.f90:299,360-402,460-497); vdW acts on the lead
// pair only; the erfc is the A&S 7.1.26 form with the exact-derivative force
// (src/forcem.f90:333-340).  A group of G lanes owns one lead: lanes take the partners of its
// list rows, accumulate the forces of the lead's own sites (the full list visits every pair
// from both sides, so no atomics and no write conflicts), then reduce with lane shuffles.
// Energies and virial of moving pairs are halved for the double visit; frozen pairs are
// visited once.  Per-lead sums go to column-major partial arrays for a deterministic
// reduction.  Arithmetic type R is double (FP64) or float (FP32 site-pair arithmetic with
// FP64 accumulation); the lead-pair minimum image and the cut-off test are FP64 without
// contraction in both, so the set of interacting pairs is the reference's.
#include "mdf_internal.h"
#include "mdf_math.h"

namespace mdf {

// FP32 form: the hardware estimates v_rsq_f32, v_rcp_f32 and v_exp_f32 (2^x), each within
// about an ulp, with no refinement: the FP32 form's error is set by its FP32 arithmetic
__device__ __forceinline__ float rsq_hw(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float rcp_hw(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float exp_hw(float x) { return __builtin_amdgcn_exp2f(x * 1.4426950408889634f); }

// ACC = false: the reference's A&S 7.1.26 erfc (src/forcem.f90:11-13,335-337); true: libm erfc
template <typename R, bool ACC>
struct Coul {
    R alfa, spii, p, a1, a2, a3, a4, a5;
    const double *e2t;   // 2^(j/64) in LDS (FP64 e^x)
    // energy (erfc(ar) q/r) and force/r  for a site pair at squared distance r2,
    // with qq = facq q_i q_j
    __device__ __forceinline__ void eval(R r2, R qq, R &e, R &f) const {
        R ri;
        if constexpr (sizeof(R) == 8) ri = rsqrt_nr(r2); else ri = rsq_hw(r2);   // 1/r
        const R r = r2 * ri;
        const R r2i = ri * ri;
        const R aux = qq * ri;
        const R ar = alfa * r;
        R ex;
        if constexpr (sizeof(R) == 8) ex = exp_neg(-ar * ar, e2t); else ex = exp_hw(-ar * ar);
        R erfx;
        if constexpr (!ACC) {
            R t;
            if constexpr (sizeof(R) == 8) t = rcp_nr(R(1) + p * ar); else t = rcp_hw(R(1) + p * ar);
            erfx = ((((a5 * t + a4) * t + a3) * t + a2) * t + a1) * t * ex;
        } else {
            erfx = erfc(ar);
        }
        e = erfx * aux;
        f = (spii * ar * ex + erfx) * aux * r2i;
    }
};

template <typename R>
__device__ __forceinline__ void vdw_eval(int kind, R rsq, R eps4, R sig6, R A, R rhoi, R Cc, const double *e2t, R &e,
                                         R &f) {
    R rsqi;
    if constexpr (sizeof(R) == 8) rsqi = rcp_nr(rsq); else rsqi = rcp_hw(rsq);
    if (kind == 1) {  // Lennard-Jones, eps4 = 4 eps, sig6 = sigma^6 (src/forcem.f90:315-319)
        R r6 = sig6 * (rsqi * rsqi * rsqi);
        e = eps4 * r6 * (r6 - R(1));
        f = R(6) * rsqi * eps4 * r6 * (R(2) * r6 - R(1));
    } else if (kind == 2) {  // Buckingham A exp(-r/rho) - C/r^6 (src/forcem.f90:320-325)
        R rinv;
        if constexpr (sizeof(R) == 8) rinv = rsqrt_nr(rsq); else rinv = rsq_hw(rsq);
        R rij = rsq * rinv;
        R r6 = Cc * (rsqi * rsqi * rsqi);
        R aux;
        if constexpr (sizeof(R) == 8) aux = A * exp_neg(-rij * rhoi, e2t); else aux = A * exp_hw(-rij * rhoi);
        e = aux - r6;
        f = aux * rhoi * rinv - R(6) * rsqi * r6;
    } else {
        e = R(0);
        f = R(0);
    }
}

// sum over the G lanes of an aligned group, the same in every lane of it: data-parallel
// lane moves (quad swaps, half-row and row mirrors) within 16 lanes, cross-lane shuffles
// beyond; a symmetric tree, so every lane adds the same pairs
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, CTRL, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), CTRL, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}
template <int G>
__device__ __forceinline__ double group_sum(double v) {
    if constexpr (G >= 2) v += dpp_d<0xB1>(v);    // quad_perm [1,0,3,2]
    if constexpr (G >= 4) v += dpp_d<0x4E>(v);    // quad_perm [2,3,0,1]
    if constexpr (G >= 8) v += dpp_d<0x141>(v);   // row_half_mirror
    if constexpr (G >= 16) v += dpp_d<0x140>(v);  // row_mirror
    if constexpr (G >= 32) v += __shfl_xor(v, 16, 64);
    if constexpr (G >= 64) v += __shfl_xor(v, 32, 64);
    return v;
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

// the wave-local barrier between the lanes that fill a group's queue and those that read it
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One row of a group of G lanes: every partner inside the lead-pair cut-off is handed to
// eval(j), in row order lane by lane.  Compacted form: the row goes in segments of PAIR_SEG
// entries; the G lanes test the segment's partners (one position load each, all issued before
// the tests) and append those inside the cut-off to the group's queue in LDS (ballot prefix
// count, so the queue keeps the row's order); then the lanes take the queue G entries at a
// time, the next entry's partner record in flight while one is evaluated (LOAD(j) -> record,
// EVAL(record); two records alternate, LOAD(-1) gives a record EVAL skips).  STRIDED: lane
// sub takes row entries sub, sub + G, ... directly and EVAL itself skips a partner outside
// the cut-off.
template <int G, bool STRIDED, typename POS, typename LOAD, typename EVAL>
__device__ __forceinline__ void row_pass(const DevSys &S, const int *row, int n, int *queue, int sub, int gbase,
                                         double xi, double yi, double zi, POS pos, LOAD load, EVAL eval) {
    if constexpr (STRIDED) {
        int k = sub;
        auto ra = load(k < n ? row[k] : -1);
#pragma unroll 1
        for (; k < n; k += 2 * G) {
            const auto rb = load(k + G < n ? row[k + G] : -1);
            eval(ra);
            ra = load(k + 2 * G < n ? row[k + 2 * G] : -1);
            eval(rb);
        }
    } else {
#pragma unroll 1
        for (int s0 = 0; s0 < n; s0 += PAIR_SEG) {
            const int se = min(n, s0 + PAIR_SEG);
            int qn = 0;
            constexpr int PER = PAIR_SEG / G;
            int jj[PER];
            double3 hh[PER];
#pragma unroll
            for (int u = 0; u < PER; u++) {
                const int k = s0 + u * G + sub;
                jj[u] = k < se ? row[k] : -1;
            }
#pragma unroll
            for (int u = 0; u < PER; u++) hh[u] = jj[u] >= 0 ? pos(jj[u]) : make_double3(0.0, 0.0, 0.0);
#pragma unroll
            for (int u = 0; u < PER; u++) {
                if (s0 + u * G >= se) break;   // uniform over the group
                double dx, dy, dz;
                const bool in = jj[u] >= 0 && !(lead_d2(S, xi, yi, zi, hh[u].x, hh[u].y, hh[u].z, dx, dy, dz) > S.rcutsq);
                const unsigned long long m = (__ballot(in) >> gbase) & ((G == 64) ? ~0ull : ((1ull << G) - 1ull));
                if (in) queue[qn + __popcll(m & ((1ull << sub) - 1ull))] = jj[u];
                qn += __popcll(m);
            }
            wave_sync();
            int q = sub;
            auto ra = load(q < qn ? queue[q] : -1);
#pragma unroll 1
            for (; q < qn; q += 2 * G) {
                const auto rb = load(q + G < qn ? queue[q + G] : -1);
                eval(ra);
                ra = load(q + 2 * G < qn ? queue[q + 2 * G] : -1);
                eval(rb);
            }
            wave_sync();   // the queue is rewritten by the next segment
        }
    }
}

// a moving partner as the evaluation takes it: lead position (FP64), the other two sites as
// offsets from the lead and the three charges, in R; cls = its lead slot (energy class)
// a frozen partner: position and charge, crystal class
struct WzRec {
    double4 p;
    int cw;
    bool ok;
};

template <typename R>
struct MvRec {
    double x, y, z;
    R rx[3], ry[3], rz[3], q[3];
    int cls;
    bool ok;
};

// G lanes share one lead: they take its moving row (loop A, src/forcem.f90:288-403) and its
// frozen row (loop B, src/forcem.f90:418-498) through row_pass, each lane accumulating the
// forces of the lead's own sites; the full list visits every moving pair from both sides, so
// there are no atomics and no write conflicts.  All nine site pairs of two molecules are
// evaluated without branches: records carry zero charges (and zero offsets) past a
// molecule's sites, which add exact zeros.  Site-pair arithmetic is in R: FP64, or FP32 with
// the per-molecule-pair force, energy and virial sums added to FP64 accumulators.  The lead
// pair's minimum image and cut-off test are FP64 without contraction in both, so the set of
// interacting pairs is the reference's.  Energies of moving pairs are halved for the double
// visit (frozen pairs are visited once); NMC energy classes of moving partners (1: one
// moving lead type, its class 0), two of frozen ones.  The G lanes reduce their sums with a
// fixed butterfly and write per-lead partial columns for a deterministic reduction.
template <typename R, int G, bool VIR, bool ACC, bool STRIDED, int NMC>
__device__ __forceinline__ void pair_body(const DevSys &S, int bid, int nblk) {
    __shared__ double s_tab[MAX_LEAD_TYPES * NCLS * PTAB_W];
    __shared__ int s_q[256 / G][PAIR_SEG];
    __shared__ double s_e2[64];
    for (int k = threadIdx.x; k < MAX_LEAD_TYPES * NCLS * PTAB_W; k += blockDim.x) s_tab[k] = S.ptab[k];
    if constexpr (sizeof(R) == 8) exp2_table_load(s_e2);
    __syncthreads();
    const int tid = threadIdx.x;
    const int grp = tid / G;
    const int gslot = bid * (blockDim.x / G) + grp;
    const int sub = tid % G;
    const int gbase = (tid & 63) - sub;   // the group's first lane in the wave
    const bool active = gslot < S.n_lead;
    // the lead of this group: in lead order, or in the cell order of the last list build
    // (neighbouring groups then share most partners in the caches)
    const int gid = (active && S.pair_cell_order) ? S.cell_items[gslot] : gslot;
    int *queue = s_q[grp];
    constexpr bool F32 = sizeof(R) == 4;

    double f[3][3];
    double vir[6];
    double ev[NMC], ec[NMC];      // moving partner classes
    double evw[2], ecw[2], w[6];  // frozen partner classes, wall force per side
#pragma unroll
    for (int a = 0; a < 3; a++) f[a][0] = f[a][1] = f[a][2] = 0.0;
#pragma unroll
    for (int c = 0; c < 6; c++) { vir[c] = 0.0; w[c] = 0.0; }
#pragma unroll
    for (int c = 0; c < NMC; c++) { ev[c] = 0.0; ec[c] = 0.0; }
    evw[0] = evw[1] = ecw[0] = ecw[1] = 0.0;

    Coul<R, ACC> cl;
    cl.alfa = (R)S.alfa; cl.spii = (R)S.spii; cl.p = (R)S.p;
    cl.a1 = (R)S.a1; cl.a2 = (R)S.a2; cl.a3 = (R)S.a3; cl.a4 = (R)S.a4; cl.a5 = (R)S.a5;
    cl.e2t = s_e2;

    int i = 0, nsi_i = 0;
    bool wall_any = false;   // frozen partners or surface-site bookkeeping for this lead
    if (active) {
        i = S.lead[gid];
        const int mi = S.pmeta[i];
        const int slot_i = mi & 15;
        nsi_i = mi >> 4;
        const double4 *prec4 = reinterpret_cast<const double4 *>(S.prec);
        const double4 m0 = prec4[(size_t)i * 3], m1 = prec4[(size_t)i * 3 + 1], m2 = prec4[(size_t)i * 3 + 2];
        const double xi = m0.x, yi = m0.y, zi = m0.z;
        // the lead's sites: offsets from the lead and facq q (zero past nsi_i)
        const R rxi[3] = {R(0), (R)m1.x, (R)m2.x}, ryi[3] = {R(0), (R)m1.y, (R)m2.y}, rzi[3] = {R(0), (R)m1.z, (R)m2.z};
        const R qi[3] = {(R)(S.facq * m0.w), (R)(S.facq * m1.w), (R)(S.facq * m2.w)};
        const double *tab_i = s_tab + slot_i * NCLS * PTAB_W;

        // ---- moving partners (loop A) ----
        auto mv_pos = [&](int j) {
            const double4 h = prec4[(size_t)j * 3];
            return make_double3(h.x, h.y, h.z);
        };
        auto mv_load = [&](int j) {
            MvRec<R> m;
            m.ok = j >= 0;
            const int jj = m.ok ? j : i;
            const double4 a = prec4[(size_t)jj * 3], b = prec4[(size_t)jj * 3 + 1], c = prec4[(size_t)jj * 3 + 2];
            m.x = a.x; m.y = a.y; m.z = a.z;
            m.rx[0] = R(0); m.ry[0] = R(0); m.rz[0] = R(0); m.q[0] = (R)a.w;
            m.rx[1] = (R)b.x; m.ry[1] = (R)b.y; m.rz[1] = (R)b.z; m.q[1] = (R)b.w;
            m.rx[2] = (R)c.x; m.ry[2] = (R)c.y; m.rz[2] = (R)c.z; m.q[2] = (R)c.w;
            m.cls = NMC > 1 ? (S.pmeta[jj] & 15) : 0;
            return m;
        };
        auto mv_eval = [&](const MvRec<R> &pj) {
            if (!pj.ok) return;
            double dx, dy, dz;
            const double rsq = lead_d2(S, xi, yi, zi, pj.x, pj.y, pj.z, dx, dy, dz);
            if (STRIDED && rsq > S.rcutsq) return;
            const double *pt = tab_i + pj.cls * PTAB_W;
            R evd, fvd;
            vdw_eval<R>((int)pt[0], (R)rsq, (R)pt[1], (R)pt[2], (R)pt[3], (R)pt[4], (R)pt[5], s_e2, evd, fvd);
            const R dxr = (R)dx, dyr = (R)dy, dzr = (R)dz;
            R e_c = R(0);
            R fa[3][3];      // this molecule pair's force on the lead's sites
            R vr[6];
#pragma unroll
            for (int c = 0; c < 6; c++) vr[c] = R(0);
#pragma unroll
            for (int a = 0; a < 3; a++) {
                fa[a][0] = fa[a][1] = fa[a][2] = R(0);
                const R xa = dxr + rxi[a], ya = dyr + ryi[a], za = dzr + rzi[a];
#pragma unroll
                for (int b = 0; b < 3; b++) {
                    R xd = dxr, yd = dyr, zd = dzr, r2 = (R)rsq;
                    if (a != 0 || b != 0) {   // (0, 0) is the lead pair, at the cut-off test's rsq
                        xd = xa - pj.rx[b]; yd = ya - pj.ry[b]; zd = za - pj.rz[b];
                        r2 = xd * xd + yd * yd + zd * zd;
                    }
                    R es, fs;
                    cl.eval(r2, qi[a] * pj.q[b], es, fs);
                    e_c += es;
                    const R fd = (a == 0 && b == 0) ? fvd + fs : fs;
                    if constexpr (F32) {
                        fa[a][0] += xd * fd; fa[a][1] += yd * fd; fa[a][2] += zd * fd;
                    } else {
                        f[a][0] += xd * fd; f[a][1] += yd * fd; f[a][2] += zd * fd;
                    }
                    if (VIR) {
                        vr[0] += xd * xd * fd; vr[1] += xd * yd * fd; vr[2] += xd * zd * fd;
                        vr[3] += yd * yd * fd; vr[4] += yd * zd * fd; vr[5] += zd * zd * fd;
                    }
                }
            }
            if constexpr (F32) {
#pragma unroll
                for (int a = 0; a < 3; a++) {
                    f[a][0] += (double)fa[a][0]; f[a][1] += (double)fa[a][1]; f[a][2] += (double)fa[a][2];
                }
            }
            if (VIR) {
#pragma unroll
                for (int c = 0; c < 6; c++) vir[c] += 0.5 * (double)vr[c];
            }
            if constexpr (NMC == 1) {
                ev[0] += 0.5 * (double)evd;
                ec[0] += 0.5 * (double)e_c;
            } else {
                acc_at<NMC>(ev, pj.cls, 0.5 * (double)evd);
                acc_at<NMC>(ec, pj.cls, 0.5 * (double)e_c);
            }
        };
        row_pass<G, STRIDED>(S, S.nbr_mm + (size_t)gid * S.cap_mm, min(S.cnt_mm[gid], S.cap_mm), queue, sub, gbase,
                             xi, yi, zi, mv_pos, mv_load, mv_eval);

        // surface-site bookkeeping between the two loops (src/forcem.f90:408-413)
        if (i >= S.n_liquid_sites) {
            if (signbit(zi)) w[2] += f[0][2]; else w[5] += f[0][2];
        }
        const int nw = S.n_wall > 0 ? min(S.cnt_mw[gid], S.cap_mw) : 0;
        wall_any = nw > 0 || i >= S.n_liquid_sites;

        // ---- frozen partners (loop B): single sites (x, y, z, q), crystal class 0 / 1 ----
        const bool wall_bk = S.itype[i] <= S.n_ltype;
        auto wz_pos = [&](int j) {
            const double4 h = S.wrec[j - S.n_move];
            return make_double3(h.x, h.y, h.z);
        };
        auto wz_loadp = [&](int j) {
            WzRec r;
            r.ok = j >= 0;
            const int wj = (r.ok ? j : S.n_move) - S.n_move;
            r.p = S.wrec[wj];
            r.cw = S.wcls[wj];
            return r;
        };
        auto wz_eval = [&](const WzRec &pr) {
            if (!pr.ok) return;
            const double4 pw = pr.p;
            double dx, dy, dz;
            const double rsq = lead_d2(S, xi, yi, zi, pw.x, pw.y, pw.z, dx, dy, dz);
            if (STRIDED && rsq > S.rcutsq) return;
            const int sgo = signbit(pw.z) ? 0 : 3;
            const int cw = pr.cw;   // classes MAX_LEAD_TYPES + cw
            const double *pt = tab_i + (MAX_LEAD_TYPES + cw) * PTAB_W;
            R evd, fvd;
            vdw_eval<R>((int)pt[0], (R)rsq, (R)pt[1], (R)pt[2], (R)pt[3], (R)pt[4], (R)pt[5], s_e2, evd, fvd);
            const R qw = (R)pw.w;
            const R dxr = (R)dx, dyr = (R)dy, dzr = (R)dz;
            R e_c = R(0);
            R wx = R(0), wy = R(0), wz = R(0);
            R vr[6];
#pragma unroll
            for (int c = 0; c < 6; c++) vr[c] = R(0);
#pragma unroll
            for (int a = 0; a < 3; a++) {
                R xd = dxr, yd = dyr, zd = dzr, r2 = (R)rsq;
                if (a != 0) {
                    xd = dxr + rxi[a]; yd = dyr + ryi[a]; zd = dzr + rzi[a];
                    r2 = xd * xd + yd * yd + zd * zd;
                }
                R es, fs;
                cl.eval(r2, qw * qi[a], es, fs);
                e_c += es;
                const R fij = (a == 0) ? fvd + fs : fs;
                const R fx = xd * fij, fy = yd * fij, fz = zd * fij;
                f[a][0] += (double)fx; f[a][1] += (double)fy; f[a][2] += (double)fz;
                wx -= fx; wy -= fy; wz -= fz;
                if (VIR) {
                    vr[0] += xd * fx; vr[1] += xd * fy; vr[2] += xd * fz;
                    vr[3] += yd * fy; vr[4] += yd * fz; vr[5] += zd * fz;
                }
            }
            if (VIR) {
#pragma unroll
                for (int c = 0; c < 6; c++) vir[c] += (double)vr[c];
            }
            if (wall_bk) {   // selects, not an index by side: w stays in registers
                const bool lo = sgo == 0;
                w[0] += lo ? (double)wx : 0.0; w[1] += lo ? (double)wy : 0.0; w[2] += lo ? (double)wz : 0.0;
                w[3] += lo ? 0.0 : (double)wx; w[4] += lo ? 0.0 : (double)wy; w[5] += lo ? 0.0 : (double)wz;
            }
            acc_at<2>(evw, cw, (double)evd);
            acc_at<2>(ecw, cw, (double)e_c);
        };
        if (S.n_wall > 0)
            row_pass<G, STRIDED>(S, S.nbr_mw + (size_t)gid * S.cap_mw, nw, queue, sub, gbase, xi, yi, zi, wz_pos,
                                 wz_loadp, wz_eval);
    }
    // ---- reduce over the G lanes of the group (fixed tree); the frozen-partner sums only
    // where a lead of the wave had frozen partners or is a surface site ----
#pragma unroll
    for (int a = 0; a < 3; a++) {
        f[a][0] = group_sum<G>(f[a][0]); f[a][1] = group_sum<G>(f[a][1]); f[a][2] = group_sum<G>(f[a][2]);
    }
    if (VIR) {
#pragma unroll
        for (int c = 0; c < 6; c++) vir[c] = group_sum<G>(vir[c]);
    }
#pragma unroll
    for (int c = 0; c < NMC; c++) { ev[c] = group_sum<G>(ev[c]); ec[c] = group_sum<G>(ec[c]); }
    if (__ballot(wall_any)) {
#pragma unroll
        for (int c = 0; c < 2; c++) { evw[c] = group_sum<G>(evw[c]); ecw[c] = group_sum<G>(ecw[c]); }
#pragma unroll
        for (int c = 0; c < 6; c++) w[c] = group_sum<G>(w[c]);
    }
    if (active && sub == 0) {
        for (int a = 0; a < nsi_i && a < 3; a++) {
            S.fx[i + a] = f[a][0]; S.fy[i + a] = f[a][1]; S.fz[i + a] = f[a][2];
        }
        const size_t L = (size_t)S.n_lead;
        double *P = S.part;
#pragma unroll
        for (int c = 0; c < MAX_LEAD_TYPES; c++) {
            P[(COL_EV + c) * L + gid] = c < NMC ? ev[c < NMC ? c : 0] : 0.0;
            P[(COL_EC + c) * L + gid] = c < NMC ? ec[c < NMC ? c : 0] : 0.0;
        }
#pragma unroll
        for (int c = 0; c < 2; c++) {
            P[(COL_EV + MAX_LEAD_TYPES + c) * L + gid] = evw[c];
            P[(COL_EC + MAX_LEAD_TYPES + c) * L + gid] = ecw[c];
        }
#pragma unroll
        for (int c = 0; c < 6; c++) P[(COL_FWS + c) * L + gid] = w[c];
        const double v9[9] = {vir[0], vir[1], vir[2], vir[1], vir[3], vir[4], vir[2], vir[4], vir[5]};
#pragma unroll
        for (int c = 0; c < 9; c++) P[(COL_VIR + c) * L + gid] = VIR ? v9[c] : 0.0;
    }
}

template <typename R, int G, bool VIR, bool ACC, bool STRIDED, int NMC>
__global__ __launch_bounds__(256) void k_pair(DevSys S) {
    pair_body<R, G, VIR, ACC, STRIDED, NMC>(S, blockIdx.x, gridDim.x);
}

#define MDF_PAIR_INST3(G, ST, NM)                                           \
    template __global__ void k_pair<double, G, true, false, ST, NM>(DevSys);  \
    template __global__ void k_pair<float, G, true, false, ST, NM>(DevSys);   \
    template __global__ void k_pair<double, G, false, false, ST, NM>(DevSys); \
    template __global__ void k_pair<float, G, false, false, ST, NM>(DevSys);  \
    template __global__ void k_pair<double, G, true, true, ST, NM>(DevSys);   \
    template __global__ void k_pair<float, G, true, true, ST, NM>(DevSys);    \
    template __global__ void k_pair<double, G, false, true, ST, NM>(DevSys);  \
    template __global__ void k_pair<float, G, false, true, ST, NM>(DevSys);
#define MDF_PAIR_INST(G) \
    MDF_PAIR_INST3(G, false, 1) MDF_PAIR_INST3(G, true, 1) MDF_PAIR_INST3(G, false, MAX_LEAD_TYPES) MDF_PAIR_INST3(G, true, MAX_LEAD_TYPES)
MDF_PAIR_INST(8)
MDF_PAIR_INST(16)
MDF_PAIR_INST(32)
#undef MDF_PAIR_INST3
#undef MDF_PAIR_INST

// forces of moving sites that are neither leads nor lead-carried are zero before the
// reciprocal part: clear every moving-site force the pair kernel does not write
__global__ void k_zero_forces(DevSys S) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < S.n_move) { S.fx[t] = 0.0; S.fy[t] = 0.0; S.fz[t] = 0.0; }
}

}  // namespace mdf

namespace mdf {
// batched form (one lanes-per-lead choice, loop form, erfc form and class count for the
// whole batch)
template <typename R, int G, bool VIR, bool ACC, bool STRIDED, int NMC>
__global__ __launch_bounds__(256) void k_pair_b(const RepArgs *__restrict__ RA, const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    pair_body<R, G, VIR, ACC, STRIDED, NMC>(RA[m.x].S, m.y, 0);
}
#define MDF_PAIR_B_INST3(G, ST, NM)                                                                        \
    template __global__ void k_pair_b<double, G, true, false, ST, NM>(const RepArgs *, const int2 *);  \
    template __global__ void k_pair_b<double, G, false, false, ST, NM>(const RepArgs *, const int2 *); \
    template __global__ void k_pair_b<float, G, true, false, ST, NM>(const RepArgs *, const int2 *);   \
    template __global__ void k_pair_b<float, G, false, false, ST, NM>(const RepArgs *, const int2 *);  \
    template __global__ void k_pair_b<double, G, true, true, ST, NM>(const RepArgs *, const int2 *);   \
    template __global__ void k_pair_b<double, G, false, true, ST, NM>(const RepArgs *, const int2 *);  \
    template __global__ void k_pair_b<float, G, true, true, ST, NM>(const RepArgs *, const int2 *);    \
    template __global__ void k_pair_b<float, G, false, true, ST, NM>(const RepArgs *, const int2 *);
#define MDF_PAIR_B_INST(G)                                                                   \
    MDF_PAIR_B_INST3(G, false, 1) MDF_PAIR_B_INST3(G, true, 1) MDF_PAIR_B_INST3(G, false, MAX_LEAD_TYPES) \
        MDF_PAIR_B_INST3(G, true, MAX_LEAD_TYPES)
MDF_PAIR_B_INST(8)
MDF_PAIR_B_INST(16)
MDF_PAIR_B_INST(32)
#undef MDF_PAIR_B_INST3
#undef MDF_PAIR_B_INST
}  // namespace mdf
