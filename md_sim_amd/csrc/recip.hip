// Reciprocal-space Ewald sum over the reference's half-space k-set (src/forcem.f90:827-918).
//
// S(k) = sum_i q_i exp(i k.r_i) over moving sites plus the static frozen-site S_w(k);
// E = 1/2 sum_k A_k |S(k)|^2;  F_i = sum_k A_k Im(conj(S(k)) q_i exp(i k.r_i)) k.
// The k-set is stored as rows (kl, km, kn0..kn1): kn is contiguous inside a row because the
// selection rule is a sphere in (kl, km, kn*kmax/knmax) (src/forcem.f90:756-759).
//
// Phase 1 (k_sf): each thread owns NS sites and a chunk of up to KC consecutive kn of one
// row; the phase exp(i pi t) with t = kl u + km v + kn w (u = x * 2pi/Lx / pi, ...) is
// started with one sincospi per (site, chunk) and advanced along kn by recursion in
// registers; per-thread partial sums over its sites are reduced over the block in LDS and
// written per site block - no atomics, fixed order.
// Phase 2 (k_bp): one thread per site walks whole rows (one sincospi per row), and the row's
// k data A_k conj(S(k)) are the same for the whole wave, so they come through scalar loads.
// FP64 advances with the Chebyshev recursion e_{n+1} = 2 cos(theta) e_n - e_{n-1}
// (2 FMA per step); FP32 uses the complex product e_{n+1} = e_n exp(i theta), whose
// rounding error grows linearly rather than quadratically along the row.
#include "mdf_internal.h"

namespace mdf {

template <typename R>
__device__ __forceinline__ void sincospi_r(double t, R &s, R &c);
template <>
__device__ __forceinline__ void sincospi_r<double>(double t, double &s, double &c) { sincospi(t, &s, &c); }
template <>
__device__ __forceinline__ void sincospi_r<float>(double t, float &s, float &c) {
    // reduce in FP64 to [-1, 1] half turns, then the FP32 function
    double r = t - 2.0 * rint(0.5 * t);
    sincospif((float)r, &s, &c);
}

// Recursion along kn: FP64 uses the Chebyshev form (state e_n, e_{n-1}); FP32 the
// complex product with exp(i pi w) (state e_n only).  cs = cos(pi w), sn = sin(pi w); dir is
// +1 going up in kn, -1 going down.
template <typename R>
struct KnWalk {
    R er, ei, pr, pi_;
    __device__ __forceinline__ void start(R Pr, R Pi, R cs, R sn, int dir) {
        // dir +1: e_0 = P (kn = 0), e_-1 = P conj(step); dir -1: e_0 = P conj(step) (kn = -1), prev = P
        R qr = Pr * cs + Pi * sn, qi = Pi * cs - Pr * sn;
        if (dir > 0) { er = Pr; ei = Pi; pr = qr; pi_ = qi; }
        else { er = qr; ei = qi; pr = Pr; pi_ = Pi; }
    }
    __device__ __forceinline__ void step(R cs, R sn, int dir) {
        if constexpr (sizeof(R) == 8) {
            R c2 = R(2) * cs;
            R nr = c2 * er - pr, ni = c2 * ei - pi_;
            pr = er; pi_ = ei; er = nr; ei = ni;
        } else {
            R s = dir > 0 ? sn : -sn;
            R nr = er * cs - ei * s, ni = er * s + ei * cs;
            er = nr; ei = ni;
        }
    }
};

// Row phase P = q exp(i pi (kl u + km v)): one sincospi at a block's first row and when kl
// changes, otherwise one complex product with exp(i pi v) per km step.
template <typename R>
__device__ __forceinline__ void row_phase(bool fresh, int kl, int km, double u, double v, R q, R vc, R vs, R &Pr,
                                          R &Pi) {
    if (fresh) {
        R s0, c0;
        sincospi_r<R>((double)kl * u + (double)km * v, s0, c0);
        Pr = q * c0; Pi = q * s0;
    } else {
        R nr = Pr * vc - Pi * vs, ni = Pr * vs + Pi * vc;
        Pr = nr; Pi = ni;
    }
}

// the same, reloading the site's coordinates and charge for a fresh start (keeps them out of
// registers in kernels that hold several sites per thread)
// out of line on purpose: several inlined copies of sincospi would be interleaved by the
// scheduler and hold their temporaries at once; a fresh phase is rare (once per kl)
template <typename R>
__device__ __attribute__((noinline)) void fresh_phase(int kl, int km, const double *__restrict__ x,
                                                      const double *__restrict__ y, const double *__restrict__ q,
                                                      int i, bool ok, double cux, double cuy, R *P) {
    R s0, c0;
    const double u = ok ? x[i] * cux : 0.0, v = ok ? y[i] * cuy : 0.0;
    const R qq = ok ? (R)q[i] : R(0);
    sincospi_r<R>((double)kl * u + (double)km * v, s0, c0);
    P[0] = qq * c0;
    P[1] = qq * s0;
}

template <typename R>
__device__ __forceinline__ void row_phase_reload(bool fresh, int kl, int km, const double *__restrict__ x,
                                                 const double *__restrict__ y, const double *__restrict__ q, int i,
                                                 bool ok, double cux, double cuy, R vc, R vs, R &Pr, R &Pi) {
    if (fresh) {
        R P[2];
        fresh_phase<R>(kl, km, x, y, q, i, ok, cux, cuy, P);
        Pr = P[0]; Pi = P[1];
    } else {
        R nr = Pr * vc - Pi * vs, ni = Pr * vs + Pi * vc;
        Pr = nr; Pi = ni;
    }
}

constexpr int SF_KC = 8;   // kn per accumulation chunk

template <typename R, int NS>
__device__ __forceinline__ void sf_body(const double *__restrict__ x, const double *__restrict__ y,
                                                 const double *__restrict__ z, const double *__restrict__ q, int n,
                                                 double cux, double cuy, double cuz,
                                                 const KRow *__restrict__ rows, int n_rows, int rpb,
                                                 double2 *__restrict__ partial, int nk, int bx, int by, int nbx, int nby) {
    constexpr int NW = SF_BLOCK / 64;
    __shared__ double red[NW][2 * SF_KC][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int sb = bx;
    double su[NS], sv[NS];
    R sq[NS], wc[NS], ws[NS], vc[NS], vs[NS], Pr[NS], Pi[NS];
#pragma unroll
    for (int m = 0; m < NS; m++) {
        const int i = sb * SF_BLOCK * NS + m * SF_BLOCK + threadIdx.x;
        double w = 0.0;
        if (i < n) { su[m] = x[i] * cux; sv[m] = y[i] * cuy; w = z[i] * cuz; sq[m] = (R)q[i]; }
        else { su[m] = 0.0; sv[m] = 0.0; sq[m] = R(0); }
        sincospi_r<R>(w, ws[m], wc[m]);
        sincospi_r<R>(sv[m], vs[m], vc[m]);
    }
    const int r_begin = by * rpb, r_end = min(n_rows, r_begin + rpb);
    int pkl = -1, pkm = 0;
    for (int r = r_begin; r < r_end; r++) {
        const KRow row = rows[r];
        const bool fresh = (r == r_begin) || row.kl != pkl || row.km != pkm + 1;
        pkl = row.kl; pkm = row.km;
#pragma unroll
        for (int m = 0; m < NS; m++) row_phase<R>(fresh, row.kl, row.km, su[m], sv[m], sq[m], vc[m], vs[m], Pr[m], Pi[m]);
        // two walks from kn = 0: upward over [max(kn0,0), kn1], downward over [kn0, -1]
        for (int dir = 1; dir >= -1; dir -= 2) {
            int s0, s1;   // first and last kn of the walk, inclusive, in walk order
            if (dir > 0) { s0 = max(row.kn0, 0); s1 = row.kn1; }
            else { s0 = -1; s1 = row.kn0; }
            if ((dir > 0 && s0 > s1) || (dir < 0 && s0 < s1)) continue;
            KnWalk<R> st[NS];
#pragma unroll
            for (int m = 0; m < NS; m++) {
                st[m].start(Pr[m], Pi[m], wc[m], ws[m], dir);
                if (dir > 0 && s0 == 1) st[m].step(wc[m], ws[m], dir);   // the (0,0) row starts at kn = 1
            }
            const int total = dir > 0 ? s1 - s0 + 1 : s0 - s1 + 1;
            for (int c = 0; c < total; c += SF_KC) {
                const int len = min(SF_KC, total - c);
                R ar[SF_KC], ai[SF_KC];
#pragma unroll
                for (int k = 0; k < SF_KC; k++) { ar[k] = R(0); ai[k] = R(0); }
#pragma unroll
                for (int m = 0; m < NS; m++) {
#pragma unroll
                    for (int k = 0; k < SF_KC; k++) {
                        ar[k] += st[m].er; ai[k] += st[m].ei;
                        st[m].step(wc[m], ws[m], dir);
                    }
                }
                // block reduction through LDS: thread t sums value t/16 over lanes (t%16)*4..+3
                // of every wave, then a 16-lane shuffle tree
#pragma unroll
                for (int k = 0; k < SF_KC; k++) {
                    red[wave][2 * k][lane] = (double)ar[k];
                    red[wave][2 * k + 1][lane] = (double)ai[k];
                }
                __syncthreads();
                const int vidx = threadIdx.x >> 4, p = threadIdx.x & 15;
                double sum = 0.0;
#pragma unroll
                for (int wv = 0; wv < NW; wv++)
#pragma unroll
                    for (int l = 0; l < 4; l++) sum += red[wv][vidx][p * 4 + l];
                sum += __shfl_xor(sum, 1, 64);
                sum += __shfl_xor(sum, 2, 64);
                sum += __shfl_xor(sum, 4, 64);
                sum += __shfl_xor(sum, 8, 64);
                const int k = vidx >> 1;
                if (p == 0 && k < len) {
                    const int kn = dir > 0 ? s0 + c + k : s0 - c - k;
                    double *dst = reinterpret_cast<double *>(&partial[(size_t)sb * nk + row.koff + (kn - row.kn0)]);
                    dst[vidx & 1] = sum;
                }
                __syncthreads();
            }
        }
    }
}

// ---- S(k) on the matrix cores (FP32 form) ------------------------------------------------
// For a row (kl, km) and kn, S = sum_j P_j E_j with P_j = q_j exp(i pi (kl u_j + km v_j)) and
// E_j = exp(i pi kn w_j): over a block of sites that is a complex product of a (rows x sites)
// and a (sites x kn) matrix.  As real operands, A = [Pr Pi] per site pair of k entries and
// B = [Er -Ei] (real part) or [Ei Er] (imaginary part).  Each FP32 operand is split into two
// FP16 parts, v = hi + lo, and A B ~ A_hi B_hi + A_hi B_lo + A_lo B_hi (the dropped lo lo
// term is 2^-22 relative), accumulated in FP32 by v_mfma_f32_16x16x32_f16 over a block of
// sites and written as FP32 partial rows, which k_sf_reduce sums in FP64.  Operand
// layout of that instruction (one wave): lane l holds A[l%16][8(l/16)..+7], B[8(l/16)..+7][l%16]
// and D[4(l/16)+e][l%16]; 8 k entries are 4 complex values.
// A block takes up to 512 sites (sb_sites, a multiple of 32) and a row group of up to four
// 16-row tiles, one per wave: the row tiles are spread evenly over the gridDim.y groups.  Per
// 32-site stage all threads build the P and E tables (split, in LDS) by recursion in steps of
// 8 along km and kn, then each wave runs its tile over the kn tiles its rows reach.
typedef _Float16 h16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int SFM_SB = 512;                  // sites per block (one partial row)
constexpr int SFM_SUB = 32;                  // sites per LDS stage
constexpr int SFM_ROWS = 64;                 // rows per block: 4 waves x 16
constexpr int SFM_NT_MAX = 8;                // kn tiles of 16 the E table holds (knmax <= 63)
// tables are [site][kn] and [site][row]; strides = 4 mod 8 put the two site quads a half
// wave reads in one step 16 banks apart
constexpr int SFM_KP = SFM_NT_MAX * 16 + 4;
constexpr int SFM_RP = SFM_ROWS + 4;

// (re, im) -> packed FP16 pairs hi and lo with re + i im = hi + lo to about 21 bits: hi
// rounds toward zero (v_cvt_pkrtz), so re - hi is exact in FP32 and lo carries it
typedef __fp16 h16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void split_c(float re, float im, unsigned &hi, unsigned &lo) {
    const h16x2 h = __builtin_amdgcn_cvt_pkrtz(re, im);
    const h16x2 l = __builtin_amdgcn_cvt_pkrtz(re - (float)h[0], im - (float)h[1]);
    hi = __builtin_bit_cast(unsigned, h);
    lo = __builtin_bit_cast(unsigned, l);
}

// sin, cos of pi t for the FP16-split tables: reduce to [-1, 1] half turns in FP64, then the
// hardware v_sin_f32 / v_cos_f32 (arguments in revolutions, absolute error about 2^-21)
__device__ __forceinline__ void sincospi_hw(double t, float &s, float &c) {
    const float rev = (float)(0.5 * (t - 2.0 * rint(0.5 * t)));
    s = __builtin_amdgcn_sinf(rev);
    c = __builtin_amdgcn_cosf(rev);
}

__device__ __forceinline__ h16x8 frag4(unsigned a, unsigned b, unsigned c, unsigned d) {
    uint4 v = make_uint4(a, b, c, d);
    return __builtin_bit_cast(h16x8, v);
}
__device__ __forceinline__ unsigned neg_im(unsigned c) { return c ^ 0x80000000u; }          // (re, -im)
__device__ __forceinline__ unsigned swap_c(unsigned c) { return (c >> 16) | (c << 16); }    // (im, re)
// four packed (re, -im) pairs -> (im, re): swap the halves, then flip the new low half's
// sign.  Two VALU instructions per pair, written as volatile asm so that the compiler
// recomputes them where they are used instead of keeping a loop-invariant copy in registers.
__device__ __forceinline__ unsigned conj_swap1(unsigned c) {
    unsigned r;
    asm volatile("v_alignbit_b32 %0, %1, %1, 16\n\tv_xor_b32 %0, 0x8000, %0" : "=&v"(r) : "v"(c));
    return r;
}
__device__ __forceinline__ h16x8 conj_swap8(h16x8 v) {
    uint4 u = __builtin_bit_cast(uint4, v);
    u.x = conj_swap1(u.x); u.y = conj_swap1(u.y);
    u.z = conj_swap1(u.z); u.w = conj_swap1(u.w);
    return __builtin_bit_cast(h16x8, u);
}

__device__ __forceinline__ f32x4 mfma16h(h16x8 a, h16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ int wave_min(int v) {
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ int wave_max(int v) {
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}

// block (bx, by) of a (site blocks) x (nby row groups) grid
template <int NT>
__device__ __forceinline__ void sfm_body(const double *__restrict__ x, const double *__restrict__ y,
                                         const double *__restrict__ z, const double *__restrict__ q, int n,
                                         double cux, double cuy, double cuz, const KRow *__restrict__ rows,
                                         int n_rows, int knmax, int n_nt, float2 *__restrict__ partial,
                                         int nk, int sb_sites, const MtsState *__restrict__ mts, int bx, int by,
                                         int nby) {
    if (mts && !mts->exact) return;   // a surrogate step: no k-sum
    // dynamic LDS, laid out as sfm_lds_bytes(sb_sites, n_nt) counts it
    extern __shared__ __align__(16) unsigned char sfm_smem[];
    const int kp = 16 * n_nt + 4;   // E row stride: = 4 mod 8 like the P stride
    float4 *sStep = reinterpret_cast<float4 *>(sfm_smem);              // cos, sin of 8 pi v and of 8 pi w
    double *sU = reinterpret_cast<double *>(sfm_smem + 16 * sb_sites);   // site phases in half turns
    double *sV = sU + sb_sites, *sW = sV + sb_sites;
    float *sQ = reinterpret_cast<float *>(sW + sb_sites);
    unsigned *sEh = reinterpret_cast<unsigned *>(sQ + sb_sites);
    unsigned *sEl = sEh + SFM_SUB * kp;
    unsigned *sPh = sEl + SFM_SUB * kp;
    unsigned *sPl = sPh + SFM_SUB * SFM_RP;
    __shared__ int sKl[SFM_ROWS], sKm[SFM_ROWS], sFresh[SFM_ROWS];
    const int sb = bx;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // the group's row tiles [t0, t1) of n_mt, at most SFM_ROWS / 16 of them
    const int n_mt = (n_rows + 15) >> 4;
    const int t0 = (int)(((long long)by * n_mt) / nby);
    const int t1 = (int)(((long long)(by + 1) * n_mt) / nby);
    const int row0 = t0 * 16;
    const int nr_g = min(16 * (t1 - t0), n_rows - row0);   // rows of the group
    const int wrow0 = row0 + wave * 16;
    const bool active = wave < t1 - t0;   // uniform over the wave
    const int s_begin = sb * sb_sites, s_end = min(n, s_begin + sb_sites);
    // prologue: the block's sites and rows, once
    for (int t = threadIdx.x; t < sb_sites; t += blockDim.x) {
        const int i = s_begin + t;
        const bool ok = i < s_end;
        const double u = ok ? x[i] * cux : 0.0, v = ok ? y[i] * cuy : 0.0, w = ok ? z[i] * cuz : 0.0;
        float vs, vc, ws, wc;
        sincospi_hw(8.0 * v, vs, vc);
        sincospi_hw(8.0 * w, ws, wc);
        sU[t] = u; sV[t] = v; sW[t] = w;
        sStep[t] = make_float4(vc, vs, wc, ws);
        sQ[t] = ok ? (float)q[i] : 0.f;
    }
    if (threadIdx.x < SFM_ROWS) {
        const int r = row0 + threadIdx.x;
        int kl = 0, km = 0, fresh = 1;
        if ((int)threadIdx.x < nr_g) {
            const KRow row = rows[r];
            kl = row.kl; km = row.km;
            if (threadIdx.x >= 8) {   // P walks rows in steps of 8: km + 8 in the same kl continues
                const KRow pr = rows[r - 8];
                fresh = (row.kl != pr.kl || row.km != pr.km + 8) ? 1 : 0;
            }
        }
        sKl[threadIdx.x] = kl; sKm[threadIdx.x] = km; sFresh[threadIdx.x] = fresh;
    }
    int lo = 1 << 30, hi = -(1 << 30);
    if (lane < 16 && wrow0 + lane < n_rows) {
        const KRow r = rows[wrow0 + lane];
        lo = r.kn0;
        hi = r.kn1;
    }
    lo = wave_min(lo);
    hi = wave_max(hi);
    const int nt_lo = active ? (lo + knmax) >> 4 : 1, nt_hi = active ? (hi + knmax) >> 4 : 0;
    f32x4 accr[NT], acci[NT];
#pragma unroll
    for (int t = 0; t < NT; t++) {
        accr[t] = f32x4{0.f, 0.f, 0.f, 0.f};
        acci[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    }
    __syncthreads();
    // table roles: site gj of the stage and phase gs of 8; a thread walks kn = gs + 8 e >= 0
    // (and writes E(-kn) = conj E(kn) beside it) and rows gs + 8 rr, so the 8 lanes of a site
    // write 8 consecutive LDS words
    const int gj = threadIdx.x >> 3, gs = threadIdx.x & 7;
    const int per_e = (knmax + 8) >> 3;   // kn = 0 .. knmax over 8 phases
    for (int c0 = 0; c0 < s_end - s_begin; c0 += SFM_SUB) {
        const int c = c0 + gj;
        const float4 st = sStep[c];
        {   // E = exp(i pi kn w) at table index knmax + kn, kn = gs + 8 e, and its conjugate at knmax - kn
            float es, ec;
            sincospi_hw((double)gs * sW[c], es, ec);
            for (int e = 0; e < per_e; e++) {
                const int kn = gs + 8 * e;
                if (kn <= knmax) {
                    unsigned h, l;
                    split_c(ec, es, h, l);
                    sEh[gj * kp + knmax + kn] = h;
                    sEl[gj * kp + knmax + kn] = l;
                    sEh[gj * kp + knmax - kn] = neg_im(h);
                    sEl[gj * kp + knmax - kn] = neg_im(l);
                }
                const float nr = ec * st.z - es * st.w, ni = ec * st.w + es * st.z;
                ec = nr; es = ni;
            }
        }
        {   // P = q exp(i pi (kl u + km v)) for the block's rows gs + 8 rr
            const float qq = sQ[c];
            const double u = sU[c], v = sV[c];
            float pr = 0.f, pi_ = 0.f;
#pragma unroll
            for (int rr = 0; rr < 8; rr++) {
                const int lr = gs + 8 * rr;
                if (8 * rr >= nr_g) break;   // uniform: no row of the group from here on
                if (sFresh[lr]) {
                    float s0, cc;
                    sincospi_hw((double)sKl[lr] * u + (double)sKm[lr] * v, s0, cc);
                    pr = cc; pi_ = s0;
                } else {
                    const float nr = pr * st.x - pi_ * st.y, ni = pr * st.y + pi_ * st.x;
                    pr = nr; pi_ = ni;
                }
                unsigned h, l;
                split_c(qq * pr, qq * pi_, h, l);
                sPh[gj * SFM_RP + lr] = h;
                sPl[gj * SFM_RP + lr] = l;
            }
        }
        __syncthreads();
        if (active) {
            // the matrix products at raised wave priority: they go first while the other
            // resident blocks build their tables
            __builtin_amdgcn_s_setprio(1);
            const int ar = wave * 16 + (lane & 15), g4 = 4 * (lane >> 4);
#pragma unroll 1
            for (int ks = 0; ks < SFM_SUB / 16; ks++) {
                const int sa = ks * 16 + g4;
                // real part: [Pr -Pi] . [Er Ei]; imaginary part: [Pi Pr] . [Er Ei]
                const unsigned *pph = sPh + sa * SFM_RP + ar, *ppl = sPl + sa * SFM_RP + ar;
                const unsigned ph0 = pph[0], ph1 = pph[SFM_RP], ph2 = pph[2 * SFM_RP], ph3 = pph[3 * SFM_RP];
                const unsigned pl0 = ppl[0], pl1 = ppl[SFM_RP], pl2 = ppl[2 * SFM_RP], pl3 = ppl[3 * SFM_RP];
                const h16x8 Arh = frag4(neg_im(ph0), neg_im(ph1), neg_im(ph2), neg_im(ph3));
                const h16x8 Arl = frag4(neg_im(pl0), neg_im(pl1), neg_im(pl2), neg_im(pl3));
                const h16x8 Aih = frag4(swap_c(ph0), swap_c(ph1), swap_c(ph2), swap_c(ph3));
                const h16x8 Ail = frag4(swap_c(pl0), swap_c(pl1), swap_c(pl2), swap_c(pl3));
#pragma unroll
                for (int t = 0; t < NT; t++) {
                    if (t < nt_lo || t > nt_hi) continue;
                    const int ec = t * 16 + (lane & 15);
                    const unsigned *peh = sEh + sa * kp + ec, *pel = sEl + sa * kp + ec;
                    const h16x8 Bh = frag4(peh[0], peh[kp], peh[2 * kp], peh[3 * kp]);
                    const h16x8 Bl = frag4(pel[0], pel[kp], pel[2 * kp], pel[3 * kp]);
                    accr[t] = mfma16h(Arh, Bh, accr[t]);
                    acci[t] = mfma16h(Aih, Bh, acci[t]);
                    accr[t] = mfma16h(Arh, Bl, accr[t]);
                    acci[t] = mfma16h(Aih, Bl, acci[t]);
                    accr[t] = mfma16h(Arl, Bh, accr[t]);
                    acci[t] = mfma16h(Ail, Bh, acci[t]);
                }
            }
            __builtin_amdgcn_s_setprio(0);
        }
        __syncthreads();
    }
    if (!active) return;
    // lane holds rows 4 (lane/16) + e of the tile at kn column lane % 16
#pragma unroll
    for (int t = 0; t < NT; t++) {
        if (t < nt_lo || t > nt_hi) continue;
        const int kn = t * 16 + (lane & 15) - knmax;
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int r = wrow0 + 4 * (lane >> 4) + e;
            if (r >= n_rows) continue;
            const KRow row = rows[r];
            if (kn < row.kn0 || kn > row.kn1) continue;
            partial[(size_t)sb * nk + row.koff + (kn - row.kn0)] = make_float2(accr[t][e], acci[t][e]);
        }
    }
}

template <int NT>
__global__ __launch_bounds__(256) void k_sf_mma(const double *__restrict__ x, const double *__restrict__ y,
                                                const double *__restrict__ z, const double *__restrict__ q, int n,
                                                double cux, double cuy, double cuz, const KRow *__restrict__ rows,
                                                int n_rows, int knmax, int n_nt, float2 *__restrict__ partial,
                                                int nk, int sb_sites, const MtsState *__restrict__ mts) {
    sfm_body<NT>(x, y, z, q, n, cux, cuy, cuz, rows, n_rows, knmax, n_nt, partial, nk, sb_sites, mts, blockIdx.x,
                 blockIdx.y, gridDim.y);
}
// batched: block map entry (system, local block), local block = row group * site blocks + site block
template <int NT>
__global__ __launch_bounds__(256) void k_sf_mma_b(const RepArgs *__restrict__ RA, const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    const RepArgs &r = RA[m.x];
    sfm_body<NT>(r.S.x, r.S.y, r.S.z, r.S.q, r.S.n_move, r.cux, r.cuy, r.cuz, r.rows, r.nrows, r.knmax, r.n_nt,
                 reinterpret_cast<float2 *>(r.sf_partial), r.nk, r.sfm_sites, nullptr, m.y % r.sfm_nsb,
                 m.y / r.sfm_nsb, r.sfm_nrg);
}

#define SFM_INST(NT)                                                                                            \
    template __global__ void k_sf_mma<NT>(const double *, const double *, const double *, const double *, int,   \
                                          double, double, double, const KRow *, int, int, int, float2 *, int, int, \
                                          const MtsState *); \
    template __global__ void k_sf_mma_b<NT>(const RepArgs *, const int2 *);
SFM_INST(1) SFM_INST(2) SFM_INST(3) SFM_INST(4) SFM_INST(5) SFM_INST(6) SFM_INST(7) SFM_INST(8)
#undef SFM_INST

// ---- back-projection on the matrix cores (FP32 form) ---------------------------------------
// F_j = q_j sum_rows c(row) Im(P_j(row) G_j(row)) with G_j(row) = sum_kn T(row, kn) E_j(kn),
// T = A conj S (and kn T for the z component): per 16-row tile, G for 16 sites is a complex
// product of T (rows x kn) and E (kn x sites) on the matrix cores (FP16 hi/lo split as in
// k_sf_mma, 12 products per K-step of 16 kn: G and Gz, real and imaginary).  k_sf_reduce_rows
// leaves T already split, in a dense zero-padded layout tpack[(row * 4 + plane) * kn_pad +
// kn + knmax] with planes T hi, T lo, kn T hi, kn T lo, each row's planes scaled by powers of
// two (tscale) that the contraction takes out again.  A wave owns 16 sites and keeps their
// E fragments in registers for the whole kernel; the block copies each T tile into LDS once
// for its waves, one tile ahead (two buffers, one barrier per tile).  The contraction with
// P runs per lane (a tile's four rows in FP32, the sum over tiles in FP64), and the four
// lanes that share a site add up at the end: every site's force is complete in one wave
// (one partial group).
constexpr int BPM_WAVES = 8;                 // most waves per block (blockDim.x / 64 may be fewer)
constexpr int BPM_ROWS_MAX = 1024;           // rows whose (kl, km) the block keeps in LDS

// block bx of the site blocks (16 sites per wave)
template <int NT>
__device__ __forceinline__ void bpm_body(const double *__restrict__ x, const double *__restrict__ y,
                                         const double *__restrict__ z, const double *__restrict__ q, int n,
                                         double cux, double cuy, double cuz, const int2 *__restrict__ rowkk,
                                         const int2 *__restrict__ tile_ks, int n_rows, int knmax, int n_nt,
                                         const unsigned *__restrict__ tpack, const int *__restrict__ tscale,
                                         double pipix, double pipiy, double pipiz, double *__restrict__ fpart,
                                         const MtsState *__restrict__ mts, int bx) {
    if (mts && !mts->exact) return;
    __shared__ unsigned sT[2][4][16][16 * NT + 4];   // row stride = 4 mod 8 words
    // per row: kl, km and the exponents of its two planes' scales, one signed byte each
    __shared__ int sInfo[BPM_ROWS_MAX];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, col = lane & 15;
    const int site = (bx * (int)(blockDim.x >> 6) + wave) * 16 + col;
    const bool ok = site < n;
    const int kn_pad = 16 * NT, q4 = kn_pad >> 2;   // uint4 per (row, plane); n_nt == NT
    const int n_mt = (n_rows + 15) >> 4;
    for (int r = threadIdx.x; r < n_rows; r += blockDim.x) {
        const int2 v = rowkk[r];
        sInfo[r] = (v.x & 0xff) | ((v.y & 0xff) << 8) | (tscale[r] << 16);
    }
    // stage tile mt (16 rows x 4 planes x kn_pad) into buffer b
    auto stage = [&](int mt, int b) {
        for (int e = threadIdx.x; e < 64 * q4; e += blockDim.x) {
            const int rp = e / q4, c4 = e - rp * q4;   // rp = row-in-tile * 4 + plane
            const int rr = rp >> 2, pl = rp & 3, r = mt * 16 + rr;
            uint4 v = make_uint4(0u, 0u, 0u, 0u);
            if (r < n_rows) v = *reinterpret_cast<const uint4 *>(&tpack[((size_t)r * 4 + pl) * kn_pad + 4 * c4]);
            *reinterpret_cast<uint4 *>(&sT[b][pl][rr][4 * c4]) = v;
        }
    };
    stage(0, 0);
    const double u = ok ? x[site] * cux : 0.0, v = ok ? y[site] * cuy : 0.0, w = ok ? z[site] * cuz : 0.0;
    const double qq = ok ? q[site] : 0.0;
    // E fragments for kn columns 16 ks + 4 g .. +3 (site col), in the two forms the complex
    // product needs: [Er -Ei] (real part) and [Ei Er] (imaginary part), hi and lo
    // [Ei Er] lo is derived from [Er -Ei] lo where it is used; at 6 and 7 kn tiles [Ei Er] hi
    // too, which brings them to 128 VGPRs (at 8 tiles that still leaves 3 waves per SIMD)
    constexpr bool IH_KEEP = NT <= 5 || NT >= 8;
    h16x8 Brh[NT], Brl[NT], Bih[IH_KEEP ? NT : 1];
    {
        float s1, c1, s16, c16, as, ac;
        sincospi_hw(w, s1, c1);
        sincospi_hw(16.0 * w, s16, c16);
        sincospi_hw((double)(4 * g - knmax) * w, as, ac);
#pragma unroll
        for (int ks = 0; ks < NT; ks++) {
            unsigned h[4], l[4];
            float er = ac, ei = as;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                split_c(er, ei, h[j], l[j]);
                const float t0 = er * c1 - ei * s1, t1 = er * s1 + ei * c1;
                er = t0; ei = t1;
            }
            Brh[ks] = frag4(neg_im(h[0]), neg_im(h[1]), neg_im(h[2]), neg_im(h[3]));
            Brl[ks] = frag4(neg_im(l[0]), neg_im(l[1]), neg_im(l[2]), neg_im(l[3]));
            if constexpr (IH_KEEP) Bih[ks] = frag4(swap_c(h[0]), swap_c(h[1]), swap_c(h[2]), swap_c(h[3]));
            const float t0 = ac * c16 - as * s16, t1 = ac * s16 + as * c16;
            ac = t0; as = t1;
        }
    }
    float vs1, vc1, vs16, vc16;
    sincospi_hw(v, vs1, vc1);
    sincospi_hw(16.0 * v, vs16, vc16);
    double fx = 0.0, fy = 0.0, fz = 0.0;
    float p0r = 0.f, p0i = 0.f;   // P of the lane's first row in the previous tile
    int p0kl = -1 << 20, p0km = 0;
    for (int mt = 0; mt < n_mt; mt++) {
        __syncthreads();   // tile mt staged; the other buffer's readers are done
        if (mt + 1 < n_mt) stage(mt + 1, (mt + 1) & 1);
        const int b = mt & 1;
        const int2 kr = tile_ks[mt];
        f32x4 gr = {0.f, 0.f, 0.f, 0.f}, gi = gr, zr = gr, zi = gr;
        __builtin_amdgcn_s_setprio(1);   // the matrix products first, as in k_sf_mma
#pragma unroll
        for (int ks = 0; ks < NT; ks++) {
            if (ks < kr.x || ks > kr.y) continue;
            const int kc = 16 * ks + 4 * g;
            const h16x8 Th = *reinterpret_cast<const h16x8 *>(&sT[b][0][col][kc]);
            const h16x8 Tl = *reinterpret_cast<const h16x8 *>(&sT[b][1][col][kc]);
            const h16x8 Zh = *reinterpret_cast<const h16x8 *>(&sT[b][2][col][kc]);
            const h16x8 Zl = *reinterpret_cast<const h16x8 *>(&sT[b][3][col][kc]);
            // real part: [Tr Ti] . [Er -Ei]; imaginary part: [Tr Ti] . [Ei Er]
            gr = mfma16h(Th, Brh[ks], gr);
            h16x8 Bhi;
            if constexpr (IH_KEEP) Bhi = Bih[ks]; else Bhi = conj_swap8(Brh[ks]);
            gi = mfma16h(Th, Bhi, gi);
            zr = mfma16h(Zh, Brh[ks], zr);
            zi = mfma16h(Zh, Bhi, zi);
            gr = mfma16h(Th, Brl[ks], gr);
            const h16x8 Bil = conj_swap8(Brl[ks]);
            gi = mfma16h(Th, Bil, gi);
            zr = mfma16h(Zh, Brl[ks], zr);
            zi = mfma16h(Zh, Bil, zi);
            gr = mfma16h(Tl, Brh[ks], gr);
            gi = mfma16h(Tl, Bhi, gi);
            zr = mfma16h(Zl, Brh[ks], zr);
            zi = mfma16h(Zl, Bhi, zi);
        }
        __builtin_amdgcn_s_setprio(0);
        // lane holds G for rows 4 g + e of the tile at its site: contract with P(row, site);
        // P walks km by 1 inside the tile and by 16 from tile to tile while kl stays
        float pr = 0.f, pi_ = 0.f, tx = 0.f, ty = 0.f, tz = 0.f;
        int pkl = 0, pkm = 0;
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const int r = mt * 16 + 4 * g + e;
            if (r >= n_rows) break;
            const int inf = sInfo[r];
            const int kl = (signed char)inf, km = (signed char)(inf >> 8);
            // 2^-sT and 2^-sZ undo the row's scaling (exponent bits of a normal float)
            const float it = __int_as_float((127 - (signed char)(inf >> 16)) << 23);
            const float iz = __int_as_float((127 - (signed char)(inf >> 24)) << 23);
            if (e == 0) {
                if (kl == p0kl && km == p0km + 16) {
                    const float t0 = p0r * vc16 - p0i * vs16, t1 = p0r * vs16 + p0i * vc16;
                    pr = t0; pi_ = t1;
                } else {
                    float s0, c0;
                    sincospi_hw((double)kl * u + (double)km * v, s0, c0);
                    pr = c0; pi_ = s0;
                }
                p0r = pr; p0i = pi_; p0kl = kl; p0km = km;
            } else if (kl == pkl && km == pkm + 1) {
                const float t0 = pr * vc1 - pi_ * vs1, t1 = pr * vs1 + pi_ * vc1;
                pr = t0; pi_ = t1;
            } else {
                float s0, c0;
                sincospi_hw((double)kl * u + (double)km * v, s0, c0);
                pr = c0; pi_ = s0;
            }
            pkl = kl; pkm = km;
            const float im = (pr * gi[e] + pi_ * gr[e]) * it;
            tx += im * (float)kl;
            ty += im * (float)km;
            tz += (pr * zi[e] + pi_ * zr[e]) * iz;
        }
        fx += (double)tx;   // the tile's four rows in FP32, the sum over tiles in FP64
        fy += (double)ty;
        fz += (double)tz;
    }
    fx *= pipix; fy *= pipiy; fz *= pipiz;
    // the four lanes of a site (g = 0..3) hold its row groups
    fx += __shfl_xor(fx, 16, 64); fx += __shfl_xor(fx, 32, 64);
    fy += __shfl_xor(fy, 16, 64); fy += __shfl_xor(fy, 32, 64);
    fz += __shfl_xor(fz, 16, 64); fz += __shfl_xor(fz, 32, 64);
    if (g == 0 && ok) {
        fpart[site] = qq * fx;
        fpart[(size_t)n + site] = qq * fy;
        fpart[2 * (size_t)n + site] = qq * fz;
    }
}

template <int NT>
__global__ __launch_bounds__(64 * BPM_WAVES) __attribute__((amdgpu_waves_per_eu(NT <= 7 ? 4 : 1))) void k_bp_mma(
    const double *__restrict__ x, const double *__restrict__ y, const double *__restrict__ z,
    const double *__restrict__ q, int n, double cux, double cuy, double cuz, const int2 *__restrict__ rowkk,
    const int2 *__restrict__ tile_ks, int n_rows, int knmax, int n_nt, const unsigned *__restrict__ tpack,
    const int *__restrict__ tscale, double pipix, double pipiy, double pipiz, double *__restrict__ fpart,
    const MtsState *__restrict__ mts) {
    bpm_body<NT>(x, y, z, q, n, cux, cuy, cuz, rowkk, tile_ks, n_rows, knmax, n_nt, tpack, tscale, pipix, pipiy,
                 pipiz, fpart, mts, blockIdx.x);
}
// batched: block map entry (system, site block); every system of the launch has NT kn tiles
// and the launch's waves per block
template <int NT>
__global__ __launch_bounds__(64 * BPM_WAVES) __attribute__((amdgpu_waves_per_eu(NT <= 7 ? 4 : 1))) void k_bp_mma_b(
    const RepArgs *__restrict__ RA, const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    const RepArgs &r = RA[m.x];
    bpm_body<NT>(r.S.x, r.S.y, r.S.z, r.S.q, r.S.n_move, r.cux, r.cuy, r.cuz, r.rowkk, r.tile_ks, r.nrows, r.knmax,
                 r.n_nt, r.tpack, r.tscale, r.pipix, r.pipiy, r.pipiz, r.fpart, nullptr, m.y);
}

#define BPM_INST(NT)                                                                                     \
    template __global__ void k_bp_mma<NT>(const double *, const double *, const double *, const double *, int, \
                                          double, double, double, const int2 *, const int2 *, int, int, int,    \
                                          const unsigned *, const int *, double, double, double, double *,      \
                                          const MtsState *); \
    template __global__ void k_bp_mma_b<NT>(const RepArgs *, const int2 *);
BPM_INST(1) BPM_INST(2) BPM_INST(3) BPM_INST(4) BPM_INST(5) BPM_INST(6) BPM_INST(7) BPM_INST(8)
#undef BPM_INST

template <typename R, int NS>
__global__ __launch_bounds__(SF_BLOCK) void k_sf(const double *__restrict__ x, const double *__restrict__ y,
                                                 const double *__restrict__ z, const double *__restrict__ q, int n,
                                                 double cux, double cuy, double cuz,
                                                 const KRow *__restrict__ rows, int n_rows, int rpb,
                                                 double2 *__restrict__ partial, int nk,
                                                 const MtsState *__restrict__ mts) {
    if (mts && !mts->exact) return;   // a surrogate step: no k-sum
    sf_body<R, NS>(x, y, z, q, n, cux, cuy, cuz, rows, n_rows, rpb, partial, nk, blockIdx.x, blockIdx.y, gridDim.x, gridDim.y);
}

template __global__ void k_sf<double, 4>(const double *, const double *, const double *, const double *, int, double,
                                        double, double, const KRow *, int, int, double2 *, int, const MtsState *);
template __global__ void k_sf<float, 4>(const double *, const double *, const double *, const double *, int, double,
                                       double, double, const KRow *, int, int, double2 *, int, const MtsState *);

// S(k) = sum over site blocks + static frozen part; T(k) = A_k conj(S(k)); per-block
// partial energies and virial:  eK = A|S|^2, eLL = A|S_moving|^2, vir_ab = 1/2 eK (d_ab - g k_a k_b).
// A block of SFR_THREADS owns SFR_K k-vectors; each of its waves sums every SFR_THREADS /
// SFR_K-th site block (fixed order, many loads in flight), the first wave combines the
// slices and finishes the k-vectors.
__device__ __forceinline__ void sf_reduce_body(const double2 *__restrict__ partial, int n_sb, int nk,
                                                   const double2 *__restrict__ s_wall, const double *__restrict__ ak,
                                                   const int *__restrict__ kvec, double pipix, double pipiy,
                                                   double pipiz, double lamb, double2 *__restrict__ s_out,
                                                   double4 *__restrict__ t_out, double *__restrict__ epart, int bid, int nblk,
                                                   unsigned *__restrict__ tpack = nullptr, const int *__restrict__ krow = nullptr,
                                                   int kn_pad = 0, int knmax = 0,
                                                   const float2 *__restrict__ partial32 = nullptr) {
    constexpr int SL = SFR_THREADS / SFR_K;
    __shared__ double2 slice[SL][SFR_K];
    const int kk = threadIdx.x % SFR_K, sl = threadIdx.x / SFR_K;
    const int k = bid * SFR_K + kk;
    // the finishing wave's k data, fetched before the partial rows so the loads overlap
    double2 w = make_double2(0.0, 0.0);
    double a = 0.0;
    int3 kv = make_int3(0, 0, 0);
    int kr = 0;
    if (sl == 0 && k < nk) {
        w = s_wall[k];
        a = ak[k];
        kv = make_int3(kvec[3 * k], kvec[3 * k + 1], kvec[3 * k + 2]);
        if (tpack) kr = krow[k];
    }
    double sr = 0.0, si = 0.0;
    if (k < nk && partial32) {   // FP32 partial rows (the matrix-core structure factor)
#pragma unroll 4
        for (int b = sl; b < n_sb; b += SL) {
            const float2 v = partial32[(size_t)b * nk + k];
            sr += (double)v.x; si += (double)v.y;
        }
    } else if (k < nk) {
#pragma unroll 4
        for (int b = sl; b < n_sb; b += SL) {
            const double2 v = partial[(size_t)b * nk + k];
            sr += v.x; si += v.y;
        }
    }
    slice[sl][kk] = make_double2(sr, si);
    __syncthreads();
    if (sl != 0) return;
    double acc[RE_N];
#pragma unroll
    for (int c = 0; c < RE_N; c++) acc[c] = 0.0;
    if (k < nk) {
        sr = 0.0; si = 0.0;
#pragma unroll
        for (int j = 0; j < SL; j++) { sr += slice[j][kk].x; si += slice[j][kk].y; }
        const double qr = sr + w.x, qi = si + w.y;
        s_out[k] = make_double2(qr, qi);
        const double kn = (double)kv.z;
        t_out[k] = make_double4(a * qr, -a * qi, kn * (a * qr), kn * (-a * qi));
        if (tpack) {   // the same, split for k_bp_mma (planes T hi, T lo, kn T hi, kn T lo)
            const size_t base = (size_t)kr * 4 * kn_pad + (size_t)(kv.z + knmax);
            unsigned h, l;
            split_c((float)(a * qr), (float)(-a * qi), h, l);
            tpack[base] = h;
            tpack[base + kn_pad] = l;
            split_c((float)(kn * (a * qr)), (float)(kn * (-a * qi)), h, l);
            tpack[base + 2 * kn_pad] = h;
            tpack[base + 3 * kn_pad] = l;
        }
        const double ek = a * (qr * qr + qi * qi);
        acc[0] = ek;
        acc[1] = a * (sr * sr + si * si);
        const double kx = pipix * kv.x, ky = pipiy * kv.y, kz = pipiz * kv.z;
        const double g = 2.0 * (1.0 / (kx * kx + ky * ky + kz * kz) + lamb);
        const double h = 0.5 * ek;
        acc[2] = h * (1.0 - g * kx * kx);
        acc[3] = h * (1.0 - g * ky * ky);
        acc[4] = h * (1.0 - g * kz * kz);
        acc[5] = -h * g * kx * ky;
        acc[6] = -h * g * kx * kz;
        acc[7] = -h * g * ky * kz;
    }
    // fixed butterfly over the wave: the same additions in the same order on every run
#pragma unroll
    for (int c = 0; c < RE_N; c++)
        for (int o = 32; o > 0; o >>= 1) acc[c] += __shfl_xor(acc[c], o, 64);
    if (kk == 0)
#pragma unroll
        for (int c = 0; c < RE_N; c++) epart[bid * RE_N + c] = acc[c];
}

__global__ __launch_bounds__(SFR_THREADS) void k_sf_reduce(const double2 *__restrict__ partial, int n_sb, int nk,
                                                   const double2 *__restrict__ s_wall, const double *__restrict__ ak,
                                                   const int *__restrict__ kvec, double pipix, double pipiy,
                                                   double pipiz, double lamb, double2 *__restrict__ s_out,
                                                   double4 *__restrict__ t_out, double *__restrict__ epart,
                                                   const MtsState *__restrict__ mts, unsigned *__restrict__ tpack,
                                                   const int *__restrict__ krow, int kn_pad, int knmax,
                                                   const float2 *__restrict__ partial32) {
    if (mts && !mts->exact) return;   // k partials keep the last full sum's values
    sf_reduce_body(partial, n_sb, nk, s_wall, ak, kvec, pipix, pipiy, pipiz, lamb, s_out, t_out, epart, blockIdx.x,
                   gridDim.x, tpack, krow, kn_pad, knmax, partial32);
}

// The same reduction with one block per k-set row, for the matrix-core back-projection: it
// also splits T = A conj S and kn T into FP16 hi/lo pairs for k_bp_mma.  FP16 holds
// magnitudes below 65504 only, so each row's two planes are first scaled by exact powers of
// two, 2^sT and 2^sZ, that bring the row's largest |T| (|kn T|) into [2^14, 2^15): hi never
// saturates, and lo keeps 11 bits below hi down to 2^-29 of the row's largest entry, beneath
// which entries contribute less than FP32 rounding.  k_bp_mma multiplies the row's products
// by 2^-sT and 2^-sZ.  tscale[row] = (sT & 0xff) | (sZ & 0xff) << 8, each clamped to
// [-100, 100] (|T| <= A_k sum |q| stays many decades inside that range).
constexpr int SFRR_CHUNKS = 2;   // 64-wide chunks of a row: rows hold at most 2 * 63 + 1 k-vectors

__device__ __forceinline__ int pow2_scale(float m) {
    if (!(m > 0.f) || !(m < 3.0e38f)) return 0;   // zero or non-finite row: no scaling
    int ex;
    (void)frexpf(m, &ex);                         // m = f 2^ex, f in [0.5, 1)
    return min(100, max(-100, 15 - ex));          // m 2^(15 - ex) in [2^14, 2^15)
}

// row rb of the k-set
__device__ __forceinline__ void sfr_rows_body(
    const double2 *__restrict__ partial, const float2 *__restrict__ partial32, int n_sb, int nk,
    const KRow *__restrict__ rows, const double2 *__restrict__ s_wall, const double *__restrict__ ak,
    double pipix, double pipiy, double pipiz, double lamb, double2 *__restrict__ s_out, double4 *__restrict__ t_out,
    double *__restrict__ epart, const MtsState *__restrict__ mts, unsigned *__restrict__ tpack,
    int *__restrict__ tscale, int kn_pad, int knmax, int rb) {
    if (mts && !mts->exact) return;
    constexpr int SL = SFR_THREADS / SFR_K;
    __shared__ double2 slice[SL][SFR_K];
    const int kk = threadIdx.x % SFR_K, sl = threadIdx.x / SFR_K;
    const KRow row = rows[rb];
    const int len = row.kn1 - row.kn0 + 1;
    double acc[RE_N];
#pragma unroll
    for (int c = 0; c < RE_N; c++) acc[c] = 0.0;
    double tr[SFRR_CHUNKS], ti[SFRR_CHUNKS];
#pragma unroll
    for (int ch = 0; ch < SFRR_CHUNKS; ch++) {
        tr[ch] = 0.0; ti[ch] = 0.0;
        const int j = ch * SFR_K + kk;
        const bool ok = j < len;
        const int k = row.koff + j;
        double sr = 0.0, si = 0.0;
        if (ok && partial32) {
#pragma unroll 4
            for (int b = sl; b < n_sb; b += SL) {
                const float2 v = partial32[(size_t)b * nk + k];
                sr += (double)v.x; si += (double)v.y;
            }
        } else if (ok) {
#pragma unroll 4
            for (int b = sl; b < n_sb; b += SL) {
                const double2 v = partial[(size_t)b * nk + k];
                sr += v.x; si += v.y;
            }
        }
        slice[sl][kk] = make_double2(sr, si);
        __syncthreads();
        if (sl == 0 && ok) {
            sr = 0.0; si = 0.0;
#pragma unroll
            for (int q = 0; q < SL; q++) { sr += slice[q][kk].x; si += slice[q][kk].y; }
            const double2 w = s_wall[k];
            const double a = ak[k];
            const int kn = row.kn0 + j;
            const double qr = sr + w.x, qi = si + w.y;
            s_out[k] = make_double2(qr, qi);
            tr[ch] = a * qr; ti[ch] = -a * qi;
            t_out[k] = make_double4(tr[ch], ti[ch], kn * tr[ch], kn * ti[ch]);
            const double ek = a * (qr * qr + qi * qi);
            acc[0] += ek;
            acc[1] += a * (sr * sr + si * si);
            const double kx = pipix * row.kl, ky = pipiy * row.km, kz = pipiz * kn;
            const double g = 2.0 * (1.0 / (kx * kx + ky * ky + kz * kz) + lamb);
            const double h = 0.5 * ek;
            acc[2] += h * (1.0 - g * kx * kx);
            acc[3] += h * (1.0 - g * ky * ky);
            acc[4] += h * (1.0 - g * kz * kz);
            acc[5] += -h * g * kx * ky;
            acc[6] += -h * g * kx * kz;
            acc[7] += -h * g * ky * kz;
        }
        if ((ch + 1) * SFR_K >= len) break;   // uniform over the block
        __syncthreads();
    }
    if (sl != 0) return;
    // the row's largest |T| and |kn T| (max is exact and order-free: the same bits every run)
    float mt = 0.f, mz = 0.f;
#pragma unroll
    for (int ch = 0; ch < SFRR_CHUNKS; ch++) {
        const int j = ch * SFR_K + kk;
        if (j < len) {
            const float m = (float)fmax(fabs(tr[ch]), fabs(ti[ch]));
            mt = fmaxf(mt, m);
            mz = fmaxf(mz, m * (float)abs(row.kn0 + j));
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        mt = fmaxf(mt, __shfl_xor(mt, o, 64));
        mz = fmaxf(mz, __shfl_xor(mz, o, 64));
    }
    // the scaled maxima stay below 2^15 after FP32 rounding of the scaled doubles
    const int st = pow2_scale(mt), sz = pow2_scale(mz);
#pragma unroll
    for (int ch = 0; ch < SFRR_CHUNKS; ch++) {
        const int j = ch * SFR_K + kk;
        if (j < len) {
            const int kn = row.kn0 + j;
            const size_t base = (size_t)rb * 4 * kn_pad + (size_t)(kn + knmax);
            unsigned h, l;
            split_c((float)ldexp(tr[ch], st), (float)ldexp(ti[ch], st), h, l);
            tpack[base] = h;
            tpack[base + kn_pad] = l;
            split_c((float)ldexp(kn * tr[ch], sz), (float)ldexp(kn * ti[ch], sz), h, l);
            tpack[base + 2 * kn_pad] = h;
            tpack[base + 3 * kn_pad] = l;
        }
    }
    if (kk == 0) tscale[rb] = (st & 0xff) | ((sz & 0xff) << 8);
#pragma unroll
    for (int c = 0; c < RE_N; c++)
        for (int o = 32; o > 0; o >>= 1) acc[c] += __shfl_xor(acc[c], o, 64);
    if (kk == 0)
#pragma unroll
        for (int c = 0; c < RE_N; c++) epart[rb * RE_N + c] = acc[c];
}

__global__ __launch_bounds__(SFR_THREADS) void k_sf_reduce_rows(
    const double2 *__restrict__ partial, const float2 *__restrict__ partial32, int n_sb, int nk,
    const KRow *__restrict__ rows, const double2 *__restrict__ s_wall, const double *__restrict__ ak,
    double pipix, double pipiy, double pipiz, double lamb, double2 *__restrict__ s_out, double4 *__restrict__ t_out,
    double *__restrict__ epart, const MtsState *__restrict__ mts, unsigned *__restrict__ tpack,
    int *__restrict__ tscale, int kn_pad, int knmax) {
    sfr_rows_body(partial, partial32, n_sb, nk, rows, s_wall, ak, pipix, pipiy, pipiz, lamb, s_out, t_out, epart, mts,
                  tpack, tscale, kn_pad, knmax, blockIdx.x);
}
// batched: block map entry (system, row)
__global__ __launch_bounds__(SFR_THREADS) void k_sf_reduce_rows_b(const RepArgs *__restrict__ RA,
                                                                  const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    const RepArgs &r = RA[m.x];
    sfr_rows_body(r.sf_partial, r.S.n_move > 0 ? reinterpret_cast<const float2 *>(r.sf_partial) : nullptr,
                  r.S.n_move > 0 ? r.sfm_nsb : 0, r.nk, r.rows, r.s_wall, r.ak, r.pipix, r.pipiy, r.pipiz, r.lamb,
                  r.s_out, r.t_out, r.kpart, nullptr, r.tpack, r.tscale, 16 * r.n_nt, r.knmax, m.y);
}

// T4[k] = (A conj S, kn A conj S): per k-vector the force pieces are Im(T e) and kn Im(T e),
// each two FMAs on the separated real and imaginary products
template <typename R>
__device__ __forceinline__ void bp_acc(const double4 &t, R er, R ei, R &a1, R &a2, R &z1, R &z2) {
    a1 += (R)t.x * ei;
    a2 += (R)t.y * er;
    z1 += (R)t.z * ei;
    z2 += (R)t.w * er;
}

// Rows are split into groups of similar k count (host-balanced, row_start[g]..row_start[g+1]);
// a block stages its group's k data (at most BP_LDS_K entries) in LDS once and every wave
// reads it with uniform-address LDS loads, instead of scalar loads that miss the scalar cache.
constexpr int BP_LDS_K = 1024;

template <typename R, int NS>
__device__ __forceinline__ void bp_body(const double *__restrict__ x, const double *__restrict__ y,
                                            const double *__restrict__ z, const double *__restrict__ q, int n,
                                            double cux, double cuy, double cuz, const KRow *__restrict__ rows,
                                            const int *__restrict__ row_start, const double4 *__restrict__ T,
                                            double pipix, double pipiy, double pipiz, double *__restrict__ fpart, int bx, int by, int nbx, int nby) {
    // NS sites per thread: every k datum read from LDS serves NS sites
    __shared__ double4 sT[BP_LDS_K + 2];
    const int g = by;
    const int r_begin = row_start[g], r_end = row_start[g + 1];
    const int k_begin = rows[r_begin].koff;
    const int k_count = rows[r_end - 1].koff + (rows[r_end - 1].kn1 - rows[r_end - 1].kn0 + 1) - k_begin;
    for (int t = threadIdx.x; t < k_count + 2; t += blockDim.x)
        sT[t] = T[k_begin + t];   // T is padded past its last entry
    __syncthreads();
    int idx[NS];
    R wc[NS], ws[NS], vc[NS], vs[NS], Pr[NS], Pi[NS];
    double fx[NS], fy[NS], fz[NS];
#pragma unroll
    for (int m = 0; m < NS; m++) {
        idx[m] = bx * blockDim.x * NS + m * blockDim.x + threadIdx.x;
        const bool ok = idx[m] < n;
        const double w = ok ? z[idx[m]] * cuz : 0.0, v = ok ? y[idx[m]] * cuy : 0.0;
        sincospi_r<R>(w, ws[m], wc[m]);
        sincospi_r<R>(v, vs[m], vc[m]);
        Pr[m] = R(0); Pi[m] = R(0);
        fx[m] = 0.0; fy[m] = 0.0; fz[m] = 0.0;
    }
    int pkl = -1, pkm = 0;
    for (int r = r_begin; r < r_end; r++) {
        const KRow row = rows[r];
        const bool fresh = (r == r_begin) || row.kl != pkl || row.km != pkm + 1;
        pkl = row.kl; pkm = row.km;
        R acc[NS], accz[NS];
#pragma unroll
        for (int m = 0; m < NS; m++) {
            row_phase_reload<R>(fresh, row.kl, row.km, x, y, q, idx[m], idx[m] < n, cux, cuy, vc[m], vs[m], Pr[m],
                                Pi[m]);
            acc[m] = R(0); accz[m] = R(0);
        }
        for (int dir = 1; dir >= -1; dir -= 2) {
            int s0, s1;
            if (dir > 0) { s0 = max(row.kn0, 0); s1 = row.kn1; }
            else { s0 = -1; s1 = row.kn0; }
            if ((dir > 0 && s0 > s1) || (dir < 0 && s0 < s1)) continue;
            KnWalk<R> st[NS];
#pragma unroll
            for (int m = 0; m < NS; m++) {
                st[m].start(Pr[m], Pi[m], wc[m], ws[m], dir);
                if (dir > 0 && s0 == 1) st[m].step(wc[m], ws[m], dir);
            }
            const int total = dir > 0 ? s1 - s0 + 1 : s0 - s1 + 1;
            const double4 *Tr = sT + (row.koff - k_begin) + (s0 - row.kn0);
            for (int j = 0; j < total; j++) {
                const double4 tv = Tr[dir * j];
#pragma unroll
                for (int m = 0; m < NS; m++) {
                    acc[m] += (R)tv.x * st[m].ei;
                    acc[m] += (R)tv.y * st[m].er;
                    accz[m] += (R)tv.z * st[m].ei;
                    accz[m] += (R)tv.w * st[m].er;
                    st[m].step(wc[m], ws[m], dir);
                }
            }
        }
#pragma unroll
        for (int m = 0; m < NS; m++) {
            fx[m] += (double)acc[m] * (pipix * (double)row.kl);
            fy[m] += (double)acc[m] * (pipiy * (double)row.km);
            fz[m] += (double)accz[m] * pipiz;
        }
    }
#pragma unroll
    for (int m = 0; m < NS; m++) {
        if (idx[m] < n) {
            fpart[((size_t)g * 3 + 0) * n + idx[m]] = fx[m];
            fpart[((size_t)g * 3 + 1) * n + idx[m]] = fy[m];
            fpart[((size_t)g * 3 + 2) * n + idx[m]] = fz[m];
        }
    }
}

template <typename R, int NS>
__global__ __launch_bounds__(256) void k_bp(const double *__restrict__ x, const double *__restrict__ y,
                                            const double *__restrict__ z, const double *__restrict__ q, int n,
                                            double cux, double cuy, double cuz, const KRow *__restrict__ rows,
                                            const int *__restrict__ row_start, const double4 *__restrict__ T,
                                            double pipix, double pipiy, double pipiz, double *__restrict__ fpart,
                                            const MtsState *__restrict__ mts) {
    if (mts && !mts->exact) return;
    bp_body<R, NS>(x, y, z, q, n, cux, cuy, cuz, rows, row_start, T, pipix, pipiy, pipiz, fpart, blockIdx.x, blockIdx.y, gridDim.x, gridDim.y);
}

template __global__ void k_bp<double, 4>(const double *, const double *, const double *, const double *, int, double,
                                         double, double, const KRow *, const int *, const double4 *, double, double,
                                         double, double *, const MtsState *);
template __global__ void k_bp<float, 4>(const double *, const double *, const double *, const double *, int, double,
                                        double, double, const KRow *, const int *, const double4 *, double, double,
                                        double, double *, const MtsState *);

}  // namespace mdf

namespace mdf {
// batched forms
template <typename R>
__global__ __launch_bounds__(SF_BLOCK) void k_sf_b(const RepArgs *__restrict__ RA, const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    const RepArgs &r = RA[m.x];
    const int bx = m.y % r.n_sb, by = m.y / r.n_sb;
    sf_body<R, 4>(r.S.x, r.S.y, r.S.z, r.S.q, r.S.n_move, r.cux, r.cuy, r.cuz, r.rows, r.nrows, r.cpb, r.sf_partial,
                  r.nk, bx, by, r.n_sb, r.n_cg);
}
template __global__ void k_sf_b<double>(const RepArgs *, const int2 *);
template __global__ void k_sf_b<float>(const RepArgs *, const int2 *);

__global__ __launch_bounds__(SFR_THREADS) void k_sf_reduce_b(const RepArgs *__restrict__ RA, const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    const RepArgs &r = RA[m.x];
    sf_reduce_body(r.sf_partial, r.S.n_move > 0 ? r.n_sb : 0, r.nk, r.s_wall, r.ak, r.kvec, r.pipix, r.pipiy, r.pipiz,
                   r.lamb, r.s_out, r.t_out, r.kpart, m.y, r.n_kblk);
}

template <typename R>
__global__ __launch_bounds__(256) void k_bp_b(const RepArgs *__restrict__ RA, const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    const RepArgs &r = RA[m.x];
    const int bx = m.y % r.n_bpsb, by = m.y / r.n_bpsb;
    bp_body<R, 4>(r.S.x, r.S.y, r.S.z, r.S.q, r.S.n_move, r.cux, r.cuy, r.cuz, r.rows, r.bp_row_start, r.t_out,
                  r.pipix, r.pipiy, r.pipiz, r.fpart, bx, by, r.n_bpsb, r.n_rg);
}
template __global__ void k_bp_b<double>(const RepArgs *, const int2 *);
template __global__ void k_bp_b<float>(const RepArgs *, const int2 *);
}  // namespace mdf
