// Reciprocal-force surrogate with multiple time steps.
//
// No counterpart in the reference, which evaluates the full k-sum every step
// (src/forcem.f90:827-918).  Between full sums the k-space force on a site is predicted, and
// the full sum every `interval`-th step is both the correction and the judge:
//   F_k(t) ~ S(t) + delta(t0) + (t - t0) [delta(t0) - delta(t1)] / (t0 - t1),
// with S a per-site network of the site's surroundings, delta = F_k - S the correction at
// the last two full sums t0 > t1 (the linear term only when both came from the same
// weights: the part S misses is mostly the far field, which changes smoothly).  The judged
// error is that prediction against the next full sum, relative to the RMS of the whole site
// force; above the tolerance the scheme falls back to full sums every step (and trains
// again) before it resumes.
//
// Per moving site i and every site j within SG_RC (taper fc = (cos(pi r/SG_RC) + 1)/2,
// r_ij = r_i - r_j), inputs x_i (48; 39 used) and vector features U_i (16 x 3):
//   x[0:8], x[8:16]   sum over q_j > 0 / q_j < 0 of |q_j| fc exp(-2 (r - 0.75 (m+1))^2)
//   x[16:37]          dot products of the moments A_{c,s} = sum |q_j| fc exp(-(r - R_s)^2) r_ij/r
//                     (classes c as above, R_s = 1.5, 3, 4.5 A), a <= b in a-major order
//   x[37], x[38]      q_i, q_i^2
// The vector features run over whole molecules (the reference's group cut-off,
// src/forcem.f90:299,360-402, which keeps each neighbour molecule neutral): every site of a
// molecule whose lead lies within rc of i's lead (frozen sites: within rc of i's lead), with
// the smooth step wg(R) = 1 - s^2 (3 - 2 s), s = clamp((R - rc + 2)/2, 0, 1), of that lead
// distance, and the other sites of i's own molecule with weight 1:
//   U_0               facq q_i sum_j q_j wg (erf(a r)/r^2 - 2a/sqrt(pi) exp(-a^2 r^2)/r) r_ij/r
//   U_m (m = 1..7)    facq q_i sum_j q_j wg T_{m-1}(2 r/rc - 1) r_ij/r   (Chebyshev T)
//   U_8 .. U_15       0 (the network's 16 outputs fill one matrix-core tile)
// Network (bfloat16 weights, FP32 accumulation on the matrix cores): h1 = silu(W1 xn + b1),
// h2 = silu(W2 h1 + b2), c = W3 h2 + b3, S_i = sum_m c_m U_i,m / su_m, xn = (x - mu)/sigma.
// Training: Adam on the change between two full-sum frames a, b (what the scheme uses),
// L = sum_i |(S_a,i - S_b,i) - (F_a,i - F_b,i)|^2 / (n mean|F_a - F_b|^2).
//
// Matrix-core layout (v_mfma_f32_16x16x16_bf16, one wave): lane l holds A[l%16][4(l/16)..+3],
// B[4(l/16)..+3][l%16] and D[4(l/16)+r][l%16].  Activations run transposed (units x sites)
// with the weights as A, so a layer's output tile is the next layer's B operand as it
// stands.  Weight gradients contract over sites, so their operands pass through a small
// per-wave LDS transpose.
#include "mdf_internal.h"

namespace mdf {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short f2bf(float f) {   // round to nearest even (finite)
    unsigned u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ bf16x4 pack4(float a, float b, float c, float d) {
    bf16x4 r;
    r[0] = (short)f2bf(a); r[1] = (short)f2bf(b); r[2] = (short)f2bf(c); r[3] = (short)f2bf(d);
    return r;
}
__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ bf16x4 frag(const unsigned short *p) { return *reinterpret_cast<const bf16x4 *>(p); }
__device__ __forceinline__ float silu(float z) { return z / (1.0f + expf(-z)); }
__device__ __forceinline__ float dsilu(float z) {
    const float s = 1.0f / (1.0f + expf(-z));
    return s * (1.0f + z * (1.0f - s));
}
__device__ __forceinline__ bool overflowed(const DevSys &S) {
    return S.flags->overflow_mm != 0 || S.flags->overflow_mw != 0;
}
__device__ __forceinline__ f32x4 zero4() { f32x4 z = {0.f, 0.f, 0.f, 0.f}; return z; }

// ---- control: does this step run the full k-sum? --------------------------------------
__global__ void k_sg_control(MtsArgs a, int again) {
    if (threadIdx.x != 0 || again) return;   // a redone step keeps its decision
    MtsState &st = *a.st;
    st.step += 1;
    st.train_now = 0;   // only a full-sum step's judge may set it (training needs that step's target)
    if (st.state == 1) {
        st.phase += 1;
        st.exact = st.phase >= a.interval;
        if (st.exact) st.phase = 0;
    } else {
        st.exact = 1;
        st.phase = 0;
    }
    if (st.exact) st.n_exact++;
    else st.n_surrogate++;
}

// ---- inputs and vector features ----------------------------------------------------------
// per-site accumulators as one array: radial [0, 16), moments [16, 34), vectors [34, 58)
constexpr int SA_A = 16, SA_U = 34, SA_N = 34 + 3 * SG_MV;
struct SiteAcc {
    float v[SA_N];
};

// invariant inputs: one site j within SG_RC of site i.  The Gaussians of equally spaced
// centres follow from one exponential each by the ratio of neighbours,
// g_{m+1} / g_m = exp(2 eta dc r) exp(-eta (c_{m+1}^2 - c_m^2)), with eta dc = 1.5 for both
// families (radial: eta 2, centres 0.75 (m+1); angular: eta 1, centres 1.5 (s+1)).
__device__ __forceinline__ void sg_inputs(SiteAcc &s, float dx, float dy, float dz, float r2, float qj) {
    const float ri = __builtin_amdgcn_rsqf(r2), r = r2 * ri;
    const float fc = 0.5f * (__cosf(3.14159265358979f / SG_RC * r) + 1.0f);
    const float ux = dx * ri, uy = dy * ri, uz = dz * ri;
    const float aq = fabsf(qj) * fc;
    const float wp = qj > 0.0f ? aq : 0.0f, wn = qj < 0.0f ? aq : 0.0f;
    // exp(-2 (c_{m+1}^2 - c_m^2)), c_m = 0.75 (m + 1), and exp(-(c_{s+1}^2 - c_s^2)), c_s = 1.5 (s + 1)
    constexpr float KR[8] = {3.421811831e-02f, 3.606563136e-03f, 3.801289579e-04f, 4.006529739e-05f, 4.222851277e-06f, 4.450852500e-07f, 4.691164022e-08f, 4.944450502e-09f};
    constexpr float KA[3] = {1.170879621e-03f, 1.300729765e-05f, 1.444980246e-07f};
    const float e3 = __expf(3.0f * r);
    float g = __expf(-2.0f * (r - 0.75f) * (r - 0.75f));
#pragma unroll
    for (int m = 0; m < 8; m++) {
        s.v[m] += wp * g;
        s.v[8 + m] += wn * g;
        g *= e3 * KR[m];
    }
    float h = __expf(-(r - 1.5f) * (r - 1.5f));
#pragma unroll
    for (int c = 0; c < 3; c++) {
        const float hp = wp * h, hn = wn * h;
        s.v[SA_A + 3 * (c) + (0)] += hp * ux; s.v[SA_A + 3 * (c) + (1)] += hp * uy; s.v[SA_A + 3 * (c) + (2)] += hp * uz;
        s.v[SA_A + 3 * (3 + c) + (0)] += hn * ux; s.v[SA_A + 3 * (3 + c) + (1)] += hn * uy; s.v[SA_A + 3 * (3 + c) + (2)] += hn * uz;
        h *= e3 * KA[c];
    }
}

// erf by Abramowitz & Stegun 7.1.26 (|error| < 1.5e-7), sharing exp(-x^2) with the caller
__device__ __forceinline__ float erf_as(float x, float ex) {
    const float t = __builtin_amdgcn_rcpf(1.0f + 0.3275911f * x);
    const float p = ((((1.061405429f * t - 1.453152027f) * t + 1.421413741f) * t - 0.284496736f) * t + 0.254829592f) * t;
    return 1.0f - p * ex;
}

// vector features: one site j of a molecule inside the group cut-off, group weight wg
__device__ __forceinline__ void sg_vectors(SiteAcc &s, float dx, float dy, float dz, float r2, float qj, float wg,
                                           float alpha, float x_scale) {
    const float ri = __builtin_amdgcn_rsqf(r2), r = r2 * ri;
    const float ux = dx * ri, uy = dy * ri, uz = dz * ri;
    const float w = qj * wg;
    const float ar = alpha * r;
    const float ex = __expf(-ar * ar);
    const float kew = w * (erf_as(ar, ex) * ri * ri - 1.12837916709551f * alpha * ex * ri);
    s.v[SA_U + 3 * (0) + (0)] += kew * ux; s.v[SA_U + 3 * (0) + (1)] += kew * uy; s.v[SA_U + 3 * (0) + (2)] += kew * uz;
    const float xr = x_scale * r - 1.0f;     // 2 r / rc - 1
    const float wx = w * ux, wy = w * uy, wz = w * uz;
    float t0 = 1.0f, t1 = xr;
    s.v[SA_U + 3 * (1) + (0)] += wx; s.v[SA_U + 3 * (1) + (1)] += wy; s.v[SA_U + 3 * (1) + (2)] += wz;
    s.v[SA_U + 3 * (2) + (0)] += wx * t1; s.v[SA_U + 3 * (2) + (1)] += wy * t1; s.v[SA_U + 3 * (2) + (2)] += wz * t1;
#pragma unroll
    for (int m = 3; m < SG_MV; m++) {
        const float t2 = 2.0f * xr * t1 - t0;
        t0 = t1;
        t1 = t2;
        s.v[SA_U + 3 * (m) + (0)] += wx * t2; s.v[SA_U + 3 * (m) + (1)] += wy * t2; s.v[SA_U + 3 * (m) + (2)] += wz * t2;
    }
}

__device__ __forceinline__ float group_weight(float R, float rg) {
    const float sx = fminf(fmaxf((R - rg + 2.0f) * 0.5f, 0.0f), 1.0f);
    return 1.0f - sx * sx * (3.0f - 2.0f * sx);
}

// positions and charges of every site as float4 for the input kernel
__global__ __launch_bounds__(256) void k_sg_pack(DevSys S, MtsArgs a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < S.n_total) a.pos4[i] = make_float4((float)S.x[i], (float)S.y[i], (float)S.z[i], (float)S.q[i]);
}

// G lanes per lead walk its list rows (moving partners: every site of the partner
// molecule; frozen partners: the listed site) once per site of the lead's molecule, then
// reduce over the group.  The rows hold every lead within rc (the skin rule) and every
// frozen site within rc of the lead, which covers the group cut-off at rc; the host checks
// SG_RC + twice the largest site-to-lead distance < rc for the inputs.  FP32 throughout:
// the inputs need relative vectors to ~1e-5 A, not the list's bit-exact predicate.
template <int G>
__global__ __launch_bounds__(256) void k_sg_features(DevSys S, MtsArgs a) {
    const int tid = threadIdx.x;
    const int gid = blockIdx.x * (256 / G) + tid / G;
    const int sub = tid % G;
    const bool active = gid < S.n_lead;
    int i = 0, nsi_i = 0;
    float4 pi = make_float4(0.f, 0.f, 0.f, 0.f);
    if (active) {
        i = S.lead[gid];
        nsi_i = S.nsi[S.itype[i]];
        pi = a.pos4[i];
    }
    const float alpha = (float)S.alfa;
    const float rg = (float)sqrt(S.rcutsq), rg2 = rg * rg;
    const float x_scale = 2.0f / rg;
    const float lx = (float)S.lx, ly = (float)S.ly, lz = (float)S.lz;
    const float lxi = 1.0f / lx, lyi = 1.0f / ly, lzi = 1.0f / lz;
    const bool zp = S.zper_force != 0;
    constexpr float RC2 = SG_RC * SG_RC;
#pragma unroll 1
    for (int as = 0; as < 3; as++) {
        const bool on = active && as < nsi_i;
        SiteAcc acc;
#pragma unroll
        for (int m = 0; m < SA_N; m++) acc.v[m] = 0.0f;
        const int ia = i + as;
        if (on) {
            const float4 pa = a.pos4[ia];
            // the other sites of the molecule
            if (sub == 0)
                for (int b = 0; b < nsi_i; b++)
                    if (b != as) {
                        const float4 pb = a.pos4[i + b];
                        const float dx = pa.x - pb.x, dy = pa.y - pb.y, dz = pa.z - pb.z;
                        const float r2 = dx * dx + dy * dy + dz * dz;
                        if (r2 < RC2) sg_inputs(acc, dx, dy, dz, r2, pb.w);
                        sg_vectors(acc, dx, dy, dz, r2, pb.w, 1.0f, alpha, x_scale);
                    }
            const int n = min(S.cnt_mm[gid], S.cap_mm);
            const int *row = S.nbr_mm + (size_t)gid * S.cap_mm;
            const int last = S.n_total - 1;
            // two partner molecules per pass: all their loads are issued before either is used
#pragma unroll 1
            for (int k = sub; k < n; k += 2 * G) {
                const bool two = k + G < n;
                const int j0 = row[k], j1 = two ? row[k + G] : j0;
                float4 p0[3], p1[3];
                int ns0 = a.msize[j0], ns1 = a.msize[j1];
#pragma unroll
                for (int b = 0; b < 3; b++) {
                    p0[b] = a.pos4[min(j0 + b, last)];
                    p1[b] = a.pos4[min(j1 + b, last)];
                }
                if (!two) ns1 = 0;
#pragma unroll
                for (int u = 0; u < 2; u++) {
                    const float4 *pm = u == 0 ? p0 : p1;
                    const int nsj = u == 0 ? ns0 : ns1;
                    if (nsj == 0) continue;
                    // lead-lead image shift, applied to every site pair of the two molecules
                    float dx = pi.x - pm[0].x, dy = pi.y - pm[0].y, dz = pi.z - pm[0].z;
                    const float sx = -lx * rintf(dx * lxi), sy = -ly * rintf(dy * lyi),
                                sz = zp ? -lz * rintf(dz * lzi) : 0.0f;
                    dx += sx; dy += sy; dz += sz;
                    const float R2 = dx * dx + dy * dy + dz * dz;
                    if (R2 >= rg2) continue;
                    const float wg = group_weight(sqrtf(R2), rg);
#pragma unroll
                    for (int b = 0; b < 3; b++) {
                        if (b >= nsj) break;
                        const float4 pb = pm[b];
                        const float ex = pa.x - pb.x + sx, ey = pa.y - pb.y + sy, ez = pa.z - pb.z + sz;
                        const float r2 = ex * ex + ey * ey + ez * ez;
                        if (r2 < RC2) sg_inputs(acc, ex, ey, ez, r2, pb.w);
                        if (wg > 0.0f) sg_vectors(acc, ex, ey, ez, r2, pb.w, wg, alpha, x_scale);
                    }
                }
            }
            const int nw = min(S.cnt_mw[gid], S.cap_mw);
            const int *wrow = S.nbr_mw + (size_t)gid * S.cap_mw;
#pragma unroll 2
            for (int k = sub; k < nw; k += G) {
                const int j = wrow[k];
                const float4 pj = a.pos4[j];
                float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
                const float sx = -lx * rintf(dx * lxi), sy = -ly * rintf(dy * lyi), sz = zp ? -lz * rintf(dz * lzi) : 0.0f;
                dx += sx; dy += sy; dz += sz;
                const float R2 = dx * dx + dy * dy + dz * dz;
                if (R2 >= rg2) continue;
                const float wg = group_weight(sqrtf(R2), rg);
                const float ex = pa.x - pj.x + sx, ey = pa.y - pj.y + sy, ez = pa.z - pj.z + sz;
                const float r2 = ex * ex + ey * ey + ez * ez;
                if (r2 < RC2) sg_inputs(acc, ex, ey, ez, r2, pj.w);
                if (wg > 0.0f) sg_vectors(acc, ex, ey, ez, r2, pj.w, wg, alpha, x_scale);
            }
        }
        // reduce over the group in chunks of 6 values: the scheduling barriers keep the
        // compiler from overlapping all 82 shuffles (and their temporaries) at once
#pragma unroll
        for (int c0 = 0; c0 < SA_N; c0 += 6) {
#pragma unroll
            for (int o = G / 2; o > 0; o >>= 1)
#pragma unroll
                for (int c = c0; c < c0 + 6 && c < SA_N; c++) acc.v[c] += __shfl_xor(acc.v[c], o, 64);
            __builtin_amdgcn_sched_barrier(0);
        }
        if (on && sub == 0) {
            float *f = a.feat + (size_t)ia * SG_D;
#pragma unroll
            for (int m = 0; m < 16; m++) f[m] = acc.v[m];
            int k = 16;
#pragma unroll
            for (int p = 0; p < 6; p++)
#pragma unroll
                for (int q = p; q < 6; q++, k++)
                    f[k] = acc.v[SA_A + 3 * (p) + (0)] * acc.v[SA_A + 3 * (q) + (0)] + acc.v[SA_A + 3 * (p) + (1)] * acc.v[SA_A + 3 * (q) + (1)] + acc.v[SA_A + 3 * (p) + (2)] * acc.v[SA_A + 3 * (q) + (2)];
            const float qi = (float)S.q[ia];
            f[37] = qi;
            f[38] = qi * qi;
#pragma unroll
            for (int m = 39; m < SG_D; m++) f[m] = 0.0f;
            const float sc = (float)S.facq * qi;
            float *u = a.vec + (size_t)ia * (SG_M * 3);
#pragma unroll
            for (int m = 0; m < SG_MV; m++) {
                u[3 * m + 0] = sc * acc.v[SA_U + 3 * m + 0];
                u[3 * m + 1] = sc * acc.v[SA_U + 3 * m + 1];
                u[3 * m + 2] = sc * acc.v[SA_U + 3 * m + 2];
            }
#pragma unroll
            for (int m = 3 * SG_MV; m < 3 * SG_M; m++) u[m] = 0.0f;
        }
    }
}
template __global__ void k_sg_features<8>(DevSys, MtsArgs);
template __global__ void k_sg_features<16>(DevSys, MtsArgs);

// ---- input normalisation from the first frame -------------------------------------------
// mu, sigma per input (sigma 1 where an input does not vary) and su_m = rms |U_m|; sets the
// physics prior c_0 = su_0 (the network then starts as the tapered near-field k-space force)
// unless weights were given.  One block, fixed summation order.
__global__ __launch_bounds__(1024) void k_sg_norm(MtsArgs a, int n) {
    MtsState &st = *a.st;
    if (st.norm_ready) return;
    __shared__ double r1[16][64], r2[16][64];
    const int f = threadIdx.x & 63, sl = threadIdx.x >> 6;
    double s1 = 0.0, s2 = 0.0;
    for (int i = sl; i < n; i += 16) {
        if (f < SG_D) {
            const double v = a.feat[(size_t)i * SG_D + f];
            s1 += v;
            s2 += v * v;
        } else {
            const float *u = a.vec + (size_t)i * (SG_M * 3) + 3 * (f - SG_D);
            s2 += (double)u[0] * u[0] + (double)u[1] * u[1] + (double)u[2] * u[2];
        }
    }
    r1[sl][f] = s1;
    r2[sl][f] = s2;
    __syncthreads();
    if (threadIdx.x < 64) {
        double t1 = 0.0, t2 = 0.0;
        for (int k = 0; k < 16; k++) { t1 += r1[k][f]; t2 += r2[k][f]; }
        const double nn = n > 0 ? (double)n : 1.0;
        if (f < SG_D) {
            const double mu = t1 / nn;
            double sd = sqrt(fmax(t2 / nn - mu * mu, 0.0));
            if (!(sd > 1e-6 * fmax(1.0, fabs(mu)))) sd = 1.0;
            a.norm[f] = (float)mu;
            a.norm[SG_D + f] = (float)sd;
        } else {
            double su = sqrt(t2 / nn);
            if (!(su > 0.0)) su = 1.0;
            a.norm[2 * SG_D + (f - SG_D)] = (float)su;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        if (st.prior) {
            a.params[SG_OFF_B3] = a.norm[2 * SG_D];
            st.prior = 0;
        }
        st.norm_ready = 1;
    }
}

// ---- inference -------------------------------------------------------------------------
struct InferSmem {
    unsigned short w1[SG_H * SG_D], w2[SG_H * SG_H], w3[SG_M * SG_H];
    float mu[SG_D], isd[SG_D], isu[SG_M], b1[SG_H], b2[SG_H], b3[SG_M];
};

__device__ __forceinline__ void load_infer_smem(InferSmem &L, const MtsArgs &a) {
    for (int k = threadIdx.x; k < SG_H * SG_D; k += blockDim.x) L.w1[k] = a.wb[SG_WB_W1 + k];
    for (int k = threadIdx.x; k < SG_H * SG_H; k += blockDim.x) L.w2[k] = a.wb[SG_WB_W2 + k];
    for (int k = threadIdx.x; k < SG_M * SG_H; k += blockDim.x) L.w3[k] = a.wb[SG_WB_W3 + k];
    for (int k = threadIdx.x; k < SG_D; k += blockDim.x) {
        L.mu[k] = a.norm[k];
        L.isd[k] = 1.0f / a.norm[SG_D + k];
    }
    for (int k = threadIdx.x; k < SG_M; k += blockDim.x) {
        L.isu[k] = 1.0f / a.norm[2 * SG_D + k];
        L.b3[k] = a.params[SG_OFF_B3 + k];
    }
    for (int k = threadIdx.x; k < SG_H; k += blockDim.x) {
        L.b1[k] = a.params[SG_OFF_B1 + k];
        L.b2[k] = a.params[SG_OFF_B2 + k];
    }
}

// forward pass of one 16-site tile; z1/z2 keep the pre-activations (bias included)
struct TileFwd {
    bf16x4 x[3], h1[4], h2[4];
    f32x4 z1[4], z2[4], c;
};

__device__ __forceinline__ void tile_forward(const InferSmem &L, const float *__restrict__ feat, int s, bool valid,
                                             int col, int kg, TileFwd &T) {
#pragma unroll
    for (int kt = 0; kt < 3; kt++) {
        const int f0 = 16 * kt + 4 * kg;
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (valid) {
            v = *reinterpret_cast<const float4 *>(feat + (size_t)s * SG_D + f0);
            v.x = (v.x - L.mu[f0 + 0]) * L.isd[f0 + 0];
            v.y = (v.y - L.mu[f0 + 1]) * L.isd[f0 + 1];
            v.z = (v.z - L.mu[f0 + 2]) * L.isd[f0 + 2];
            v.w = (v.w - L.mu[f0 + 3]) * L.isd[f0 + 3];
        }
        T.x[kt] = pack4(v.x, v.y, v.z, v.w);
    }
#pragma unroll
    for (int ot = 0; ot < 4; ot++) {
        f32x4 z = zero4();
#pragma unroll
        for (int kt = 0; kt < 3; kt++) z = mfma16(frag(&L.w1[(16 * ot + col) * SG_D + 16 * kt + 4 * kg]), T.x[kt], z);
#pragma unroll
        for (int r = 0; r < 4; r++) z[r] += L.b1[16 * ot + 4 * kg + r];
        T.z1[ot] = z;
        T.h1[ot] = pack4(silu(z[0]), silu(z[1]), silu(z[2]), silu(z[3]));
    }
#pragma unroll
    for (int ot = 0; ot < 4; ot++) {
        f32x4 z = zero4();
#pragma unroll
        for (int kt = 0; kt < 4; kt++) z = mfma16(frag(&L.w2[(16 * ot + col) * SG_H + 16 * kt + 4 * kg]), T.h1[kt], z);
#pragma unroll
        for (int r = 0; r < 4; r++) z[r] += L.b2[16 * ot + 4 * kg + r];
        T.z2[ot] = z;
        T.h2[ot] = pack4(silu(z[0]), silu(z[1]), silu(z[2]), silu(z[3]));
    }
    f32x4 c = zero4();
#pragma unroll
    for (int kt = 0; kt < 4; kt++) c = mfma16(frag(&L.w3[col * SG_H + 16 * kt + 4 * kg]), T.h2[kt], c);
#pragma unroll
    for (int r = 0; r < 4; r++) c[r] += L.b3[4 * kg + r];
    T.c = c;
}

// S_i = sum_m c_m U_m / su_m: each lane its 4 outputs, then across the 4 lane groups
__device__ __forceinline__ void tile_force(const InferSmem &L, const float *__restrict__ vec, int s, bool valid, int kg,
                                           const f32x4 &c, float &fx, float &fy, float &fz) {
    fx = fy = fz = 0.0f;
    if (valid) {
        const float *u = vec + (size_t)s * (SG_M * 3) + 12 * kg;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const float cm = c[r] * L.isu[4 * kg + r];
            fx += cm * u[3 * r + 0];
            fy += cm * u[3 * r + 1];
            fz += cm * u[3 * r + 2];
        }
    }
    fx += __shfl_xor(fx, 16, 64); fy += __shfl_xor(fy, 16, 64); fz += __shfl_xor(fz, 16, 64);
    fx += __shfl_xor(fx, 32, 64); fy += __shfl_xor(fy, 32, 64); fz += __shfl_xor(fz, 32, 64);
}

// after_training = 0: this step's frame; 1: this frame again after training, 2: the
// previous full-sum frame after training (into sg_prev, for its correction with the new
// weights); 1 and 2 only on steps that trained
__global__ __launch_bounds__(256) void k_sg_infer(MtsArgs a, int n, int after_training, const DevFlags *flags) {
    const MtsState &st = *a.st;
    if (!st.norm_ready) return;
    if (after_training && !st.train_now) return;
    if (flags && (flags->overflow_mm || flags->overflow_mw)) return;
    __shared__ InferSmem L;
    load_infer_smem(L, a);
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int col = lane & 15, kg = lane >> 4;
    const int ntiles = (n + 15) / 16;
    for (int tile = blockIdx.x * 4 + wave; tile < ntiles; tile += gridDim.x * 4) {
        const int s = tile * 16 + col;
        const bool valid = s < n;
        TileFwd T;
        const bool prev = after_training == 2;
        tile_forward(L, prev ? a.feat_prev : a.feat, s, valid, col, kg, T);
        float fx, fy, fz;
        tile_force(L, prev ? a.vec_prev : a.vec, s, valid, kg, T.c, fx, fy, fz);
        float *out = prev ? a.sg_prev : a.sg;
        if (kg == 0 && valid) {
            out[s] = fx;
            out[n + s] = fy;
            out[2 * (size_t)n + s] = fz;
        }
    }
}

// ---- training: forward + backward over a pair of frames, gradient partials per block ---
struct TrainSmem {
    InferSmem L;
    unsigned short w2t[SG_H * SG_H], w3t[SG_H * SG_M];
    // per wave: operand tiles [rows][16 sites] for the weight-gradient products
    unsigned short tx[4][SG_D * 16], th1[4][SG_H * 16], th2[4][SG_H * 16], tdz1[4][SG_H * 16], tdz2[4][SG_H * 16],
        tdc[4][SG_M * 16];
    float gacc[SG_GP];
};

struct GradAcc {
    f32x4 w1[4][3], w2[4][4], w3[4];
    float b1[4][4], b2[4][4], b3[4];
};

// back-propagate dL/dS = (gx, gy, gz) of site col through the tile's forward pass T and add
// the weight gradients (the operands of the site contractions pass through the wave's LDS)
__device__ __forceinline__ void tile_backward(TrainSmem &M, int wave, const TileFwd &T, const float *__restrict__ vec,
                                              int s, bool valid, int col, int kg, float gx, float gy, float gz,
                                              GradAcc &G) {
    const InferSmem &L = M.L;
    float dc[4] = {0.f, 0.f, 0.f, 0.f};
    if (valid) {
        const float *u = vec + (size_t)s * (SG_M * 3) + 12 * kg;
#pragma unroll
        for (int r = 0; r < 4; r++)
            dc[r] = (gx * u[3 * r] + gy * u[3 * r + 1] + gz * u[3 * r + 2]) * L.isu[4 * kg + r];
    }
#pragma unroll
    for (int r = 0; r < 4; r++) G.b3[r] += dc[r];
    const bf16x4 dcf = pack4(dc[0], dc[1], dc[2], dc[3]);
    bf16x4 dz2f[4], dz1f[4];
#pragma unroll
    for (int ot = 0; ot < 4; ot++) {
        const f32x4 d = mfma16(frag(&M.w3t[(16 * ot + col) * SG_M + 4 * kg]), dcf, zero4());
        float dz[4];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            dz[r] = d[r] * dsilu(T.z2[ot][r]);
            G.b2[ot][r] += dz[r];
        }
        dz2f[ot] = pack4(dz[0], dz[1], dz[2], dz[3]);
    }
#pragma unroll
    for (int ot = 0; ot < 4; ot++) {
        f32x4 d = zero4();
#pragma unroll
        for (int kt = 0; kt < 4; kt++) d = mfma16(frag(&M.w2t[(16 * ot + col) * SG_H + 16 * kt + 4 * kg]), dz2f[kt], d);
        float dz[4];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            dz[r] = d[r] * dsilu(T.z1[ot][r]);
            G.b1[ot][r] += dz[r];
        }
        dz1f[ot] = pack4(dz[0], dz[1], dz[2], dz[3]);
    }
    unsigned short *TX = M.tx[wave], *TH1 = M.th1[wave], *TH2 = M.th2[wave], *TDZ1 = M.tdz1[wave],
                   *TDZ2 = M.tdz2[wave], *TDC = M.tdc[wave];
    // operand tiles to [rows][sites]: a lane holds rows 4 kg + r of site col
#pragma unroll
    for (int r = 0; r < 4; r++) {
#pragma unroll
        for (int kt = 0; kt < 3; kt++) TX[(16 * kt + 4 * kg + r) * 16 + col] = (unsigned short)T.x[kt][r];
#pragma unroll
        for (int ot = 0; ot < 4; ot++) {
            const int row = (16 * ot + 4 * kg + r) * 16 + col;
            TH1[row] = (unsigned short)T.h1[ot][r];
            TH2[row] = (unsigned short)T.h2[ot][r];
            TDZ1[row] = (unsigned short)dz1f[ot][r];
            TDZ2[row] = (unsigned short)dz2f[ot][r];
        }
        TDC[(4 * kg + r) * 16 + col] = (unsigned short)dcf[r];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const bf16x4 adc = frag(&TDC[col * 16 + 4 * kg]);
#pragma unroll
    for (int ct = 0; ct < 4; ct++) G.w3[ct] = mfma16(adc, frag(&TH2[(16 * ct + col) * 16 + 4 * kg]), G.w3[ct]);
#pragma unroll
    for (int rt = 0; rt < 4; rt++) {
        const bf16x4 a2 = frag(&TDZ2[(16 * rt + col) * 16 + 4 * kg]);
#pragma unroll
        for (int ct = 0; ct < 4; ct++) G.w2[rt][ct] = mfma16(a2, frag(&TH1[(16 * ct + col) * 16 + 4 * kg]), G.w2[rt][ct]);
        const bf16x4 a1 = frag(&TDZ1[(16 * rt + col) * 16 + 4 * kg]);
#pragma unroll
        for (int ct = 0; ct < 3; ct++) G.w1[rt][ct] = mfma16(a1, frag(&TX[(16 * ct + col) * 16 + 4 * kg]), G.w1[rt][ct]);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Frame a = (feat, vec, frec), frame b = (feat_prev, vec_prev, frec_prev):
// e = (S_a - S_b) - (F_a - F_b); frame a back-propagates +2e/(n t2), frame b -2e/(n t2).
// Iteration `it` of a step trains on the tiles t with t = adam_base + it (mod train_stride),
// a different subset each Adam step (the gradient of a random-like subset of the frame).
__global__ __launch_bounds__(256) void k_sg_train(MtsArgs a, int n, int it, const DevFlags *flags) {
    const MtsState &st = *a.st;
    if (!st.train_now) return;
    if (flags && (flags->overflow_mm || flags->overflow_mw)) return;
    __shared__ TrainSmem M;
    const int stride = max(a.train_stride, 1);
    const int phase = (st.adam_base + it) % stride;
    load_infer_smem(M.L, a);
    for (int k = threadIdx.x; k < SG_H * SG_H; k += blockDim.x) M.w2t[k] = a.wb[SG_WB_W2T + k];
    for (int k = threadIdx.x; k < SG_H * SG_M; k += blockDim.x) M.w3t[k] = a.wb[SG_WB_W3T + k];
    for (int k = threadIdx.x; k < SG_GP; k += blockDim.x) M.gacc[k] = 0.0f;
    __syncthreads();
    const InferSmem &L = M.L;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int col = lane & 15, kg = lane >> 4;
    const int ntiles = (n + 15) / 16;
    const int nsel = (ntiles - phase + stride - 1) / stride;   // tiles of this subset
    const int nsites = min(n, 16 * nsel);
    const float gscale = (float)(2.0 / ((double)nsites * st.t2));
    GradAcc G;
#pragma unroll
    for (int p = 0; p < 4; p++) {
        G.w3[p] = zero4();
        G.b3[p] = 0.0f;
#pragma unroll
        for (int q = 0; q < 3; q++) G.w1[p][q] = zero4();
#pragma unroll
        for (int q = 0; q < 4; q++) G.w2[p][q] = zero4();
#pragma unroll
        for (int r = 0; r < 4; r++) G.b1[p][r] = G.b2[p][r] = 0.0f;
    }
    float lacc = 0.0f;
    for (int q = blockIdx.x * 4 + wave; q < nsel; q += gridDim.x * 4) {
        const int tile = phase + q * stride;
        const int s = tile * 16 + col;
        const bool valid = s < n;
        TileFwd T;
        float bx, by, bz, ax, ay, az;
        tile_forward(L, a.feat_prev, s, valid, col, kg, T);
        tile_force(L, a.vec_prev, s, valid, kg, T.c, bx, by, bz);
        tile_forward(L, a.feat, s, valid, col, kg, T);
        tile_force(L, a.vec, s, valid, kg, T.c, ax, ay, az);
        float gx = 0.f, gy = 0.f, gz = 0.f;
        if (valid) {
            const size_t n2 = 2 * (size_t)n;
            const float ex = (ax - bx) - (float)(a.frec[s] - a.frec_prev[s]);
            const float ey = (ay - by) - (float)(a.frec[n + s] - a.frec_prev[n + s]);
            const float ez = (az - bz) - (float)(a.frec[n2 + s] - a.frec_prev[n2 + s]);
            if (kg == 0) lacc += ex * ex + ey * ey + ez * ez;
            gx = gscale * ex; gy = gscale * ey; gz = gscale * ez;
        }
        tile_backward(M, wave, T, a.vec, s, valid, col, kg, gx, gy, gz, G);
        tile_forward(L, a.feat_prev, s, valid, col, kg, T);
        tile_backward(M, wave, T, a.vec_prev, s, valid, col, kg, -gx, -gy, -gz, G);
    }
    // bias and loss partials: sum over the 16 sites (lanes) of each lane group
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) {
#pragma unroll
        for (int p = 0; p < 4; p++) {
            G.b3[p] += __shfl_xor(G.b3[p], o, 64);
#pragma unroll
            for (int r = 0; r < 4; r++) {
                G.b1[p][r] += __shfl_xor(G.b1[p][r], o, 64);
                G.b2[p][r] += __shfl_xor(G.b2[p][r], o, 64);
            }
        }
    }
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) lacc += __shfl_xor(lacc, o, 64);
    // the block's waves in order
    for (int w = 0; w < 4; w++) {
        if (wave == w) {
#pragma unroll
            for (int r = 0; r < 4; r++) {
#pragma unroll
                for (int ct = 0; ct < 4; ct++) M.gacc[SG_OFF_W3 + (4 * kg + r) * SG_H + 16 * ct + col] += G.w3[ct][r];
#pragma unroll
                for (int rt = 0; rt < 4; rt++) {
#pragma unroll
                    for (int ct = 0; ct < 4; ct++)
                        M.gacc[SG_OFF_W2 + (16 * rt + 4 * kg + r) * SG_H + 16 * ct + col] += G.w2[rt][ct][r];
#pragma unroll
                    for (int ct = 0; ct < 3; ct++)
                        M.gacc[SG_OFF_W1 + (16 * rt + 4 * kg + r) * SG_D + 16 * ct + col] += G.w1[rt][ct][r];
                }
            }
            if (col == 0) {
#pragma unroll
                for (int p = 0; p < 4; p++) {
                    M.gacc[SG_OFF_B3 + 4 * kg + p] += G.b3[p];
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        M.gacc[SG_OFF_B1 + 16 * p + 4 * kg + r] += G.b1[p][r];
                        M.gacc[SG_OFF_B2 + 16 * p + 4 * kg + r] += G.b2[p][r];
                    }
                }
            }
            if (lane == 0) M.gacc[SG_NP] += lacc * (float)(1.0 / ((double)nsites * st.t2));
        }
        __syncthreads();
    }
    for (int k = threadIdx.x; k < SG_GP; k += blockDim.x) a.gpart[(size_t)blockIdx.x * SG_GP + k] = M.gacc[k];
}

// Sum of the block partials (fixed order), Adam step (bias-corrected moments) on the FP32
// parameters, and fresh bfloat16 copies.  lr = 0: the gradient only (gsum), no update.
__global__ __launch_bounds__(256) void k_sg_adam(MtsArgs a, int nb, int it, const DevFlags *flags) {
    MtsState &st = *a.st;
    if (!st.train_now) return;
    if (flags && (flags->overflow_mm || flags->overflow_mw)) return;
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= SG_GP) return;
    double g = 0.0;
    for (int b = 0; b < nb; b++) g += a.gpart[(size_t)b * SG_GP + p];
    a.gsum[p] = (float)g;
    if (p == SG_NP) {
        st.loss_last = g;
        return;
    }
    if (!(a.lr > 0.0f)) return;
    const int t = st.adam_base + it + 1;
    const float gf = (float)g;
    const float m = 0.9f * a.adam_m[p] + 0.1f * gf;
    const float v = 0.999f * a.adam_v[p] + 0.001f * gf * gf;
    a.adam_m[p] = m;
    a.adam_v[p] = v;
    const float mh = m / (float)(1.0 - pow(0.9, (double)t));
    const float vh = v / (float)(1.0 - pow(0.999, (double)t));
    const float w = a.params[p] - a.lr * mh / (sqrtf(vh) + 1e-8f);
    a.params[p] = w;
    const unsigned short wb = f2bf(w);
    if (p < SG_OFF_B1) {
        a.wb[SG_WB_W1 + p] = wb;
    } else if (p >= SG_OFF_W2 && p < SG_OFF_B2) {
        const int k = p - SG_OFF_W2, u2 = k / SG_H, u1 = k % SG_H;
        a.wb[SG_WB_W2 + k] = wb;
        a.wb[SG_WB_W2T + u1 * SG_H + u2] = wb;
    } else if (p >= SG_OFF_W3 && p < SG_OFF_B3) {
        const int k = p - SG_OFF_W3, m3 = k / SG_H, u = k % SG_H;
        a.wb[SG_WB_W3 + k] = wb;
        a.wb[SG_WB_W3T + u * SG_M + m3] = wb;
    }
}

// Set-up of one evaluation outside any handle that allocates nothing (mdf_mts_eval_device):
// the FP32 parameters and input normalisation from the caller's weights, their bfloat16
// copies (and transposes) as k_sg_adam makes them, zero Adam moments, and the scheme state
// with t2 = mean over sites of |T_a - T_b|^2 (the loss scale).  One block; every buffer the
// later kernels read is written here, so nothing of an earlier call survives in them.
__global__ __launch_bounds__(1024) void k_sg_eval_prep(MtsArgs a, const float *__restrict__ weights, int n,
                                                       const double *__restrict__ target,
                                                       const double *__restrict__ target_b, int adam_t) {
    __shared__ double red[1024];
    const bool train = target != nullptr;
    double acc = 0.0;
    if (train)
        for (size_t k = threadIdx.x; k < 3 * (size_t)n; k += blockDim.x) {
            const double d = target[k] - target_b[k];
            acc += d * d;
        }
    red[threadIdx.x] = acc;
    for (int p = threadIdx.x; p < SG_NP; p += blockDim.x) {
        a.params[p] = weights[p];
        a.adam_m[p] = 0.0f;
        a.adam_v[p] = 0.0f;
    }
    for (int k = threadIdx.x; k < SG_NORM_N; k += blockDim.x) a.norm[k] = weights[SG_NP + k];
    for (int k = threadIdx.x; k < SG_H * SG_D; k += blockDim.x) a.wb[SG_WB_W1 + k] = f2bf(weights[SG_OFF_W1 + k]);
    for (int k = threadIdx.x; k < SG_H * SG_H; k += blockDim.x) {
        const int u2 = k / SG_H, u1 = k % SG_H;
        const unsigned short v = f2bf(weights[SG_OFF_W2 + k]);
        a.wb[SG_WB_W2 + k] = v;
        a.wb[SG_WB_W2T + u1 * SG_H + u2] = v;
    }
    for (int k = threadIdx.x; k < SG_M * SG_H; k += blockDim.x) {
        const int m3 = k / SG_H, u = k % SG_H;
        const unsigned short v = f2bf(weights[SG_OFF_W3 + k]);
        a.wb[SG_WB_W3 + k] = v;
        a.wb[SG_WB_W3T + u * SG_M + m3] = v;
    }
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        MtsState st = {};
        st.norm_ready = 1;
        st.have_prev = train ? 1 : 0;
        st.train_now = train ? 1 : 0;
        st.t2 = train ? red[0] / n : 0.0;
        st.adam_base = adam_t;
        st.adam_t = adam_t + (train ? 1 : 0);
        *a.st = st;
    }
}

// weight of the correction's slope at this step: (t - t0) / (t0 - t1), or 0
__device__ __forceinline__ double slope_weight(const MtsState &st) {
    if (!st.have_slope || st.delta_step <= st.delta_prev_step) return 0.0;
    return (double)(st.step - st.delta_step) / (double)(st.delta_step - st.delta_prev_step);
}

// ---- steps without the full sum: surrogate + correction, energy along the path --------
// F = S + delta per site; the k-space energy follows by the trapezoid rule along the moved
// path, dE = -sum_i (F_prev + F) / 2 . (r - r_prev), accumulated since the last full sum.
__global__ __launch_bounds__(256) void k_sg_combine(DevSys S, MtsArgs a) {
    const MtsState &st = *a.st;
    if (st.exact || overflowed(S)) return;
    const int n = S.n_move;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool net = a.model != 0 && st.norm_ready;
    const double w = slope_weight(st);
    double de = 0.0;
    if (i < n) {
        const double r[3] = {S.x[i], S.y[i], S.z[i]};
        const double L[3] = {S.lx, S.ly, S.lz}, Li[3] = {S.lxi, S.lyi, S.lzi};
        for (int c = 0; c < 3; c++) {
            const size_t k = (size_t)c * n + i;
            const double f = (net ? (double)a.sg[k] : 0.0) + a.delta[k] + w * (a.delta[k] - a.delta_prev[k]);
            double d = r[c] - a.rprev[k];
            if (c < 2 || S.zper_force) d -= L[c] * anint_d(d * Li[c]);
            de -= 0.5 * (a.fprev[k] + f) * d;
            a.fmts[k] = f;
            a.rprev[k] = r[c];
            a.fprev[k] = f;
        }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) de += __shfl_xor(de, o, 64);
    __shared__ double ws[4];
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = de;
    __syncthreads();
    if (threadIdx.x == 0) a.epart[blockIdx.x] = ws[0] + ws[1] + ws[2] + ws[3];
}

// energies of a step without the full sum: the last full sum's plus the path integral
__global__ void k_sg_finish(DevSys S, MtsArgs a, Scalars *scal) {
    if (threadIdx.x != 0) return;
    MtsState &st = *a.st;
    if (st.exact || overflowed(S)) return;
    double de = 0.0;
    for (int b = 0; b < a.n_sblk; b++) de += a.epart[b];
    st.de_acc += de;
    const double e = st.de_acc;
    scal->e_recip += e;
    scal->e_recip_wl += e;
    scal->e_conf += e;
    scal->e_avex += e;
    if (!isfinite(scal->e_conf)) scal->nonfinite = 1;
}

// ---- steps with the full sum: target, judge, correction -----------------------------------
__global__ __launch_bounds__(256) void k_sg_exact(DevSys S, MtsArgs a, const double *__restrict__ fpart, int n_groups) {
    const MtsState &st = *a.st;
    if (!st.exact || overflowed(S)) return;
    const int n = S.n_move;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool net = a.model != 0 && st.norm_ready;
    const double w = slope_weight(st);
    double v[4] = {0.0, 0.0, 0.0, 0.0};   // judged error^2, |F|^2, |F_k|^2, |F_k - F_k(previous)|^2
    if (i < n) {
        const double ft[3] = {S.fx[i], S.fy[i], S.fz[i]};
        for (int c = 0; c < 3; c++) {
            double fr = 0.0;
            for (int g = 0; g < n_groups; g++) fr += fpart[((size_t)g * 3 + c) * n + i];
            const size_t k = (size_t)c * n + i;
            a.frec[k] = fr;
            const double d = (net ? (double)a.sg[k] : 0.0) + a.delta[k] + w * (a.delta[k] - a.delta_prev[k]) - fr;
            v[0] += d * d;
            v[1] += ft[c] * ft[c];
            v[2] += fr * fr;
            const double dd = st.have_prev ? fr - a.frec_prev[k] : 0.0;
            v[3] += dd * dd;
        }
    }
    __shared__ double ws[4][4];
#pragma unroll
    for (int c = 0; c < 4; c++) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v[c] += __shfl_xor(v[c], o, 64);
        if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6][c] = v[c];
    }
    __syncthreads();
    if (threadIdx.x < 4) a.jpart[blockIdx.x * 4 + threadIdx.x] = ws[0][threadIdx.x] + ws[1][threadIdx.x] +
                                                                 ws[2][threadIdx.x] + ws[3][threadIdx.x];
}

__global__ __launch_bounds__(256) void k_sg_judge(DevSys S, MtsArgs a) {
    MtsState &st = *a.st;
    if (!st.exact) return;
    if (overflowed(S)) {
        if (threadIdx.x == 0) st.train_now = 0;
        return;
    }
    __shared__ double red[4][256];
    for (int c = 0; c < 4; c++) {
        double acc = 0.0;
        for (int b = threadIdx.x; b < a.n_sblk; b += 256) acc += a.jpart[b * 4 + c];
        red[c][threadIdx.x] = acc;
    }
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w)
            for (int c = 0; c < 4; c++) red[c][threadIdx.x] += red[c][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    const double n = S.n_move > 0 ? (double)S.n_move : 1.0;
    st.t2 = red[3][0] / n;
    if (st.state == 1 && st.have_delta) {
        const double err = red[1][0] > 0.0 ? sqrt(red[0][0] / red[1][0]) : 0.0;
        st.err_last = err;
        st.err_max = fmax(st.err_max, err);
        st.err_sum += err;
        st.n_judged++;
        if (!(err <= a.tol)) {   // fall back: full sums (and training) again
            st.state = 0;
            st.frames = 0;
            st.n_fallback++;
        }
    }
    // every full-sum step trains on the pair (this frame, the previous full-sum frame): every
    // step while collecting, and online at each full sum of the scheme
    st.train_now = (a.model != 0 && a.train_iters > 0 && st.norm_ready && st.have_prev && st.t2 > 0.0) ? 1 : 0;
    st.adam_base = st.adam_t;
    if (st.train_now) st.adam_t += a.train_iters;
}

// correction for the following steps, the path start, and this frame kept as the previous
// one for training pairs
__global__ __launch_bounds__(256) void k_sg_delta(DevSys S, MtsArgs a) {
    const MtsState &st = *a.st;
    if (!st.exact || overflowed(S)) return;
    const int n = S.n_move;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool net = a.model != 0 && st.norm_ready;
    const double r[3] = {S.x[i], S.y[i], S.z[i]};
    for (int c = 0; c < 3; c++) {
        const size_t k = (size_t)c * n + i;
        const double fr = a.frec[k];
        // the previous correction with the weights the next steps use
        a.delta_prev[k] = (net && st.train_now) ? a.frec_prev[k] - (double)a.sg_prev[k] : a.delta[k];
        a.delta[k] = fr - (net ? (double)a.sg[k] : 0.0);
        a.rprev[k] = r[c];
        a.fprev[k] = fr;
        a.frec_prev[k] = fr;
    }
    if (a.model != 0) {
        const float4 *f = reinterpret_cast<const float4 *>(a.feat + (size_t)i * SG_D);
        float4 *fp = reinterpret_cast<float4 *>(a.feat_prev + (size_t)i * SG_D);
        for (int k = 0; k < SG_D / 4; k++) fp[k] = f[k];
        const float4 *u = reinterpret_cast<const float4 *>(a.vec + (size_t)i * SG_M * 3);
        float4 *up = reinterpret_cast<float4 *>(a.vec_prev + (size_t)i * SG_M * 3);
        for (int k = 0; k < SG_M * 3 / 4; k++) up[k] = u[k];
    }
}

__global__ void k_sg_post(DevSys S, MtsArgs a) {
    if (threadIdx.x != 0) return;
    MtsState &st = *a.st;
    if (!st.exact || overflowed(S)) return;
    if (st.state == 0) {
        st.frames++;
        if (st.frames >= a.train_frames) {
            st.state = 1;
            st.phase = 0;
        }
    }
    // two corrections (the previous one redone with new weights after training) give the slope
    st.have_slope = st.have_delta;
    st.delta_prev_step = st.delta_step;
    st.delta_step = st.step;
    st.have_delta = 1;
    st.have_prev = 1;
    st.de_acc = 0.0;
}

}  // namespace mdf
