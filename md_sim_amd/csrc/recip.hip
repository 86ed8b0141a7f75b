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

template <typename R, int NS>
__global__ __launch_bounds__(SF_BLOCK) void k_sf(const double *__restrict__ x, const double *__restrict__ y,
                                                 const double *__restrict__ z, const double *__restrict__ q, int n,
                                                 double cux, double cuy, double cuz,
                                                 const KChunk *__restrict__ chunks, int n_chunks, int cpb,
                                                 double2 *__restrict__ partial, int nk) {
    constexpr int NW = SF_BLOCK / 64;
    __shared__ double red[NW][KC][64];   // one half (KC doubles) of the accumulators at a time
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int sb = blockIdx.x;
    // this thread's sites (coalesced: site = sb*SF_BLOCK*NS + m*SF_BLOCK + tid)
    double su[NS], sv[NS], sw[NS];
    R sq[NS], stc[NS], sts[NS];
#pragma unroll
    for (int m = 0; m < NS; m++) {
        int i = sb * SF_BLOCK * NS + m * SF_BLOCK + threadIdx.x;
        if (i < n) {
            su[m] = x[i] * cux; sv[m] = y[i] * cuy; sw[m] = z[i] * cuz; sq[m] = (R)q[i];
        } else {
            su[m] = sv[m] = sw[m] = 0.0; sq[m] = R(0);
        }
        sincospi_r<R>(sw[m], sts[m], stc[m]);
    }
    const int c_begin = blockIdx.y * cpb, c_end = min(n_chunks, c_begin + cpb);
    for (int cidx = c_begin; cidx < c_end; cidx++) {
        const KChunk ch = chunks[cidx];
        R ar[KC], ai[KC];
#pragma unroll
        for (int k = 0; k < KC; k++) { ar[k] = R(0); ai[k] = R(0); }
#pragma unroll
        for (int m = 0; m < NS; m++) {
            double t0 = (double)ch.kl * su[m] + (double)ch.km * sv[m] + (double)ch.kn0 * sw[m];
            R s0, c0;
            sincospi_r<R>(t0, s0, c0);
            R er = sq[m] * c0, ei = sq[m] * s0;
            if constexpr (sizeof(R) == 8) {
                // Chebyshev: e_{n+1} = 2c e_n - e_{n-1}
                const R c2 = R(2) * stc[m];
                R pr = er * stc[m] + ei * sts[m];   // e_{-1} = e_0 * conj(step)
                R pi_ = ei * stc[m] - er * sts[m];
#pragma unroll
                for (int k = 0; k < KC; k++) {
                    ar[k] += er; ai[k] += ei;
                    R nr = c2 * er - pr, ni = c2 * ei - pi_;
                    pr = er; pi_ = ei; er = nr; ei = ni;
                }
            } else {
#pragma unroll
                for (int k = 0; k < KC; k++) {
                    ar[k] += er; ai[k] += ei;
                    R nr = er * stc[m] - ei * sts[m];
                    R ni = er * sts[m] + ei * stc[m];
                    er = nr; ei = ni;
                }
            }
        }
        // block reduction: real parts then imaginary parts through LDS
#pragma unroll
        for (int half = 0; half < 2; half++) {
#pragma unroll
            for (int k = 0; k < KC; k++) red[wave][k][lane] = (double)(half == 0 ? ar[k] : ai[k]);
            __syncthreads();
            // thread t: value k = t / 16, slice p = t % 16 of the 64*NW partial sums
            const int k = threadIdx.x >> 4, p = threadIdx.x & 15;
            double s = 0.0;
            if (k < KC) {
#pragma unroll
                for (int w = 0; w < NW; w++)
#pragma unroll
                    for (int l = 0; l < 4; l++) s += red[w][k][p * 4 + l];
            }
            s += __shfl_xor(s, 1, 64);
            s += __shfl_xor(s, 2, 64);
            s += __shfl_xor(s, 4, 64);
            s += __shfl_xor(s, 8, 64);
            if (p == 0 && k < ch.len) {
                double *dst = reinterpret_cast<double *>(&partial[(size_t)sb * nk + ch.koff + k]);
                dst[half] = s;
            }
            __syncthreads();
        }
    }
}

template __global__ void k_sf<double, 4>(const double *, const double *, const double *, const double *, int, double,
                                        double, double, const KChunk *, int, int, double2 *, int);
template __global__ void k_sf<float, 4>(const double *, const double *, const double *, const double *, int, double,
                                       double, double, const KChunk *, int, int, double2 *, int);
template __global__ void k_sf<double, 8>(const double *, const double *, const double *, const double *, int, double,
                                        double, double, const KChunk *, int, int, double2 *, int);
template __global__ void k_sf<float, 8>(const double *, const double *, const double *, const double *, int, double,
                                       double, double, const KChunk *, int, int, double2 *, int);

// S(k) = sum over site blocks + static frozen part; T(k) = A_k conj(S(k)); per-block
// partial energies and virial:  eK = A|S|^2, eLL = A|S_moving|^2, vir_ab = 1/2 eK (d_ab - g k_a k_b)
__global__ __launch_bounds__(256) void k_sf_reduce(const double2 *__restrict__ partial, int n_sb, int nk,
                                                   const double2 *__restrict__ s_wall, const double *__restrict__ ak,
                                                   const int *__restrict__ kvec, double pipix, double pipiy,
                                                   double pipiz, double lamb, double2 *__restrict__ s_out,
                                                   double2 *__restrict__ t_out, double *__restrict__ epart) {
    __shared__ double red[RE_N][256];
    double acc[RE_N];
#pragma unroll
    for (int c = 0; c < RE_N; c++) acc[c] = 0.0;
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nk) {
        double sr = 0.0, si = 0.0;
        for (int b = 0; b < n_sb; b++) {
            double2 v = partial[(size_t)b * nk + k];
            sr += v.x; si += v.y;
        }
        double2 w = s_wall[k];
        double qr = sr + w.x, qi = si + w.y;
        s_out[k] = make_double2(qr, qi);
        double a = ak[k];
        t_out[k] = make_double2(a * qr, -a * qi);
        double ek = a * (qr * qr + qi * qi);
        acc[0] = ek;
        acc[1] = a * (sr * sr + si * si);
        double kx = pipix * kvec[3 * k], ky = pipiy * kvec[3 * k + 1], kz = pipiz * kvec[3 * k + 2];
        double g = 2.0 * (1.0 / (kx * kx + ky * ky + kz * kz) + lamb);
        double h = 0.5 * ek;
        acc[2] = h * (1.0 - g * kx * kx);
        acc[3] = h * (1.0 - g * ky * ky);
        acc[4] = h * (1.0 - g * kz * kz);
        acc[5] = -h * g * kx * ky;
        acc[6] = -h * g * kx * kz;
        acc[7] = -h * g * ky * kz;
    }
#pragma unroll
    for (int c = 0; c < RE_N; c++) red[c][threadIdx.x] = acc[c];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s)
#pragma unroll
            for (int c = 0; c < RE_N; c++) red[c][threadIdx.x] += red[c][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x < RE_N) epart[blockIdx.x * RE_N + threadIdx.x] = red[threadIdx.x][0];
}

template <typename R>
__global__ __launch_bounds__(256) void k_bp(const double *__restrict__ x, const double *__restrict__ y,
                                            const double *__restrict__ z, const double *__restrict__ q, int n,
                                            double cux, double cuy, double cuz, const KRow *__restrict__ rows,
                                            int n_rows, int rpg, const double2 *__restrict__ T, double pipix,
                                            double pipiy, double pipiz, double *__restrict__ fpart) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int g = blockIdx.y;
    const bool ok = i < n;
    const int ii = ok ? i : 0;
    const double u = x[ii] * cux, v = y[ii] * cuy, w = z[ii] * cuz;
    const R qq = ok ? (R)q[ii] : R(0);
    R sts, stc;
    sincospi_r<R>(w, sts, stc);
    double fx = 0.0, fy = 0.0, fz = 0.0;
    const int r_begin = g * rpg, r_end = min(n_rows, r_begin + rpg);
    for (int r = r_begin; r < r_end; r++) {
        const KRow row = rows[r];
        double t0 = (double)row.kl * u + (double)row.km * v + (double)row.kn0 * w;
        R s0, c0;
        sincospi_r<R>(t0, s0, c0);
        R er = qq * c0, ei = qq * s0;
        R acc = R(0), accz = R(0);
        const int len = row.kn1 - row.kn0 + 1;
        const double2 *Tr = T + row.koff;
        if constexpr (sizeof(R) == 8) {
            const R c2 = R(2) * stc;
            R pr = er * stc + ei * sts, pi_ = ei * stc - er * sts;
            R kn = (R)row.kn0;
            for (int k = 0; k < len; k++) {
                const double2 t = Tr[k];
                R tt = t.x * ei + t.y * er;
                acc += tt;
                accz += kn * tt;
                kn += R(1);
                R nr = c2 * er - pr, ni = c2 * ei - pi_;
                pr = er; pi_ = ei; er = nr; ei = ni;
            }
        } else {
            R kn = (R)row.kn0;
            for (int k = 0; k < len; k++) {
                const double2 t = Tr[k];
                R tt = (R)t.x * ei + (R)t.y * er;
                acc += tt;
                accz += kn * tt;
                kn += R(1);
                R nr = er * stc - ei * sts;
                R ni = er * sts + ei * stc;
                er = nr; ei = ni;
            }
        }
        fx += (double)acc * (pipix * (double)row.kl);
        fy += (double)acc * (pipiy * (double)row.km);
        fz += (double)accz * pipiz;
    }
    if (ok) {
        fpart[((size_t)g * 3 + 0) * n + i] = fx;
        fpart[((size_t)g * 3 + 1) * n + i] = fy;
        fpart[((size_t)g * 3 + 2) * n + i] = fz;
    }
}

template __global__ void k_bp<double>(const double *, const double *, const double *, const double *, int, double,
                                      double, double, const KRow *, int, int, const double2 *, double, double, double,
                                      double *);
template __global__ void k_bp<float>(const double *, const double *, const double *, const double *, int, double,
                                     double, double, const KRow *, int, int, const double2 *, double, double, double,
                                     double *);

}  // namespace mdf
