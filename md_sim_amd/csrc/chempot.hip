// Widom test-particle insertions of an SPC/E water into the step's configuration: the
// reference's chempot (src/forcem.f90:1226-1507; identical in src/forcem_back.f90:1192-1473)
// as five kernels over one call's insertions, whose random geometry the host draws with the
// reference's generator (the draw sequence depends on where the call ends, so it is serial).
//   k_chem_scan   one workgroup per insertion: the oxygen against every site but type 2
//                 within rcut (overlap test, vdW + erfc Coulomb), partners compacted in
//                 site order                                          (src/forcem.f90:1282-1335)
//   k_chem_first  the first overlapping insertion ends the call (`exit`, :1336)
//   k_chem_rot    one workgroup per orientation: hydrogen real-space terms over the partners,
//                 0.5 A_k Re(rho_ins conj S(k)) over the k-set from e^{i k r} tables in LDS,
//                 slab term, wall tail terms; exp(-beta u)               (:1380-1495)
//   k_chem_bsum   per insertion, its orientations' factors summed in order
//   k_chem_accum  per bin, pave(il) += those sums and inave(il) += irot  (:1277,1495)
// Everything in FP64 whatever the force path's precision.
#include "mdf_internal.h"

namespace mdf {

__device__ __forceinline__ double chem_erfc(const DevSys &S, double ar) {
    if (S.erfc_mode == 0) {
        const double t = 1.0 / (1.0 + S.p * ar);
        return ((((S.a5 * t + S.a4) * t + S.a3) * t + S.a2) * t + S.a1) * t * exp(-ar * ar);
    }
    return erfc(ar);
}

__device__ __forceinline__ double block_sum_256(double v, double *red) {
    red[threadIdx.x] = v;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    return red[0];
}

__global__ __launch_bounds__(256) void k_chem_scan(DevSys S, ChemArgs c) {
    __shared__ double red[256];
    __shared__ int s_wcnt[4];
    __shared__ int s_over, s_base;
    const int ic = blockIdx.x;
    const double ox = c.ins[3 * ic], oy = c.ins[3 * ic + 1], oz = c.ins[3 * ic + 2];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int nt1 = S.n_types + 1;
    if (threadIdx.x == 0) { s_over = 0; s_base = 0; }
    __syncthreads();
    double acc = 0.0;
    for (int j0 = 0; j0 < S.n_total; j0 += 256) {
        const int j = j0 + threadIdx.x;
        bool keep = false;
        double xd = 0.0, yd = 0.0, zd = 0.0;
        if (j < S.n_total && S.itype[j] != 2) {
            const int jt = S.itype[j];
            zd = __dsub_rn(oz, S.z[j]);
            if (!(fabs(zd) > c.rcut)) {
                xd = __dsub_rn(ox, S.x[j]); xd = __dsub_rn(xd, __dmul_rn(S.lx, anint_d(__dmul_rn(xd, S.lxi))));
                yd = __dsub_rn(oy, S.y[j]); yd = __dsub_rn(yd, __dmul_rn(S.ly, anint_d(__dmul_rn(yd, S.lyi))));
                const double rsq = __dadd_rn(__dadd_rn(__dmul_rn(xd, xd), __dmul_rn(yd, yd)), __dmul_rn(zd, zd));
                if (!(rsq > S.rcutsq)) {
                    if (rsq < c.rwinsmin && (jt == 1 || rsq < c.rinsmin)) {
                        s_over = 1;
                    } else {
                        keep = true;
                        const int ijt = S.ijtype[1 * nt1 + jt];
                        const double rsqi = 1.0 / rsq, rij = sqrt(rsq);
                        const int vk = S.vkind[ijt];
                        double e = 0.0;
                        if (vk == 1) {
                            const double r6 = S.sig6[ijt] * (rsqi * rsqi * rsqi);
                            e = S.eps4[ijt] * r6 * (r6 - 1.0);
                        } else if (vk == 2) {
                            const double r6 = S.cc[ijt] * (rsqi * rsqi * rsqi);
                            e = S.aa[ijt] * exp(-rij * S.rhoi[ijt]) - r6;
                        }
                        e += chem_erfc(S, S.alfa * rij) * (S.facq * c.qsw[0] * S.q[j] / rij);
                        acc += e;
                    }
                }
            }
        }
        // partners in site order: wave ballots, then the waves in order
        const unsigned long long m = __ballot(keep);
        const int below = __popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) s_wcnt[wv] = __popcll(m);
        __syncthreads();
        int off = s_base;
        for (int w = 0; w < wv; w++) off += s_wcnt[w];
        if (keep) {
            const int k = off + below;
            if (k < c.cap) {
                const size_t at = (size_t)ic * c.cap + k;
                c.nbr[at] = j;
                c.nbr_d[3 * at] = xd; c.nbr_d[3 * at + 1] = yd; c.nbr_d[3 * at + 2] = zd;
            }
        }
        __syncthreads();
        if (threadIdx.x == 0) s_base += s_wcnt[0] + s_wcnt[1] + s_wcnt[2] + s_wcnt[3];
        __syncthreads();
    }
    const double u = block_sum_256(acc, red);
    if (threadIdx.x == 0) {
        c.u_ox[ic] = u;
        c.n_nbr[ic] = s_base;
        c.overlap[ic] = s_over;
    }
}

__global__ __launch_bounds__(256) void k_chem_first(DevSys S, ChemArgs c) {
    __shared__ int s_first, s_max;
    if (threadIdx.x == 0) { s_first = c.n_ins; s_max = 0; }
    __syncthreads();
    for (int ic = threadIdx.x; ic < c.n_ins; ic += blockDim.x)
        if (c.overlap[ic]) atomicMin(&s_first, ic);
    __syncthreads();
    // partner lists matter only for insertions before the first overlap
    for (int ic = threadIdx.x; ic < s_first; ic += blockDim.x)
        if (c.n_nbr[ic] > c.cap) atomicMax(&s_max, c.n_nbr[ic]);
    __syncthreads();
    if (threadIdx.x == 0) {
        c.status[0] = s_first;
        c.status[1] = s_max > 0 ? 1 : 0;
        // a step whose pair list overflowed is run again: insert only then
        c.status[2] = (S.flags->overflow_mm || S.flags->overflow_mw) ? 1 : 0;
        c.status[3] = s_max;
    }
}

__global__ __launch_bounds__(256) void k_chem_rot(DevSys S, ChemArgs c) {
    extern __shared__ double2 tab[];   // per site: kl 0..klmax, km -kmmax..kmmax, kn -knmax..knmax
    __shared__ double red[256];
    const int r = blockIdx.x;
    const int ic = c.rot_ins[r];
    if (c.status[1] || c.status[2] || ic >= c.status[0]) return;   // uniform over the block
    double sx[3], sy[3], sz[3];
    sx[0] = c.ins[3 * ic]; sy[0] = c.ins[3 * ic + 1]; sz[0] = c.ins[3 * ic + 2];
    sx[1] = c.rot[6 * r]; sy[1] = c.rot[6 * r + 1]; sz[1] = c.rot[6 * r + 2];
    sx[2] = c.rot[6 * r + 3]; sy[2] = c.rot[6 * r + 4]; sz[2] = c.rot[6 * r + 5];
    const int nl = c.klmax + 1, nm = 2 * c.kmmax + 1, nn = 2 * c.knmax + 1, per = nl + nm + nn;
    for (int e = threadIdx.x; e < 3 * per; e += blockDim.x) {
        const int a = e / per, k = e - a * per;
        double ang;
        if (k < nl) ang = (double)k * (c.pipix * sx[a]);
        else if (k < nl + nm) ang = (double)(k - nl - c.kmmax) * (c.pipiy * sy[a]);
        else ang = (double)(k - nl - nm - c.knmax) * (c.pipiz * sz[a]);
        double sn, cs;
        sincos(ang, &sn, &cs);
        tab[e] = make_double2(cs, sn);
    }
    __syncthreads();

    double acc = 0.0;
    // hydrogen terms over the partners (src/forcem.f90:1381-1433)
    const int np = min(c.n_nbr[ic], c.cap);
    const double ox = sx[0], oy = sy[0], oz = sz[0];
    for (int k = threadIdx.x; k < np; k += blockDim.x) {
        const size_t at = (size_t)ic * c.cap + k;
        const int j = c.nbr[at];
        const double xio = c.nbr_d[3 * at], yio = c.nbr_d[3 * at + 1], zio = c.nbr_d[3 * at + 2];
        const int jt = S.itype[j];
        if (jt == 1) {
            const int jend = j + S.nsi[jt];
            for (int ix = 0; ix < 3; ix++) {
                const double xi = (xio + sx[ix]) - ox, yi = (yio + sy[ix]) - oy, zi = (zio + sz[ix]) - oz;
                const double qx = S.facq * c.qsw[ix];
                for (int jx = (ix == 0 ? j + 1 : j); jx < jend; jx++) {
                    const double xd = xi - (S.x[jx] - S.x[j]), yd = yi - (S.y[jx] - S.y[j]), zd = zi - (S.z[jx] - S.z[j]);
                    const double rij = sqrt(xd * xd + yd * yd + zd * zd);
                    acc += chem_erfc(S, S.alfa * rij) * (qx * S.q[jx] / rij);
                }
            }
        } else {
            for (int ix = 1; ix < 3; ix++) {
                const double xi = (xio + sx[ix]) - ox, yi = (yio + sy[ix]) - oy, zi = (zio + sz[ix]) - oz;
                const double rij = sqrt(xi * xi + yi * yi + zi * zi);
                acc += chem_erfc(S, S.alfa * rij) * (S.facq * c.qsw[ix] * S.q[j] / rij);
            }
        }
    }
    // k-space: 0.5 A_k Re(rho_ins(k) conj S(k))  (src/forcem.f90:1439-1475)
    for (int k = threadIdx.x; k < c.nk; k += blockDim.x) {
        const int kl = c.kvec[3 * k], km = c.kvec[3 * k + 1] + c.kmmax, kn = c.kvec[3 * k + 2] + c.knmax;
        double re = 0.0, im = 0.0;
#pragma unroll
        for (int a = 0; a < 3; a++) {
            const double2 ex = tab[a * per + kl], ey = tab[a * per + nl + km], ez = tab[a * per + nl + nm + kn];
            const double xr = ex.x * ey.x - ex.y * ey.y, xi = ex.x * ey.y + ex.y * ey.x;
            re += c.qsw[a] * (xr * ez.x - xi * ez.y);
            im += c.qsw[a] * (xr * ez.y + xi * ez.x);
        }
        const double2 sk = c.sk[k];
        acc += 0.5 * c.ak[k] * (re * sk.x + im * sk.y);
    }
    const double tot = block_sum_256(acc, red);
    if (threadIdx.x != 0) return;
    double u = tot + c.u_ox[ic];
    if (c.crf_scale != 0.0) {
        const double crf = c.crf_scale * (*c.dipz);
        u -= crf * (c.qsw[0] * sz[0] + c.qsw[1] * sz[1] + c.qsw[2] * sz[2]);
    }
    // wall tail terms of type 1 for both walls (src/forcem.f90:1481-1492)
    for (int w = -1; w <= 1; w += 2) {
        const double dw = c.lwh - (double)w * oz;
        if (dw < c.rcut) {
            u += c.ulr_w;
        } else {
            const double d3 = dw * dw * dw, d9 = d3 * d3 * d3;
            u += c.ulr_l;
            u += c.lja / (9.0 * d9) - c.ljc / (3.0 * d3);
            const double t = c.bur, ex = c.bua * exp(-dw / t);
            u += t * (2.0 * t + dw) * ex - c.buc / (3.0 * d3);
        }
    }
    c.boltz[r] = exp(-c.beta * u);
}

// per insertion: the sum of its orientations' factors, in orientation order (0 from the
// first overlap on)
__global__ __launch_bounds__(256) void k_chem_bsum(ChemArgs c) {
    const int ic = blockIdx.x * blockDim.x + threadIdx.x;
    if (ic >= c.n_ins || c.status[1] || c.status[2]) return;
    double p = 0.0;
    if (ic < c.status[0]) {
        const int r0 = c.ins_rot[ic], r1 = c.ins_rot[ic + 1];
        for (int r = r0; r < r1; r++) p += c.boltz[r];
    }
    c.bsum[ic] = p;
}

// per bin: pave += the insertions' sums, inave += their orientation counts (the overlapping
// insertion counts, src/forcem.f90:1277); fixed assignment and tree order, so the sums are
// the same run to run
__global__ __launch_bounds__(256) void k_chem_accum(ChemArgs c) {
    __shared__ double red[256];
    __shared__ unsigned long long redn[256];
    if (c.status[1] || c.status[2]) return;
    const int last = min(c.status[0], c.n_ins - 1);
    for (int b = blockIdx.x; b < c.n_bins; b += gridDim.x) {
        double p = 0.0;
        unsigned long long n = 0;
        for (int ic = threadIdx.x; ic <= last; ic += blockDim.x) {
            if (c.ins_il[ic] - 1 != b) continue;
            n += (unsigned long long)(c.ins_rot[ic + 1] - c.ins_rot[ic]);
            p += c.bsum[ic];
        }
        red[threadIdx.x] = p;
        redn[threadIdx.x] = n;
        __syncthreads();
        for (int st = 128; st > 0; st >>= 1) {
            if (threadIdx.x < st) {
                red[threadIdx.x] += red[threadIdx.x + st];
                redn[threadIdx.x] += redn[threadIdx.x + st];
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            c.pave[b] += red[0];
            c.inave[b] += redn[0];
        }
        __syncthreads();
    }
}

}  // namespace mdf
