// Force assembly and reductions after the pair and reciprocal kernels.
//
//  k_epilogue: site force = pair part + reciprocal part (sum over row groups) + wall
//              long-range tail (src/forcem.f90:509-543) + slab correction
//              (src/forcem.f90:923-928) + angle bending (src/forcem.f90:565-614);
//              per-block partials of ulr, the wall z force it
//              adds, and the total force (fws(0), src/forcem.f90:990-992).
//              Each block also forms its molecules' COM forces and torques
//              (src/forcem.f90:1003-1024).
//  k_colsum:   per-lead partial columns -> per-block, per-lead-type sums.
//  k_final:    energies per type pair, reciprocal energies, uconf, pressure from the wall
//              forces, virial (src/forcem.f90:1032-1081).
#include "mdf_internal.h"

namespace mdf {

// Harmonic angle bending of the force_intra term (src/forcem.f90:565-614): a = r_k - r_j,
// b = r_j - r_i (x,y minimum image; z too in the 3-D periodic form), cos = -a.b/|a||b|;
// the force of k(theta - theta0) acts on the end site k and the vertex site j, never on i.
// role 0 adds the vertex part, role 1 the end-site part.
__device__ void angle_force(const DevSys &S, int a, int role, double &fx, double &fy, double &fz) {
    const int i = S.ang_ijk[3 * a], j = S.ang_ijk[3 * a + 1], k = S.ang_ijk[3 * a + 2];
    double ax = S.x[k] - S.x[j]; ax = ax - S.lx * anint_d(ax * S.lxi);
    double ay = S.y[k] - S.y[j]; ay = ay - S.ly * anint_d(ay * S.lyi);
    double az = S.z[k] - S.z[j];
    double bx = S.x[j] - S.x[i]; bx = bx - S.lx * anint_d(bx * S.lxi);
    double by = S.y[j] - S.y[i]; by = by - S.ly * anint_d(by * S.lyi);
    double bz = S.z[j] - S.z[i];
    if (S.zper_force) {
        az = az - S.lz * anint_d(az * S.lzi);
        bz = bz - S.lz * anint_d(bz * S.lzi);
    }
    const double ai = 1.0 / sqrt(ax * ax + ay * ay + az * az);
    const double bi = 1.0 / sqrt(bx * bx + by * by + bz * bz);
    const double ab = ax * bx + ay * by + az * bz;
    const double cang = -ab * ai * bi;
    const double fij = S.ang_kt[a] * (acos(cang) - S.ang_th[a]) / sqrt(1.0 - cang * cang) * ai * bi;
    const double bb = ab * bi * bi, aa = ab * ai * ai;
    const double tx = aa * ax - bx, ty = aa * ay - by, tz = aa * az - bz;
    if (role) {
        fx += fij * tx; fy += fij * ty; fz += fij * tz;
    } else {
        fx += fij * (bb * bx - ax - tx); fy += fij * (bb * by - ay - ty); fz += fij * (bb * bz - az - tz);
    }
}

// One block covers EPI_BLOCK = 3 x 64 consecutive sites, i.e. 64 whole rigid molecules
// (rigid sites come first, three per molecule), so the block also forms those molecules'
// COM forces and torques (src/forcem.f90:1003-1024) from the forces it has just finished.
// With multiple time steps on (mts non-null), a step without the full k-sum takes the
// reciprocal forces from fmts [3][n] (surrogate plus correction) instead of the row groups.
__device__ __forceinline__ void epilogue_body(const DevSys &S, const double *__restrict__ fpart_k, int n_groups_k,
                                                        const WallLR *__restrict__ wl, int ew3dc, double crf0,
                                                        const double *__restrict__ dipz, double *__restrict__ epart,
                                                        double *__restrict__ fcom, double *__restrict__ tq, int bid, int nblk,
                                                        const MtsState *__restrict__ mts = nullptr,
                                                        const double *__restrict__ fmts = nullptr,
                                                        double *__restrict__ efi = nullptr) {
    __shared__ double sf[3][EPI_BLOCK];
    const bool surrogate = mts != nullptr && !mts->exact;
    const double *__restrict__ fpart = surrogate ? fmts : fpart_k;
    const int n_groups = surrogate ? 1 : n_groups_k;
    __shared__ double wsum[EPI_BLOCK / 64][EP_N];
    double acc[EP_N] = {0, 0, 0, 0, 0, 0};
    const int i = bid * blockDim.x + threadIdx.x;
    const int n = S.n_move;
    sf[0][threadIdx.x] = 0.0; sf[1][threadIdx.x] = 0.0; sf[2][threadIdx.x] = 0.0;
    if (i < n) {
        double fx = S.fx[i], fy = S.fy[i], fz = S.fz[i];
        const double fz_pair = fz;   // the pair kernel's part (loops A and B)
        // wall long-range tail correction, sites with |q| >= 0.7
        if (wl->on) {
            const int it = S.itype[i];
            if (fabs(S.q[i]) >= 0.7) {
                for (int jj = -1; jj <= 1; jj += 2) {
                    double aux = 0.0, sgn = (double)jj;
                    double dw = S.lwh - sgn * S.z[i];
                    if (dw < wl->rcut) {
                        aux = aux + sgn * wl->flr_w[it];
                        acc[0] += wl->ulr_w[it];
                    } else {
                        aux = aux + sgn * wl->flr_l[it];
                        acc[0] += wl->ulr_l[it];
                        double dwa = 1.0 / (dw * dw);
                        double dwa2 = dwa * dwa;
                        aux = aux + sgn * (wl->flr_lja[it] * (dwa2 * dwa2 * dwa) - wl->flr_ljc[it] * dwa2);
                        double dw3 = dw * dw * dw;
                        acc[0] += (wl->flr_lja[it] / (9.0 * (dw3 * dw3 * dw3)) - wl->flr_ljc[it] / (3.0 * dw3));
                        double t = wl->flr_bur[it];
                        double e2 = wl->flr_bua[it] * exp(-dw / t);
                        aux = aux + sgn * ((t + dw) * e2 - wl->flr_buc[it] * dwa2);
                        acc[0] += (t * (2.0 * t + dw) * e2 - wl->flr_buc[it] / (3.0 * dw3));
                    }
                    fz -= aux;
                    if (it <= S.n_ltype) acc[jj < 0 ? 1 : 2] += aux;
                }
            }
        }
        for (int g = 0; g < n_groups; g++) {
            fx += fpart[((size_t)g * 3 + 0) * n + i];
            fy += fpart[((size_t)g * 3 + 1) * n + i];
            fz += fpart[((size_t)g * 3 + 2) * n + i];
        }
        if (ew3dc) {  // crf = 2 crf0 facq dipz; fz += crf q  (src/forcem.f90:927-928)
            const double crf = 2.0 * crf0 * S.facq * *dipz;
            fz += crf * S.q[i];
        }
        // efipos of the reference: each pair's z-force on a mobile ion of type 3 or 4 in loops A
        // and B, then the ion's force after the k-sum and the slab correction
        // (src/forcem.f90:347-354,388-395,486-489,930-934); angles come later there
        if (efi) {
            const int it = S.itype[i];
            efi[i] = (i >= S.n_rigid_sites && i < S.n_liquid_sites && (it == 3 || it == 4)) ? fz_pair + fz : 0.0;
        }
        if (S.n_angles > 0)
            for (int p = S.ang_ptr[i]; p < S.ang_ptr[i + 1]; p++) {
                const int e = S.ang_inc[p];
                angle_force(S, e >> 1, e & 1, fx, fy, fz);
            }
        S.fx[i] = fx; S.fy[i] = fy; S.fz[i] = fz;
        sf[0][threadIdx.x] = fx; sf[1][threadIdx.x] = fy; sf[2][threadIdx.x] = fz;
        acc[3] = fx; acc[4] = fy; acc[5] = fz;
    }
    // block partials: a fixed butterfly per wave, then the waves in order
#pragma unroll
    for (int c = 0; c < EP_N; c++)
        for (int o = 32; o > 0; o >>= 1) acc[c] += __shfl_xor(acc[c], o, 64);
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int c = 0; c < EP_N; c++) wsum[threadIdx.x >> 6][c] = acc[c];
    __syncthreads();
    if (threadIdx.x < EP_N) {
        double v = 0.0;
        for (int w = 0; w < EPI_BLOCK / 64; w++) v += wsum[w][threadIdx.x];
        epart[bid * EP_N + threadIdx.x] = v;
    }
    // COM force and torque of the block's molecules
    const int nr = S.n_rig;
    const int m = bid * (EPI_BLOCK / 3) + threadIdx.x;
    if (threadIdx.x < EPI_BLOCK / 3 && m < nr) {
        const int j = 3 * m, lj = 3 * threadIdx.x;
        double tx = 0, ty = 0, tz = 0, gx = 0, gy = 0, gz = 0;
        for (int a = 0; a < 3; a++) {
            const double rx = S.x[j + a] - S.cx[m], ry = S.y[j + a] - S.cy[m], rz = S.z[j + a] - S.cz[m];
            const double hx = sf[0][lj + a], hy = sf[1][lj + a], hz = sf[2][lj + a];
            tx += ry * hz - rz * hy;
            ty += rz * hx - rx * hz;
            tz += rx * hy - ry * hx;
            gx += hx; gy += hy; gz += hz;
        }
        fcom[m] = gx; fcom[nr + m] = gy; fcom[2 * nr + m] = gz;
        tq[m] = tx; tq[nr + m] = ty; tq[2 * nr + m] = tz;
    }
}

__global__ __launch_bounds__(EPI_BLOCK) void k_epilogue(DevSys S, const double *__restrict__ fpart, int n_groups,
                                                        const WallLR *__restrict__ wl, int ew3dc, double crf0,
                                                        const double *__restrict__ dipz, double *__restrict__ epart,
                                                        double *__restrict__ fcom, double *__restrict__ tq,
                                                        const MtsState *__restrict__ mts, const double *__restrict__ fmts,
                                                        double *__restrict__ efi) {
    epilogue_body(S, fpart, n_groups, wl, ew3dc, crf0, dipz, epart, fcom, tq, blockIdx.x, gridDim.x, mts, fmts, efi);
}

// ngc of the reference's pair loops, binned by site-pair distance per type pair with the
// reference's operation order and no contraction.  Integer counts: exact whatever the order
// of the atomic additions.
//  - moving-frozen (every form): every site of a lead's molecule against each frozen site of
//    its row inside the cut-off (src/forcem.f90:418-425,459-466,493-495;
//    src/forcem_back.f90:405-412,443-449,472-474);
//  - moving-moving (3-D form only, count_mm): every site pair of a molecule pair i<j inside
//    the cut-off, lead pair included, at (xio + rs(ix)) - rs(jx) with rs the site offset from
//    its lead (src/forcem_back.f90:288-298,348-349,353-382).  The slab form has these counts
//    commented out (src/forcem.f90:357,399).
__global__ __launch_bounds__(256) void k_obs_ngc(DevSys S, ObsArgs o) {
    const int lane = threadIdx.x & 15;
    const int gid = blockIdx.x * 16 + (threadIdx.x >> 4);
    if (gid >= S.n_lead) return;
    // a step whose list overflowed is run again with grown lists: count it only then
    if (S.flags->overflow_mm || S.flags->overflow_mw) return;
    const int i = S.lead[gid];
    const int ti = S.itype[i], nsi_i = S.nsi[ti];
    const double xi = S.x[i], yi = S.y[i], zi = S.z[i];
    const int nt1 = S.n_types + 1;
    if (o.count_mm) {
        const int nm = min(S.cnt_mm[gid], S.cap_mm);
        const int *mrow = S.nbr_mm + (size_t)gid * S.cap_mm;
        for (int k = lane; k < nm; k += 16) {
            const int j = mrow[k];
            if (j <= i) continue;   // each pair once, from the row of its lower index (list1 < list2)
            double dx = __dsub_rn(xi, S.x[j]); dx = __dsub_rn(dx, __dmul_rn(S.lx, anint_d(__dmul_rn(dx, S.lxi))));
            double dy = __dsub_rn(yi, S.y[j]); dy = __dsub_rn(dy, __dmul_rn(S.ly, anint_d(__dmul_rn(dy, S.lyi))));
            double dz = __dsub_rn(zi, S.z[j]);
            if (S.zper_force) dz = __dsub_rn(dz, __dmul_rn(S.lz, anint_d(__dmul_rn(dz, S.lzi))));
            const double rsq0 = __dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), __dmul_rn(dz, dz));
            if (rsq0 > S.rcutsq) continue;
            const int nsi_j = S.nsi[S.itype[j]];
            const double xj = S.x[j], yj = S.y[j], zj = S.z[j];
            for (int a = 0; a < nsi_i && a < 3; a++) {
                const int ix = i + a;
                const double xa = __dadd_rn(dx, __dsub_rn(S.x[ix], xi));
                const double ya = __dadd_rn(dy, __dsub_rn(S.y[ix], yi));
                const double za = __dadd_rn(dz, __dsub_rn(S.z[ix], zi));
                const int ta = S.itype[ix];
                for (int b = 0; b < nsi_j && b < 3; b++) {
                    const int jx = j + b;
                    const double xd = __dsub_rn(xa, __dsub_rn(S.x[jx], xj));
                    const double yd = __dsub_rn(ya, __dsub_rn(S.y[jx], yj));
                    const double zd = __dsub_rn(za, __dsub_rn(S.z[jx], zj));
                    const double rsq = __dadd_rn(__dadd_rn(__dmul_rn(xd, xd), __dmul_rn(yd, yd)), __dmul_rn(zd, zd));
                    const int l = (int)__dadd_rn(__dmul_rn(sqrt(rsq), o.rdeli), 0.5);
                    const int ii = S.ijtype[ta * nt1 + S.itype[jx]];
                    if (l >= 0 && l < o.nbins && ii >= 1) atomicAdd(&o.ngc[(size_t)l * S.nijt + (ii - 1)], 1ull);
                }
            }
        }
    }
    const int nw = min(S.cnt_mw[gid], S.cap_mw);
    const int *wrow = S.nbr_mw + (size_t)gid * S.cap_mw;
    for (int k = lane; k < nw; k += 16) {
        const int j = wrow[k];
        double dx = __dsub_rn(xi, S.x[j]); dx = __dsub_rn(dx, __dmul_rn(S.lx, anint_d(__dmul_rn(dx, S.lxi))));
        double dy = __dsub_rn(yi, S.y[j]); dy = __dsub_rn(dy, __dmul_rn(S.ly, anint_d(__dmul_rn(dy, S.lyi))));
        double dz = __dsub_rn(zi, S.z[j]);
        if (S.zper_force) dz = __dsub_rn(dz, __dmul_rn(S.lz, anint_d(__dmul_rn(dz, S.lzi))));
        const double rsq0 = __dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), __dmul_rn(dz, dz));
        if (rsq0 > S.rcutsq) continue;
        const int tj = S.itype[j];
        for (int a = 0; a < nsi_i && a < 3; a++) {
            const int ix = i + a;
            const double xd = __dadd_rn(dx, __dsub_rn(S.x[ix], xi));
            const double yd = __dadd_rn(dy, __dsub_rn(S.y[ix], yi));
            const double zd = __dadd_rn(dz, __dsub_rn(S.z[ix], zi));
            const double rsq = __dadd_rn(__dadd_rn(__dmul_rn(xd, xd), __dmul_rn(yd, yd)), __dmul_rn(zd, zd));
            const double rij = sqrt(rsq);
            const int l = (int)__dadd_rn(__dmul_rn(rij, o.rdeli), 0.5);
            const int ii = S.ijtype[S.itype[ix] * nt1 + tj];
            if (l >= 0 && l < o.nbins && ii >= 1) atomicAdd(&o.ngc[(size_t)l * S.nijt + (ii - 1)], 1ull);
        }
    }
}

// efipos bins of the step's ion z-forces, in site order (one thread: a fixed summation order)
// force_intra's angle histogram (src/forcem.f90:570-625): one thread per angle term, the
// reference's cosine in its operation order without contraction (1/|a| = 1/sqrt(a.a)), bin
// l = int((cos + 1) 100) + 1 of the term's group type; integer counts, so the order of the
// additions does not matter
__global__ void k_obs_ngang(DevSys S, const int *__restrict__ ang_type, int n_atypes,
                            unsigned long long *__restrict__ hist) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= S.n_angles) return;
    const int i = S.ang_ijk[3 * a], j = S.ang_ijk[3 * a + 1], k = S.ang_ijk[3 * a + 2];
    auto img = [](double d, double L, double Li) { return __dsub_rn(d, __dmul_rn(L, anint_d(__dmul_rn(d, Li)))); };
    const double ax = img(__dsub_rn(S.x[k], S.x[j]), S.lx, S.lxi), ay = img(__dsub_rn(S.y[k], S.y[j]), S.ly, S.lyi);
    const double bx = img(__dsub_rn(S.x[j], S.x[i]), S.lx, S.lxi), by = img(__dsub_rn(S.y[j], S.y[i]), S.ly, S.lyi);
    double az = __dsub_rn(S.z[k], S.z[j]), bz = __dsub_rn(S.z[j], S.z[i]);
    if (S.zper_force) { az = img(az, S.lz, S.lzi); bz = img(bz, S.lz, S.lzi); }
    auto dot = [](double u0, double u1, double u2, double v0, double v1, double v2) {
        return __dadd_rn(__dadd_rn(__dmul_rn(u0, v0), __dmul_rn(u1, v1)), __dmul_rn(u2, v2));
    };
    const double ai = __ddiv_rn(1.0, __dsqrt_rn(dot(ax, ay, az, ax, ay, az)));
    const double bi = __ddiv_rn(1.0, __dsqrt_rn(dot(bx, by, bz, bx, by, bz)));
    const double cang = -__dmul_rn(__dmul_rn(dot(ax, ay, az, bx, by, bz), ai), bi);
    const double u = __dmul_rn(__dadd_rn(cang, 1.0), 100.0);
    const int l = (u > -1.0 && u < (double)MDF_NGANG_BINS) ? (int)u + 1 : -1;   // the reference's int(): toward zero
    const int it = ang_type[a];
    if (l >= 0 && l < MDF_NGANG_BINS && it >= 1 && it <= n_atypes)
        atomicAdd(&hist[(size_t)l * n_atypes + (it - 1)], 1ull);
}

__global__ void k_obs_efipos(DevSys S, ObsArgs o) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (S.flags->overflow_mm || S.flags->overflow_mw) return;
    for (int i = S.n_rigid_sites; i < S.n_liquid_sites; i++) {
        const int it = S.itype[i];
        if (it != 3 && it != 4) continue;
        const int l = (int)__dadd_rn(__dmul_rn(__dadd_rn(S.z[i], o.lwh), o.rdeli), 0.5);
        if (l >= 0 && l < o.nbins) o.efipos[2 * l + (it - 3)] += o.efi[i];
    }
}

// column sums of the per-lead partials, split by the lead's type slot
__device__ __forceinline__ void colsum_body(const DevSys &S, const int *__restrict__ lead_slot, int per_block,
                                                double *__restrict__ colpart, int bx, int by, int nbx, int nby) {
    __shared__ double red[MAX_LEAD_TYPES][256];
    const int col = bx, blk = by;
    const size_t L = (size_t)S.n_lead;
    double acc[MAX_LEAD_TYPES] = {0, 0, 0, 0};
    const int b = blk * per_block, e = min(S.n_lead, b + per_block);
    for (int l = b + threadIdx.x; l < e; l += blockDim.x) {
        double v = S.part[col * L + l];
        int s = lead_slot[l];
#pragma unroll
        for (int t = 0; t < MAX_LEAD_TYPES; t++) acc[t] += (s == t) ? v : 0.0;
    }
#pragma unroll
    for (int t = 0; t < MAX_LEAD_TYPES; t++) red[t][threadIdx.x] = acc[t];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s)
#pragma unroll
            for (int t = 0; t < MAX_LEAD_TYPES; t++) red[t][threadIdx.x] += red[t][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x < MAX_LEAD_TYPES)
        colpart[((size_t)col * nby + blk) * MAX_LEAD_TYPES + threadIdx.x] = red[threadIdx.x][0];
}

__global__ __launch_bounds__(256) void k_colsum(DevSys S, const int *__restrict__ lead_slot, int per_block,
                                                double *__restrict__ colpart) {
    colsum_body(S, lead_slot, per_block, colpart, blockIdx.x, blockIdx.y, gridDim.x, gridDim.y);
}


__device__ __forceinline__ void final_body(const DevSys &S, FinalArgs A, const double *__restrict__ colpart,
                                               const double *__restrict__ kpart, const double *__restrict__ eppart,
                                               const double *__restrict__ dipz, double *__restrict__ e_vdw,
                                               double *__restrict__ e_coul, Scalars *__restrict__ out, int bid, int nblk) {
    __shared__ double colsum[NCOL][MAX_LEAD_TYPES];
    __shared__ double ksum[8];
    __shared__ double epsum[EP_N];
    constexpr int NIJT1 = (MAX_TYPES + 1) * (MAX_TYPES + 1);
    __shared__ int sij[NIJT1];
    const int t = threadIdx.x;
    // every global read of the kernel is issued up front: the type-pair table and the dipole
    // here, the partial sums below; one memory latency instead of one per phase
    const int nt1 = S.n_types + 1;
    for (int k = t; k < nt1 * nt1; k += blockDim.x) sij[k] = S.ijtype[k];
    const double dz = t == 0 ? *dipz : 0.0;
    const bool rb = S.flags->rebuild != 0;
    // columns x slots
    for (int idx = t; idx < NCOL * MAX_LEAD_TYPES; idx += blockDim.x) {
        int col = idx / MAX_LEAD_TYPES, s = idx % MAX_LEAD_TYPES;
        double acc = 0.0;
#pragma unroll 8
        for (int b = 0; b < A.n_colblk; b++) acc += colpart[((size_t)col * A.n_colblk + b) * MAX_LEAD_TYPES + s];
        colsum[col][s] = acc;
    }
    // k partials and epilogue partials: 32 lanes per quantity, then a fixed-order shuffle tree
    {
        __shared__ double tmp[RE_N + EP_N][32];
        const int q = t >> 5, l = t & 31;  // 8 quantities per pass of 256 threads
        for (int pass = 0; pass < 2; pass++) {
            const int qq = q + 8 * pass;
            double acc = 0.0;
            if (qq < RE_N) {
#pragma unroll 8
                for (int b = l; b < A.n_kblk; b += 32) acc += kpart[b * RE_N + qq];
            } else if (qq < RE_N + EP_N) {
#pragma unroll 8
                for (int b = l; b < A.n_epblk; b += 32) acc += eppart[b * EP_N + (qq - RE_N)];
            }
            if (qq < RE_N + EP_N) tmp[qq][l] = acc;
        }
        __syncthreads();
        if (t < RE_N + EP_N) {
            double acc = 0.0;
            for (int l2 = 0; l2 < 32; l2++) acc += tmp[t][l2];
            if (t < RE_N) ksum[t] = acc;
            else epsum[t - RE_N] = acc;
        }
    }
    // listed pair counts (ipack, wpack) after a rebuild: the per-row counts, four at a time
    // with many loads in flight (integer sums: the same in any order)
    __shared__ long long npair[2][256];
    if (rb) {
        long long a = 0, b = 0;
        const int n4 = S.n_lead >> 2;
        const int4 *lo4 = reinterpret_cast<const int4 *>(S.cnt_lo);
        const int4 *mw4 = reinterpret_cast<const int4 *>(S.cnt_mw);
#pragma unroll 8
        for (int l = t; l < n4; l += blockDim.x) {
            const int4 u = lo4[l], w = mw4[l];
            a += (long long)u.x + u.y + u.z + u.w;
            b += (long long)w.x + w.y + w.z + w.w;
        }
        for (int l = 4 * n4 + t; l < S.n_lead; l += blockDim.x) { a += S.cnt_lo[l]; b += S.cnt_mw[l]; }
        npair[0][t] = a; npair[1][t] = b;
    }
    __syncthreads();
    if (rb && t == 0) {
        long long a = 0, b = 0;
        for (int k = 0; k < (int)blockDim.x; k++) { a += npair[0][k]; b += npair[1][k]; }
        S.flags->n_pairs_mm = (int)a;
        S.flags->n_pairs_mw = (int)b;
    }
    // energies per pair type: thread k adds the (slot, class) columns of type k in (slot,
    // class) order, in LDS, then writes them once
    constexpr int NIJM = MAX_TYPES * (MAX_TYPES + 1) / 2 + 1;
    __shared__ double sev[NIJM], sec[NIJM];
    for (int k = t; k <= S.nijt; k += blockDim.x) {
        double ev = 0.0, ec = 0.0;
        for (int s = 0; s < A.n_slots; s++) {
            const int ti = A.slot_type[s];
            for (int c = 0; c < NCLS; c++) {
                const int tc = A.cls_type[c];
                if (tc <= 0 || sij[ti * nt1 + tc] != k) continue;
                ev += colsum[COL_EV + c][s];
                ec += colsum[COL_EC + c][s];
            }
        }
        sev[k] = ev; sec[k] = ec;
        e_vdw[k] = ev; e_coul[k] = ec;
    }
    __syncthreads();
    if (t != 0) return;
    Scalars o;
    o.e_recip = 0.5 * ksum[0];
    o.e_recip_ll = 0.5 * ksum[1];
    o.e_recip_ww = A.e_recip_ww;
    o.e_recip_wl = o.e_recip - o.e_recip_ww - o.e_recip_ll;
    o.e_self = A.e_self;
    o.e_slab = A.ew3dc ? -A.crf0 * S.facq * (dz * dz) : 0.0;
    o.e_wall_lr = epsum[0];
    double sum = 0.0;
    for (int k = 1; k <= S.nijt; k++) sum += sev[k];
    for (int k = 1; k <= S.nijt; k++) sum += sec[k];
    sum += o.e_recip;
    double uconf = sum + o.e_self + o.e_slab;
    o.e_conf = uconf - o.e_self;
    o.e_avex = sum + o.e_wall_lr + o.e_slab;
    // wall forces fws(a, sgn): pair part (columns) + tail correction (z only)
    for (int a = 0; a < 3; a++) {
        double m = 0.0, p = 0.0;
        for (int s = 0; s < A.n_slots; s++) { m += colsum[COL_FWS + a][s]; p += colsum[COL_FWS + 3 + a][s]; }
        o.f_wall[a * 3 + 0] = m;
        o.f_wall[a * 3 + 2] = p;
        o.f_wall[a * 3 + 1] = epsum[3 + a];
    }
    o.f_wall[6] += epsum[1];
    o.f_wall[8] += epsum[2];
    o.press = (o.f_wall[8] - o.f_wall[6]) / (2.0 * A.lx * A.ly) * 1000.0 / (A.na_scale) * A.bar;
    // virial: real-space columns (symmetric) + reciprocal
    double vr[9];
    for (int c = 0; c < 9; c++) {
        double acc = 0.0;
        for (int s = 0; s < A.n_slots; s++) acc += colsum[COL_VIR + c][s];
        vr[c] = acc;
    }
    vr[0] += ksum[2]; vr[4] += ksum[3]; vr[8] += ksum[4];
    vr[1] += ksum[5]; vr[3] += ksum[5];
    vr[2] += ksum[6]; vr[6] += ksum[6];
    vr[5] += ksum[7]; vr[7] += ksum[7];
    for (int c = 0; c < 9; c++) o.virial[c] = vr[c];
    o.nonfinite = !(isfinite(o.e_conf) && isfinite(o.press));
    o.pad = 0;
    *out = o;
}

// the listed pair counts after the list build at creation (no force step follows it): one
// block of 256 threads, in fixed order
__global__ __launch_bounds__(256) void k_pair_totals(DevSys S) {
    __shared__ long long npair[2][256];
    const int t = threadIdx.x;
    long long a = 0, b = 0;
    for (int l = t; l < S.n_lead; l += 256) { a += S.cnt_lo[l]; b += S.cnt_mw[l]; }
    npair[0][t] = a; npair[1][t] = b;
    __syncthreads();
    if (t == 0) {
        a = 0; b = 0;
        for (int k = 0; k < 256; k++) { a += npair[0][k]; b += npair[1][k]; }
        S.flags->n_pairs_mm = (int)a;
        S.flags->n_pairs_mw = (int)b;
    }
}

__global__ __launch_bounds__(256) void k_final(DevSys S, FinalArgs A, const double *__restrict__ colpart,
                                               const double *__restrict__ kpart, const double *__restrict__ eppart,
                                               const double *__restrict__ dipz, double *__restrict__ e_vdw,
                                               double *__restrict__ e_coul, Scalars *__restrict__ out) {
    final_body(S, A, colpart, kpart, eppart, dipz, e_vdw, e_coul, out, blockIdx.x, gridDim.x);
}

}  // namespace mdf

namespace mdf {
// batched forms
__global__ __launch_bounds__(EPI_BLOCK) void k_epilogue_b(const RepArgs *__restrict__ RA,
                                                          const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    const RepArgs &r = RA[m.x];
    epilogue_body(r.S, r.fpart, r.n_groups, r.wl, r.ew3dc, r.crf0, r.dipz, r.eppart, r.fcom, r.tq, m.y, r.n_epblk);
}
__global__ __launch_bounds__(256) void k_colsum_b(const RepArgs *__restrict__ RA, const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    const RepArgs &r = RA[m.x];
    colsum_body(r.S, r.lead_slot, r.colper, r.colpart, m.y % NCOL, m.y / NCOL, NCOL, r.n_colblk);
}
__global__ __launch_bounds__(256) void k_final_b(const RepArgs *__restrict__ RA) {
    const RepArgs &r = RA[blockIdx.x];
    final_body(r.S, r.A, r.colpart, r.kpart, r.eppart, r.dipz, r.e_vdw, r.e_coul, r.scal, 0, 1);
}
}  // namespace mdf
