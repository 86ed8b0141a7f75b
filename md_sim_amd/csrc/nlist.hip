// Verlet pair lists for the force engine: skin test, cell binning, list build.
//
// Reference: nnlist_init / nnlist_make (src/nnlist.f90:14-164) enumerate every lead pair
// i<j once and keep those with dx^2+dy^2+dz^2 < rskinsq (x,y minimum image, z open) at
// each rebuild; the rebuild fires when 4*max|displacement|^2 > skin^2 for COMs and
// non-rigid sites (src/forcem.f90:245-250).  Here the candidates come from a cell grid
// whose cells are at least sqrt(rskinsq) wide, and the same predicate - same operation
// order, no FMA contraction - decides membership, so the pair set is bit-identical.
// Lists are ELL rows per lead: moving partners in both directions (full list) and frozen
// partners.  All kernels after the skin test return at once when no rebuild is due, so the
// step's launch sequence is fixed.
#include "mdf_internal.h"

namespace mdf {

// max squared displacement since the last build (src/forcem.f90:245-249) and lead z range;
// per-block partial of the moving dipole sum(q z) (src/forcem.f90:924) into dip_part
__device__ __forceinline__ void skin_body(const DevSys &S, double *dip_part, int bid, int nblk) {
    __shared__ double red[256];
    double dmax = 0.0, dip = 0.0;
    double zlo = 1e300, zhi = -1e300;
    const int n_nonrig = S.n_move - S.n_rigid_sites;
    const int total = S.n_rig + n_nonrig;
    for (int t = bid * blockDim.x + threadIdx.x; t < max(total, max(S.n_move, S.n_lead));
         t += nblk * blockDim.x) {
        if (t < S.n_rig) {
            double a = __dsub_rn(S.ocx[t], S.cx[t]), b = __dsub_rn(S.ocy[t], S.cy[t]), c = __dsub_rn(S.ocz[t], S.cz[t]);
            double d = __dadd_rn(__dadd_rn(__dmul_rn(a, a), __dmul_rn(b, b)), __dmul_rn(c, c));
            dmax = fmax(dmax, d);
        } else if (t < total) {
            int i = S.n_rigid_sites + (t - S.n_rig);
            double a = __dsub_rn(S.ox[i], S.x[i]), b = __dsub_rn(S.oy[i], S.y[i]), c = __dsub_rn(S.oz[i], S.z[i]);
            double d = __dadd_rn(__dadd_rn(__dmul_rn(a, a), __dmul_rn(b, b)), __dmul_rn(c, c));
            dmax = fmax(dmax, d);
        }
        if (t < S.n_move) dip += S.q[t] * S.z[t];
        if (t < S.n_lead) {
            const int i = S.lead[t];
            const double xi = S.x[i], yi = S.y[i], zz = S.z[i];
            zlo = fmin(zlo, zz);
            zhi = fmax(zhi, zz);
            // the pair kernel's record of this lead's molecule: lead position and charge, the
            // other sites as offsets from the lead (the reference's rsx, src/move.f90:157-161)
            const int nsi = S.pmeta[i] >> 4;
            double r[PREC];
            r[0] = xi; r[1] = yi; r[2] = zz; r[3] = S.q[i];
#pragma unroll
            for (int a = 1; a < 3; a++) {
                const bool ok = a < nsi;
                r[4 * a + 0] = ok ? S.x[i + a] - xi : 0.0;
                r[4 * a + 1] = ok ? S.y[i + a] - yi : 0.0;
                r[4 * a + 2] = ok ? S.z[i + a] - zz : 0.0;
                r[4 * a + 3] = ok ? S.q[i + a] : 0.0;
            }
            double4 *dst = reinterpret_cast<double4 *>(S.prec + (size_t)i * PREC);
            dst[0] = make_double4(r[0], r[1], r[2], r[3]);
            dst[1] = make_double4(r[4], r[5], r[6], r[7]);
            dst[2] = make_double4(r[8], r[9], r[10], r[11]);
        }
    }
    // block partials, no atomics: part[0][b] dipole (fixed tree order), part[1][b] max
    // displacement, part[2][b] / part[3][b] lead z range (order independent)
    __shared__ double ext[3][4];
    for (int o = 32; o > 0; o >>= 1) {
        dmax = fmax(dmax, __shfl_xor(dmax, o, 64));
        zlo = fmin(zlo, __shfl_xor(zlo, o, 64));
        zhi = fmax(zhi, __shfl_xor(zhi, o, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        ext[0][threadIdx.x >> 6] = dmax; ext[1][threadIdx.x >> 6] = zlo; ext[2][threadIdx.x >> 6] = zhi;
    }
    red[threadIdx.x] = dip;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const int nb = nblk;
        dip_part[bid] = red[0];
        dip_part[nb + bid] = fmax(fmax(ext[0][0], ext[0][1]), fmax(ext[0][2], ext[0][3]));
        dip_part[2 * nb + bid] = fmin(fmin(ext[1][0], ext[1][1]), fmin(ext[1][2], ext[1][3]));
        dip_part[3 * nb + bid] = fmax(fmax(ext[2][0], ext[2][1]), fmax(ext[2][2], ext[2][3]));
    }
}

__global__ __launch_bounds__(256) void k_skin(DevSys S, double *dip_part) {
    skin_body(S, dip_part, blockIdx.x, gridDim.x);
}

// rebuild decision (src/forcem.f90:250) and z grid of this build; dipole total.  One block,
// over k_skin's per-block partials.  It counts the step and publishes a pending list
// overflow (left by an earlier step's build) to the host-mapped mirror; the overflow flags
// themselves stay set until the host has read them.
__device__ __forceinline__ void decide_body(const DevSys &S, const double *dip_part, int n_dip, double *dipz_out,
                                                int force_rebuild, int bid, int nblk) {
    __shared__ double red2[256];
    __shared__ double ext[3][4];
    double acc = 0.0, dmax = 0.0, zlo = 1e300, zhi = -1e300;
    for (int b = threadIdx.x; b < n_dip; b += blockDim.x) {
        acc += dip_part[b];
        dmax = fmax(dmax, dip_part[n_dip + b]);
        zlo = fmin(zlo, dip_part[2 * n_dip + b]);
        zhi = fmax(zhi, dip_part[3 * n_dip + b]);
    }
    for (int o = 32; o > 0; o >>= 1) {
        dmax = fmax(dmax, __shfl_xor(dmax, o, 64));
        zlo = fmin(zlo, __shfl_xor(zlo, o, 64));
        zhi = fmax(zhi, __shfl_xor(zhi, o, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        ext[0][threadIdx.x >> 6] = dmax; ext[1][threadIdx.x >> 6] = zlo; ext[2][threadIdx.x >> 6] = zhi;
    }
    red2[threadIdx.x] = acc;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) red2[threadIdx.x] += red2[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x != 0) return;
    *dipz_out = red2[0];
    dmax = fmax(fmax(ext[0][0], ext[0][1]), fmax(ext[0][2], ext[0][3]));
    zlo = fmin(fmin(ext[1][0], ext[1][1]), fmin(ext[1][2], ext[1][3]));
    zhi = fmax(fmax(ext[2][0], ext[2][1]), fmax(ext[2][2], ext[2][3]));
    DevFlags *f = S.flags;
    f->steps += 1;
    if (S.flag_mirror) S.flag_mirror[0] = (f->overflow_mm | f->overflow_mw) != 0 ? 1 : 0;
    f->rebuild = (force_rebuild || __dmul_rn(4.0, dmax) > S.skin2) ? 1 : 0;
    if (f->rebuild) {
        // pair counts describe the standing list (ipack/wpack): restart them for this build
        f->n_builds += 1;
        f->n_pairs_mm = 0;
        f->n_pairs_mw = 0;
    }
    if (S.g.zper) {
        f->ncz = S.g.ncz_fixed;
        f->z0 = 0.0;
        f->csz = S.g.csz_fixed;
    } else {
        if (!(zlo <= zhi)) { zlo = 0.0; zhi = 0.0; }
        double rs = sqrt(S.rskinsq) * (1.0 + 1e-9);
        double range = zhi - zlo;
        int ncz = (int)floor(range / rs);
        if (ncz < 1) ncz = 1;
        if (ncz > S.g.ncz_max) ncz = S.g.ncz_max;
        f->ncz = ncz;
        f->z0 = zlo;
        f->csz = range > 0.0 ? range / ncz : 1.0;
    }
}

__global__ __launch_bounds__(256) void k_decide(DevSys S, const double *dip_part, int n_dip, double *dipz_out,
                                                int force_rebuild) {
    decide_body(S, dip_part, n_dip, dipz_out, force_rebuild, blockIdx.x, gridDim.x);
}

__device__ __forceinline__ int wrap_cell(double v, double L, double Li, int n) {
    double w = v - L * floor(v * Li);
    int c = (int)(w * (double)n * Li);
    return c < 0 ? 0 : (c >= n ? n - 1 : c);
}

__device__ __forceinline__ int zcell_lin(double z, double z0, double cs, int n) {
    int c = (int)floor((z - z0) / cs);
    return c < 0 ? 0 : (c >= n ? n - 1 : c);
}

__device__ __forceinline__ int lead_cell(const DevSys &S, int i, int ncz, double z0, double csz) {
    int cx = wrap_cell(S.x[i], S.lx, S.lxi, S.g.ncx);
    int cy = wrap_cell(S.y[i], S.ly, S.lyi, S.g.ncy);
    int cz = S.g.zper ? wrap_cell(S.z[i], S.lz, S.lzi, ncz) : zcell_lin(S.z[i], z0, csz, ncz);
    return (cz * S.g.ncy + cy) * S.g.ncx + cx;
}

// exclusive scan of n cell counts by one block (any size); with zero_cnt the counts are
// cleared behind it, ready for the next build's counting
__device__ void scan_cells(int *cnt, int *start, int *cursor, int n, bool zero_cnt) {
    __shared__ int part[1024];
    const int nt = blockDim.x;
    const int per = (n + nt - 1) / nt;
    const int b = threadIdx.x * per, e = min(n, b + per);
    int s = 0;
    for (int c = b; c < e; c++) s += cnt[c];
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int t = 0; t < nt; t++) { int v = part[t]; part[t] = acc; acc += v; }
    }
    __syncthreads();
    int acc = part[threadIdx.x];
    for (int c = b; c < e; c++) {
        const int v = cnt[c];
        start[c] = acc; cursor[c] = acc; acc += v;
        if (zero_cnt) cnt[c] = 0;
    }
    if (threadIdx.x == nt - 1) start[n] = acc;
}

// cell of every lead and the cell counts
__device__ __forceinline__ void cell_count_body(const DevSys &S, int *lead_cell_out, int bid, int nblk) {
    const DevFlags *f = S.flags;
    if (!f->rebuild) return;
    const int l = bid * blockDim.x + threadIdx.x;
    if (l >= S.n_lead) return;
    const int c = lead_cell(S, S.lead[l], f->ncz, f->z0, f->csz);
    lead_cell_out[l] = c;
    atomicAdd(&S.cell_cnt[c], 1);
}

__global__ __launch_bounds__(256) void k_cell_count(DevSys S, int *lead_cell_out) {
    cell_count_body(S, lead_cell_out, blockIdx.x, gridDim.x);
}

// cell starts of this build; the counts are cleared behind the scan for the next build
__device__ __forceinline__ void cell_scan_build_body(const DevSys &S, int bid, int nblk) {
    const DevFlags *f = S.flags;
    if (!f->rebuild) return;
    scan_cells(S.cell_cnt, S.cell_start, S.cell_cursor, S.g.ncx * S.g.ncy * f->ncz, true);
}

__global__ __launch_bounds__(1024) void k_cell_scan_build(DevSys S) {
    cell_scan_build_body(S, blockIdx.x, gridDim.x);
}

// exclusive scan of cell counts, one block (frozen-site cells at creation)
__global__ __launch_bounds__(1024) void k_cell_scan(int *cnt, int *start, int *cursor, int n) {
    scan_cells(cnt, start, cursor, n, false);
}

__device__ __forceinline__ void cell_fill_body(const int *item_cell, int n_items, int *cursor, int *items, const DevFlags *f,
                            int use_flag, int bid, int nblk) {
    if (use_flag && !f->rebuild) return;
    int l = bid * blockDim.x + threadIdx.x;
    if (l >= n_items) return;
    int pos = atomicAdd(&cursor[item_cell[l]], 1);
    items[pos] = l;
}

__global__ void k_cell_fill(const int *item_cell, int n_items, int *cursor, int *items, const DevFlags *f,
                            int use_flag) {
    cell_fill_body(item_cell, n_items, cursor, items, f, use_flag, blockIdx.x, gridDim.x);
}

// ascending order inside every cell (rank sort, one wave per cell), so list rows are
// deterministic; items_out gets the sorted copy
__device__ __forceinline__ void cell_sort_body(const int *start, const int *items_in, int *items_out,
                                                  const DevFlags *f, int use_flag, int ncells_xy, int ncz_given, int bid, int nblk) {
    if (use_flag && !f->rebuild) return;
    int ncz = use_flag ? f->ncz : ncz_given;
    for (int c = bid; c < ncells_xy * ncz; c += nblk) {
        int b = start[c], n = start[c + 1] - b;
        for (int a = threadIdx.x; a < n; a += 64) {
            int v = items_in[b + a];
            int rank = 0;
            for (int k = 0; k < n; k++) rank += (items_in[b + k] < v);
            items_out[b + rank] = v;
        }
    }
}

__global__ __launch_bounds__(64) void k_cell_sort(const int *start, const int *items_in, int *items_out,
                                                  const DevFlags *f, int use_flag, int ncells_xy, int ncz_given) {
    cell_sort_body(start, items_in, items_out, f, use_flag, ncells_xy, ncz_given, blockIdx.x, gridDim.x);
}

__device__ __forceinline__ bool constrained(const DevSys &S, int i, int j) {
    if (i >= S.n_move) return false;
    for (int k = S.con_ptr[i]; k < S.con_ptr[i + 1]; k++)
        if (S.con_idx[k] == j) return true;
    return false;
}

// list rows: one wave per lead.  The wave walks the lead's neighbour cells in a fixed
// order, tests 64 candidates at a time and appends the hits with a ballot prefix count,
// so every row holds its partners in exactly the order a serial scan would give.
__device__ __forceinline__ void build_body(const DevSys &S, int bid, int nblk) {
    const DevFlags *f = S.flags;
    if (!f->rebuild) return;
    const int lane = threadIdx.x & 63;
    const int l = bid * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (l >= S.n_lead) return;   // whole wave leaves together
    const int i = S.lead[l];
    const int ncx = S.g.ncx, ncy = S.g.ncy;
    const int ncz = f->ncz;
    const double xi = S.x[i], yi = S.y[i], zi = S.z[i];
    const int cx = wrap_cell(xi, S.lx, S.lxi, ncx);
    const int cy = wrap_cell(yi, S.ly, S.lyi, ncy);
    const int cz = S.g.zper ? wrap_cell(zi, S.lz, S.lzi, ncz) : zcell_lin(zi, f->z0, f->csz, ncz);
    const int nxo = ncx >= 3 ? 3 : ncx, nyo = ncy >= 3 ? 3 : ncy;
    const int nzo = S.g.zper ? (ncz >= 3 ? 3 : ncz) : 3;
    int *row = S.nbr_mm + (size_t)l * S.cap_mm;
    const bool i_surf = i >= S.n_liquid_sites;
    int cnt = 0, lower = 0;
    for (int oz = 0; oz < nzo; oz++) {
        int z2;
        if (S.g.zper) { z2 = (nzo == 3) ? (cz + oz - 1 + ncz) % ncz : oz; }
        else { z2 = cz + oz - 1; if (z2 < 0 || z2 >= ncz) continue; }
        for (int oy = 0; oy < nyo; oy++) {
            const int y2 = (nyo == 3) ? (cy + oy - 1 + ncy) % ncy : oy;
            for (int ox = 0; ox < nxo; ox++) {
                const int x2 = (nxo == 3) ? (cx + ox - 1 + ncx) % ncx : ox;
                const int c = (z2 * ncy + y2) * ncx + x2;
                const int b = S.cell_start[c], e = S.cell_start[c + 1];
                for (int p0 = b; p0 < e; p0 += 64) {
                    const int p = p0 + lane;
                    bool hit = false;
                    int j = 0;
                    if (p < e) {
                        const int m = S.cell_items[p];
                        j = S.lead[m];
                        if (m != l && !(i_surf && j >= S.n_liquid_sites && constrained(S, i, j))) {
                            const int a = min(i, j), bb = max(i, j);
                            hit = list_d2(S, a, bb) < S.rskinsq;
                        }
                    }
                    const unsigned long long msk = __ballot(hit);
                    if (hit) {
                        const int pos = cnt + __popcll(msk & ((1ull << lane) - 1ull));
                        if (pos < S.cap_mm) row[pos] = j;
                        lower += (j > i);
                    }
                    cnt += __popcll(msk);
                }
            }
        }
    }
    // wave totals
    for (int o = 32; o > 0; o >>= 1) lower += __shfl_xor(lower, o, 64);
    if (lane == 0) {
        S.cnt_mm[l] = cnt;
        S.cnt_lo[l] = lower;   // i < j pairs of this row; summed after the build
        if (cnt > S.cap_mm) {
            atomicMax(&S.flags->overflow_mm, cnt);
            S.flags->overflow_step = S.flags->steps;   // the same value from every writer
        }
    }
    // frozen partners: same x,y grid, frozen z grid built once
    int cntw = 0;
    if (S.n_wall > 0) {
        int *wrow = S.nbr_mw + (size_t)l * S.cap_mw;
        int wcz;
        if (S.g.zper) {
            wcz = wrap_cell(zi, S.lz, S.lzi, S.wncz);
        } else {
            double czd = floor((zi - S.wz0) / S.wcsz);
            czd = fmin(fmax(czd, -2.0), (double)S.wncz + 1.0);  // far from the crystal: no cell in reach
            wcz = (int)czd;
        }
        const int nwz = S.g.zper ? min(3, S.wncz) : 3;
        for (int oz = 0; oz < nwz; oz++) {
            int z2;
            if (S.g.zper) z2 = (S.wncz >= 3) ? (wcz + oz - 1 + S.wncz) % S.wncz : oz;
            else { z2 = wcz + oz - 1; if (z2 < 0 || z2 >= S.wncz) continue; }
            for (int oy = 0; oy < nyo; oy++) {
                const int y2 = (nyo == 3) ? (cy + oy - 1 + ncy) % ncy : oy;
                for (int ox = 0; ox < nxo; ox++) {
                    const int x2 = (nxo == 3) ? (cx + ox - 1 + ncx) % ncx : ox;
                    const int c = (z2 * ncy + y2) * ncx + x2;
                    const int b = S.wcell_start[c], e = S.wcell_start[c + 1];
                    for (int p0 = b; p0 < e; p0 += 64) {
                        const int p = p0 + lane;
                        bool hit = false;
                        int j = 0;
                        if (p < e) {
                            j = S.n_move + S.wcell_items[p];
                            if (!(i_surf && constrained(S, i, j))) hit = list_d2(S, i, j) < S.rskinsq;
                        }
                        const unsigned long long msk = __ballot(hit);
                        if (hit) {
                            const int pos = cntw + __popcll(msk & ((1ull << lane) - 1ull));
                            if (pos < S.cap_mw) wrow[pos] = j;
                        }
                        cntw += __popcll(msk);
                    }
                }
            }
        }
    }
    if (lane == 0) {
        S.cnt_mw[l] = cntw;
        if (cntw > S.cap_mw) {
            atomicMax(&S.flags->overflow_mw, cntw);
            S.flags->overflow_step = S.flags->steps;
        }
    }
}

__global__ __launch_bounds__(256) void k_build(DevSys S) {
    build_body(S, blockIdx.x, gridDim.x);
}

// snapshot for the next skin test (src/nnlist.f90:99-119)
// reference positions for the next skin tests (src/nnlist.f90:99-119)
__device__ __forceinline__ void snapshot_body(const DevSys &S, int bid, int nblk) {
    if (!S.flags->rebuild) return;
    const int t = bid * blockDim.x + threadIdx.x;
    if (t < S.n_rig) { S.ocx[t] = S.cx[t]; S.ocy[t] = S.cy[t]; S.ocz[t] = S.cz[t]; }
    const int i = S.n_rigid_sites + t;
    if (i < S.n_move) { S.ox[i] = S.x[i]; S.oy[i] = S.y[i]; S.oz[i] = S.z[i]; }
}

__global__ void k_snapshot(DevSys S) {
    snapshot_body(S, blockIdx.x, gridDim.x);
}

// frozen-site cells, built once at creation
__global__ void k_wall_cells(DevSys S, int *wcell_of, int ncz, double z0, double csz) {
    int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w >= S.n_wall) return;
    int i = S.n_move + w;
    int cx = wrap_cell(S.x[i], S.lx, S.lxi, S.g.ncx);
    int cy = wrap_cell(S.y[i], S.ly, S.lyi, S.g.ncy);
    int cz = S.g.zper ? wrap_cell(S.z[i], S.lz, S.lzi, ncz) : zcell_lin(S.z[i], z0, csz, ncz);
    wcell_of[w] = (cz * S.g.ncy + cy) * S.g.ncx + cx;
}

__global__ void k_count_cells(const int *item_cell, int n_items, int *cnt) {
    int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l < n_items) atomicAdd(&cnt[item_cell[l]], 1);
}

__global__ void k_zero_int(int *p, int n) {
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) p[t] = 0;
}

}  // namespace mdf

namespace mdf {
// ---- batched forms: one launch over every system of a batch ----------------------------
__global__ __launch_bounds__(256) void k_skin_b(const RepArgs *__restrict__ RA, const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    const RepArgs &r = RA[m.x];
    skin_body(r.S, r.dip_part, m.y, r.skin_blocks);
}
__global__ __launch_bounds__(256) void k_decide_b(const RepArgs *__restrict__ RA) {
    const RepArgs &r = RA[blockIdx.x];
    decide_body(r.S, r.dip_part, r.skin_blocks, r.dipz, r.force_rebuild, 0, 1);
}
__global__ __launch_bounds__(256) void k_cell_count_b(const RepArgs *__restrict__ RA, const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    const RepArgs &r = RA[m.x];
    cell_count_body(r.S, r.lead_cell, m.y, r.lead_blocks);
}
__global__ __launch_bounds__(1024) void k_cell_scan_build_b(const RepArgs *__restrict__ RA) {
    cell_scan_build_body(RA[blockIdx.x].S, 0, 1);
}
__global__ __launch_bounds__(256) void k_cell_fill_b(const RepArgs *__restrict__ RA, const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    const RepArgs &r = RA[m.x];
    cell_fill_body(r.lead_cell, r.S.n_lead, r.S.cell_cursor, r.cell_items2, r.S.flags, 1, m.y, r.lead_blocks);
}
__global__ __launch_bounds__(64) void k_cell_sort_b(const RepArgs *__restrict__ RA, const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    const RepArgs &r = RA[m.x];
    cell_sort_body(r.S.cell_start, r.cell_items2, r.S.cell_items, r.S.flags, 1, r.ncells_xy, 0, m.y, r.sort_blocks);
}
__global__ __launch_bounds__(256) void k_build_b(const RepArgs *__restrict__ RA, const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    build_body(RA[m.x].S, m.y, RA[m.x].build_blocks);
}
__global__ __launch_bounds__(256) void k_snapshot_b(const RepArgs *__restrict__ RA, const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    snapshot_body(RA[m.x].S, m.y, RA[m.x].snap_blocks);
}
// every system's flags side by side, for one copy to the host
__global__ void k_flags_b(const RepArgs *__restrict__ RA, int n, DevFlags *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = *RA[i].S.flags;
}
}  // namespace mdf
