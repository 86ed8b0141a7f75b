// Verlet pair lists for the force engine: skin test, cell binning, list build.
//
// Reference: nnlist_init / nnlist_make (src/nnlist.f90:14-164) enumerate every lead pair
// i<j once and keep those with dx^2+dy^2+dz^2 < rskinsq (x,y minimum image, z open) at
// each rebuild; the rebuild fires when 4*max|displacement|^2 > skin^2 for COMs and
// non-rigid sites (src/forcem.f90:245-250).  Here the candidates come from a cell grid
// whose cells are at least sqrt(rskinsq) wide, and the same predicate - same operation
// order, no FMA contraction - decides membership, so the pair set is bit-identical.
// Lists are ELL rows per lead: moving partners in both directions (full list, each row in
// ascending partner index) and frozen partners (in the fixed order of the static frozen-site
// cells).  All kernels after the skin test return at once when no rebuild is due, so the
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

// exclusive scan of n cell counts by one block (any size, a multiple of 64 threads); with
// zero_cnt the counts are cleared behind it, ready for the next build's counting
__device__ void scan_cells(int *cnt, int *start, int *cursor, int n, bool zero_cnt) {
    __shared__ int wsum[16];
    const int nt = blockDim.x, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int per = (n + nt - 1) / nt;
    const int b = threadIdx.x * per, e = min(n, b + per);
    int s = 0;
    for (int c = b; c < e; c++) s += cnt[c];
    // exclusive prefix of the per-thread sums: inclusive scan in the wave, then over waves
    int inc = s;
    for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(inc, o, 64);
        if (lane >= o) inc += v;
    }
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    int base = 0;
    for (int k = 0; k < w; k++) base += wsum[k];
    int acc = base + inc - s;
    for (int c = b; c < e; c++) {
        const int v = cnt[c];
        start[c] = acc; cursor[c] = acc; acc += v;
        if (zero_cnt) cnt[c] = 0;
    }
    if (threadIdx.x == nt - 1) start[n] = acc;
}

// lanes of a wave that add to the same counter share one atomic: returns, for an active
// lane, the old value of ctr[key] plus the number of lower lanes with the same key
__device__ __forceinline__ int wave_agg_add(int *ctr, int key, bool active) {
    const int lane = threadIdx.x & 63;
    unsigned long long todo = __ballot(active);
    int mine = 0;
    while (todo) {
        const int leader = __ffsll((long long)todo) - 1;
        const int k = __shfl(key, leader, 64);
        const unsigned long long m = __ballot(active && key == k);
        int base = 0;
        if (lane == leader) base = atomicAdd(&ctr[k], __popcll(m));
        base = __shfl(base, leader, 64);
        if ((m >> lane) & 1ull) mine = base + __popcll(m & ((1ull << lane) - 1ull));
        todo &= ~m;
    }
    return mine;
}

// cell of every lead and the cell counts
__device__ __forceinline__ void cell_count_body(const DevSys &S, int *lead_cell_out, int bid, int nblk) {
    const DevFlags *f = S.flags;
    if (!f->rebuild) return;
    const int l = bid * blockDim.x + threadIdx.x;
    const bool on = l < S.n_lead;
    const int c = on ? lead_cell(S, S.lead[l], f->ncz, f->z0, f->csz) : 0;
    if (on) lead_cell_out[l] = c;
    (void)wave_agg_add(S.cell_cnt, c, on);
    // the skin test's reference positions for the lists built now (src/nnlist.f90:99-119):
    // molecule centres and the moving sites outside rigid molecules (this step's skin test has
    // read the old ones)
    const int n_free = S.n_move - S.n_rigid_sites;
    for (int t = l; t < S.n_rig || t < n_free; t += nblk * blockDim.x) {
        if (t < S.n_rig) { S.ocx[t] = S.cx[t]; S.ocy[t] = S.cy[t]; S.ocz[t] = S.cz[t]; }
        const int i = S.n_rigid_sites + t;
        if (t < n_free) { S.ox[i] = S.x[i]; S.oy[i] = S.y[i]; S.oz[i] = S.z[i]; }
    }
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

// leads into their cells in atomic order (items_tmp, xyz_tmp: the lead and its position and
// site index); cell_order_body then puts every cell in ascending lead order
__device__ __forceinline__ void cell_fill_body(const DevSys &S, const int *lead_cell, int *items_tmp,
                                               double4 *xyz_tmp, int bid, int nblk) {
    if (!S.flags->rebuild) return;
    const int l = bid * blockDim.x + threadIdx.x;
    const bool on = l < S.n_lead;
    const int c = on ? lead_cell[l] : 0;
    const int pos = wave_agg_add(S.cell_cursor, c, on);
    if (on) {
        const int i = S.lead[l];
        items_tmp[pos] = l;
        xyz_tmp[pos] = make_double4(S.x[i], S.y[i], S.z[i], (double)i);
    }
}

__global__ __launch_bounds__(256) void k_cell_fill(DevSys S, const int *lead_cell, int *items_tmp, double4 *xyz_tmp) {
    cell_fill_body(S, lead_cell, items_tmp, xyz_tmp, blockIdx.x, gridDim.x);
}

// ascending lead order inside every cell (cell_items, cell_xyz), one wave per cell: a lane's
// rank is the number of smaller leads in its cell, from the other lanes' values (cells of
// more than 64 leads rank from memory)
__device__ __forceinline__ void cell_order_body(const DevSys &S, const int *items_tmp, const double4 *xyz_tmp,
                                                int bid, int nblk) {
    const DevFlags *f = S.flags;
    if (!f->rebuild) return;
    const int nc = S.g.ncx * S.g.ncy * f->ncz;
    const int lane = threadIdx.x & 63, wpb = blockDim.x >> 6;
    for (int c = bid * wpb + (threadIdx.x >> 6); c < nc; c += nblk * wpb) {
        const int b = S.cell_start[c], n = S.cell_start[c + 1] - b;
        if (n <= 64) {
            const int v = lane < n ? items_tmp[b + lane] : 0x7fffffff;
            int rank = 0;
            for (int k = 0; k < n; k++) rank += __shfl(v, k, 64) < v;
            if (lane < n) {
                S.cell_items[b + rank] = v;
                S.cell_xyz[b + rank] = xyz_tmp[b + lane];
            }
        } else {
            for (int a = lane; a < n; a += 64) {
                const int v = items_tmp[b + a];
                int rank = 0;
                for (int k = 0; k < n; k++) rank += items_tmp[b + k] < v;
                S.cell_items[b + rank] = v;
                S.cell_xyz[b + rank] = xyz_tmp[b + a];
            }
        }
    }
}

__global__ __launch_bounds__(256) void k_cell_order(DevSys S, const int *items_tmp, const double4 *xyz_tmp) {
    cell_order_body(S, items_tmp, xyz_tmp, blockIdx.x, gridDim.x);
}

// positions and site indices of cell-ordered items (the frozen sites' cells, once)
__global__ void k_cell_xyz(const int *items, int n, int offset, const double *x, const double *y, const double *z,
                           double4 *out) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int j = offset + items[p];
    out[p] = make_double4(x[j], y[j], z[j], (double)j);
}

// fill of static items (the frozen sites, at creation): atomic order, sorted after
__global__ void k_cell_fill_static(const int *item_cell, int n_items, int *cursor, int *items) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= n_items) return;
    items[atomicAdd(&cursor[item_cell[l]], 1)] = l;
}

// ascending order inside every cell (rank sort, one wave per cell) of static items, so the
// frozen partners of a list row come in a fixed order; items_out gets the sorted copy
__global__ __launch_bounds__(64) void k_cell_sort(const int *start, const int *items_in, int *items_out, int n_cells) {
    for (int c = blockIdx.x; c < n_cells; c += gridDim.x) {
        const int b = start[c], n = start[c + 1] - b;
        for (int a = threadIdx.x; a < n; a += 64) {
            const int v = items_in[b + a];
            int rank = 0;
            for (int k = 0; k < n; k++) rank += (items_in[b + k] < v);
            items_out[b + rank] = v;
        }
    }
}

__device__ __forceinline__ bool constrained(const DevSys &S, int i, int j) {
    if (i >= S.n_move) return false;
    for (int k = S.con_ptr[i]; k < S.con_ptr[i + 1]; k++)
        if (S.con_idx[k] == j) return true;
    return false;
}

// the list predicate's squared distance from coordinates in hand: list_d2's operation
// order, with the difference taken as (site with the lower index) - (the other), as the
// reference's i < j loop has it
__device__ __forceinline__ double list_d2_xyz(const DevSys &S, double xa, double ya, double za, double xb, double yb,
                                              double zb) {
    double dx = __dsub_rn(xa, xb);
    dx = __dsub_rn(dx, __dmul_rn(S.lx, anint_d(__dmul_rn(dx, S.lxi))));
    double dy = __dsub_rn(ya, yb);
    dy = __dsub_rn(dy, __dmul_rn(S.ly, anint_d(__dmul_rn(dy, S.lyi))));
    double dz = __dsub_rn(za, zb);
    if (S.zper_list) dz = __dsub_rn(dz, __dmul_rn(S.lz, anint_d(__dmul_rn(dz, S.lzi))));
    return __dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), __dmul_rn(dz, dz));
}

// List rows.  Work items are (cell, split): the leads of a cell go in rounds of BUILD_LEADS,
// and item (c, s) takes the rounds r = s, s + BUILD_SPLIT, ... of cell c (grid-stride over
// the items).  The leads of a cell share its 27-cell stencil, so a round stages the
// stencil's candidates (cell-ordered positions, ascending lead order inside each cell) in
// LDS, BUILD_CH at a time, and every wave tests two of the round's leads at a time against
// them, 64 candidates a step, appending the hits with a ballot prefix count: a row holds its
// partners in the stencil's order, the order a serial scan of the same cells gives.  Frozen
// partners: the frozen-site cells that the round's leads reach (their frozen-grid layers
// differ from lead to lead; one run of layers covers them) are staged the same way with
// their layer, and each lead keeps the candidates of its own three layers.
//
// The predicate is decided in FP32 away from the list radius.  Positions are wrapped into
// the box in FP64, rounded to FP32 and, where the stencil wraps round a periodic axis,
// shifted by the box length, so a lead and a candidate within the list radius of each other
// are at their nearest image (a grid of three or more cells on the axis: cells are at least
// the list radius wide); grids of fewer cells take the minimum image per pair.  Coordinates
// are then below 2048 A in magnitude (rounding <= 1.2e-4 A), a difference is within 3e-4 A
// per axis of the FP64 one and a squared distance near the list radius (|d| < 11.2 A) within
// 1.2e-2 A^2: outside a band of BUILD_BAND = 0.05 A^2 around rskinsq the FP32 decision is the
// FP64 one.  Inside the band (about 1e-4 of the candidates), for boxes over 1024 A and for
// open-z coordinates beyond 1024 A (staged as NaN), the reference's FP64 predicate decides,
// on the FP64 positions, with its operation order (list_d2_xyz).
constexpr int BUILD_SPLIT = BUILD_ITEMS_PER_CELL;
constexpr float BUILD_BAND = 0.05f;

// a run of stencil cells: start and size of each cell in the cell-ordered arrays, inclusive
// prefix of the sizes, frozen-grid layer (frozen stencils), periodic image shift
struct Stencil {
    int b[64], n[64], inc[64], layer[64];
    float sx[64], sy[64], sz[64];
    int n_st;
};

struct BuildSmem {
    float4 cand[BUILD_CH];           // staged candidates: x, y, z in FP32, site index (int bits)
    int cand_layer[BUILD_CH];        // frozen candidates: frozen-grid z layer
    double4 lead_xyz[BUILD_LEADS];   // the round's leads: x, y, z, site index
    int lead_l[BUILD_LEADS], lead_cnt[BUILD_LEADS], lead_low[BUILD_LEADS], lead_wz[BUILD_LEADS];
    Stencil mv, fz;                  // the cell's moving stencil (all rounds), a round's frozen one
    int wlo, whi;
};

// the stencil run's candidate q: cell k holds it (first inclusive prefix above q)
__device__ __forceinline__ int stencil_locate(const Stencil &st, int q, int &k) {
    int lo = 0, hi = st.n_st - 1;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (st.inc[mid] <= q) lo = mid + 1; else hi = mid;
    }
    k = lo;
    return st.b[lo] + q - (st.inc[lo] - st.n[lo]);
}

// inclusive prefix of the stencil sizes (wave 0; n_st <= 64), then a barrier
__device__ __forceinline__ void stencil_prefix(Stencil &st) {
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        int v = lane < st.n_st ? st.n[lane] : 0;
        for (int o = 1; o < 64; o <<= 1) {
            const int u = __shfl_up(v, o, 64);
            if (lane >= o) v += u;
        }
        if (lane < st.n_st) st.inc[lane] = v;
    }
    __syncthreads();
}

// a coordinate wrapped into [0, L) in FP64, rounded to FP32
__device__ __forceinline__ float wrap_f32(double v, double L, double Li) { return (float)(v - L * floor(v * Li)); }

// a periodic axis' neighbour cell: index and the image shift (in box lengths) of a wrap
__device__ __forceinline__ int nbr_cell(int c, int o, int n, int no, float &shift, float L) {
    if (no < 3) { shift = 0.f; return o; }
    const int r = c + o - 1;
    shift = r < 0 ? -L : (r >= n ? L : 0.f);
    return (r + n) % n;
}

// the round's lead state in the candidate loop
struct BuildLead {
    double4 me;            // FP64 position and site index
    float x, y, z;         // FP32 staged position (NaN z: no prefilter)
    int i, ld, wz, cnt, lower;
    bool surf;
    int *row;
};

__device__ __forceinline__ void build_body(const DevSys &S, int bid, int nblk) {
    __shared__ BuildSmem sm;
    const DevFlags *f = S.flags;
    if (!f->rebuild) return;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int ncx = S.g.ncx, ncy = S.g.ncy, ncz = f->ncz;
    const int nc = ncx * ncy * ncz;
    const int nxo = ncx >= 3 ? 3 : ncx, nyo = ncy >= 3 ? 3 : ncy;
    const int nzo = S.g.zper ? (ncz >= 3 ? 3 : ncz) : 3;
    const bool zper = S.g.zper != 0;
    const bool zper_d = S.zper_list != 0;   // the predicate's z image
    const float lxf = (float)S.lx, lyf = (float)S.ly, lzf = (float)S.lz;
    const float lxif = (float)S.lxi, lyif = (float)S.lyi, lzif = (float)S.lzi;
    const float r2 = (float)S.rskinsq;
    const float r2_in = r2 - BUILD_BAND, r2_out = r2 + BUILD_BAND;
    const bool fast_box = fmax(S.lx, fmax(S.ly, zper_d ? S.lz : 0.0)) <= 1024.0;
    // shifted images stand for the minimum image when every periodic axis has 3+ cells
    const bool mv_shift = nxo == 3 && nyo == 3 && (!zper || nzo == 3);
    const float fnan = __int_as_float(0x7fc00000);
    auto stage = [&](const double4 &v, float sx, float sy, float sz) {   // FP64 position -> FP32 staged one
        float z = zper_d ? wrap_f32(v.z, S.lz, S.lzi) + sz : (float)v.z;
        if (!fast_box || !(fabsf(z) <= 1024.f)) z = fnan;
        return make_float4(wrap_f32(v.x, S.lx, S.lxi) + sx, wrap_f32(v.y, S.ly, S.lyi) + sy, z,
                           __int_as_float((int)v.w));
    };
    // FP32 squared distance: staged images, or the minimum image per pair
    auto d2f = [&](const BuildLead &L, const float4 &v, bool mi) {
        float dx = L.x - v.x, dy = L.y - v.y, dz = L.z - v.z;
        if (mi) {
            dx = fmaf(-lxf, rintf(dx * lxif), dx);
            dy = fmaf(-lyf, rintf(dy * lyif), dy);
            if (zper_d) dz = fmaf(-lzf, rintf(dz * lzif), dz);
        }
        return fmaf(dx, dx, fmaf(dy, dy, dz * dz));
    };
    auto load_lead = [&](int ld, bool frozen) {
        BuildLead L;
        L.ld = ld;
        L.me = sm.lead_xyz[ld];
        const float4 m = stage(L.me, 0.f, 0.f, 0.f);
        L.x = m.x; L.y = m.y; L.z = m.z;
        L.i = (int)L.me.w;
        L.surf = L.i >= S.n_liquid_sites;
        L.wz = sm.lead_wz[ld];
        L.cnt = sm.lead_cnt[ld];
        L.lower = 0;
        const int l = sm.lead_l[ld];
        L.row = frozen ? S.nbr_mw + (size_t)l * S.cap_mw : S.nbr_mm + (size_t)l * S.cap_mm;
        return L;
    };
    auto append = [&](BuildLead &L, bool hit, int j, int cap) {
        const unsigned long long msk = __ballot(hit);
        if (hit) {
            const int pos = L.cnt + __popcll(msk & ((1ull << lane) - 1ull));
            if (pos < cap) L.row[pos] = j;
            L.lower += j > L.i;
        }
        L.cnt += __popcll(msk);
    };
    // moving candidate v (site j) against lead L: FP32 test, FP64 predicate near the radius
    auto test_mv = [&](const BuildLead &L, const float4 &v, int j, bool ok, bool mi) {
        const float d2 = d2f(L, v, mi);
        bool hit = ok && d2 < r2_in;
        const bool exact = ok && !(d2 < r2_in) && !(d2 >= r2_out);   // in the band, or NaN
        if (__ballot(exact)) {
            if (exact) {
                const double xj = S.x[j], yj = S.y[j], zj = S.z[j];
                const bool ifirst = L.i < j;
                hit = list_d2_xyz(S, ifirst ? L.me.x : xj, ifirst ? L.me.y : yj, ifirst ? L.me.z : zj,
                                  ifirst ? xj : L.me.x, ifirst ? yj : L.me.y, ifirst ? zj : L.me.z) < S.rskinsq;
            }
        }
        hit = hit && j != L.i;
        if (L.surf && hit && j >= S.n_liquid_sites) hit = !constrained(S, L.i, j);
        return hit;
    };
    auto test_fz = [&](const BuildLead &L, const float4 &v, int j, int lay, bool ok, bool mi, int wn) {
        int dz = lay - L.wz;
        if (zper) dz = (dz % wn + wn) % wn;
        const bool mine = ok && (zper ? (wn <= 3 || dz <= 1 || dz == wn - 1) : (dz >= -1 && dz <= 1));
        const float d2 = d2f(L, v, mi);
        bool hit = mine && d2 < r2_in;
        const bool exact = mine && !(d2 < r2_in) && !(d2 >= r2_out);
        if (__ballot(exact)) {
            if (exact) hit = list_d2_xyz(S, L.me.x, L.me.y, L.me.z, S.x[j], S.y[j], S.z[j]) < S.rskinsq;
        }
        if (L.surf && hit) hit = !constrained(S, L.i, j);
        return hit;
    };
    for (int item = bid; item < nc * BUILD_SPLIT; item += nblk) {
        const int c = item / BUILD_SPLIT, split = item - c * BUILD_SPLIT;
        const int b = S.cell_start[c], e = S.cell_start[c + 1];
        const int nrounds = (e - b + BUILD_LEADS - 1) / BUILD_LEADS;
        if (split >= nrounds) continue;   // uniform over the block
        const int cx = c % ncx, cy = (c / ncx) % ncy, cz = c / (ncx * ncy);
        // moving stencil, order (oz, oy, ox)
        if (threadIdx.x == 0) sm.mv.n_st = 0;
        __syncthreads();
        if (threadIdx.x < 32) {
            const int k = threadIdx.x, per = nyo * nxo;
            bool ok = k < nzo * per;
            int cc = 0;
            float sx = 0.f, sy = 0.f, sz = 0.f;
            if (ok) {
                const int oz = k / per, oy = (k / nxo) % nyo, ox = k % nxo;
                int z2;
                if (zper) z2 = nbr_cell(cz, oz, ncz, nzo, sz, lzf);
                else { z2 = cz + oz - 1; ok = z2 >= 0 && z2 < ncz; }
                const int y2 = nbr_cell(cy, oy, ncy, nyo, sy, lyf);
                const int x2 = nbr_cell(cx, ox, ncx, nxo, sx, lxf);
                cc = (z2 * ncy + y2) * ncx + x2;
            }
            // compact the valid cells in order
            const unsigned long long m = __ballot(ok);
            if (ok) {
                const int slot = __popcll(m & ((1ull << k) - 1ull));
                const int cb = S.cell_start[cc];
                sm.mv.b[slot] = cb;
                sm.mv.n[slot] = S.cell_start[cc + 1] - cb;
                sm.mv.layer[slot] = 0;
                sm.mv.sx[slot] = mv_shift ? sx : 0.f;
                sm.mv.sy[slot] = mv_shift ? sy : 0.f;
                sm.mv.sz[slot] = mv_shift ? sz : 0.f;
            }
            if (k == 0) sm.mv.n_st = __popcll(m);
        }
        __syncthreads();
        stencil_prefix(sm.mv);
        const int n_st = sm.mv.n_st;
        const int total = n_st > 0 ? sm.mv.inc[n_st - 1] : 0;
        for (int round = split; round < nrounds; round += BUILD_SPLIT) {
            const int r0 = b + round * BUILD_LEADS;
            const int nl = min(BUILD_LEADS, e - r0);
            if (threadIdx.x == 0) { sm.wlo = 0x7fffffff; sm.whi = -0x7fffffff; }
            __syncthreads();
            if (threadIdx.x < nl) {
                const double4 v = S.cell_xyz[r0 + threadIdx.x];
                sm.lead_xyz[threadIdx.x] = v;
                sm.lead_l[threadIdx.x] = S.cell_items[r0 + threadIdx.x];
                sm.lead_cnt[threadIdx.x] = 0;
                sm.lead_low[threadIdx.x] = 0;
                int wcz = 0;
                if (S.n_wall > 0) {
                    if (zper) {
                        wcz = wrap_cell(v.z, S.lz, S.lzi, S.wncz);
                    } else {
                        double czd = floor((v.z - S.wz0) / S.wcsz);
                        czd = fmin(fmax(czd, -2.0), (double)S.wncz + 1.0);  // far from the crystal: no cell in reach
                        wcz = (int)czd;
                    }
                    atomicMin(&sm.wlo, wcz);
                    atomicMax(&sm.whi, wcz);
                }
                sm.lead_wz[threadIdx.x] = wcz;
            }
            __syncthreads();
            // moving candidates
            for (int q0 = 0; q0 < total; q0 += BUILD_CH) {
                const int nch = min(BUILD_CH, total - q0);
                for (int t = threadIdx.x; t < nch; t += BUILD_THREADS) {
                    int k;
                    const int p = stencil_locate(sm.mv, q0 + t, k);
                    sm.cand[t] = stage(S.cell_xyz[p], sm.mv.sx[k], sm.mv.sy[k], sm.mv.sz[k]);
                }
                __syncthreads();
                for (int ld0 = w; ld0 < nl; ld0 += 2 * BUILD_WAVES) {
                    const int ld1 = ld0 + BUILD_WAVES;
                    const bool two = ld1 < nl;
                    BuildLead A = load_lead(ld0, false);
                    BuildLead B = load_lead(two ? ld1 : ld0, false);
                    for (int s0 = 0; s0 < nch; s0 += 64) {
                        const int a = s0 + lane;
                        const bool ok = a < nch;
                        const float4 v = sm.cand[ok ? a : 0];
                        const int j = __float_as_int(v.w);
                        const bool ha = test_mv(A, v, j, ok, !mv_shift);
                        append(A, ha, j, S.cap_mm);
                        if (two) {
                            const bool hb = test_mv(B, v, j, ok, !mv_shift);
                            append(B, hb, j, S.cap_mm);
                        }
                    }
                    for (int o = 32; o > 0; o >>= 1) A.lower += __shfl_xor(A.lower, o, 64);
                    for (int o = 32; o > 0; o >>= 1) B.lower += __shfl_xor(B.lower, o, 64);
                    if (lane == 0) {
                        sm.lead_cnt[ld0] = A.cnt; sm.lead_low[ld0] += A.lower;
                        if (two) { sm.lead_cnt[ld1] = B.cnt; sm.lead_low[ld1] += B.lower; }
                    }
                }
                __syncthreads();
            }
            // moving rows done: counts and overflow
            if (threadIdx.x < nl) {
                const int l = sm.lead_l[threadIdx.x], cnt = sm.lead_cnt[threadIdx.x];
                S.cnt_mm[l] = cnt;
                S.cnt_lo[l] = sm.lead_low[threadIdx.x];   // i < j pairs of this row; summed after the build
                if (cnt > S.cap_mm) {
                    atomicMax(&S.flags->overflow_mm, cnt);
                    S.flags->overflow_step = S.flags->steps;   // the same value from every writer
                }
                sm.lead_cnt[threadIdx.x] = 0;   // frozen partners next
            }
            if (S.n_wall == 0) {
                if (threadIdx.x < nl) S.cnt_mw[sm.lead_l[threadIdx.x]] = 0;
                __syncthreads();
                continue;
            }
            // frozen stencil: the layers the round's leads reach, then (oy, ox)
            const int wlo = sm.wlo, whi = sm.whi, wn = S.wncz;
            int zl, nlay;
            bool all_layers = false;
            if (zper) {
                all_layers = wn <= 3 || whi - wlo + 3 >= wn;
                zl = all_layers ? 0 : wlo - 1;
                nlay = all_layers ? wn : whi - wlo + 3;
            } else {
                zl = max(0, wlo - 1);
                nlay = max(0, min(wn - 1, whi + 1) - zl + 1);
            }
            // frozen z images are per pair with a periodic z; x, y as the moving stencil
            const bool fz_shift = nxo == 3 && nyo == 3 && !zper;
            __syncthreads();
            if (threadIdx.x < 64) {
                const int k = threadIdx.x, per = nyo * nxo;
                const bool ok = k < nlay * per;
                if (ok) {
                    const int oz = k / per, oy = (k / nxo) % nyo, ox = k % nxo;
                    float sx, sy;
                    const int z2 = zper ? ((zl + oz) % wn + wn) % wn : zl + oz;
                    const int y2 = nbr_cell(cy, oy, ncy, nyo, sy, lyf);
                    const int x2 = nbr_cell(cx, ox, ncx, nxo, sx, lxf);
                    const int cc = (z2 * ncy + y2) * ncx + x2;
                    const int cb = S.wcell_start[cc];
                    sm.fz.b[k] = cb;
                    sm.fz.n[k] = S.wcell_start[cc + 1] - cb;
                    sm.fz.layer[k] = z2;
                    sm.fz.sx[k] = fz_shift ? sx : 0.f;
                    sm.fz.sy[k] = fz_shift ? sy : 0.f;
                    sm.fz.sz[k] = 0.f;
                }
                if (k == 0) sm.fz.n_st = min(64, nlay * per);
            }
            __syncthreads();
            stencil_prefix(sm.fz);
            const int n_wst = sm.fz.n_st;
            const int wtotal = n_wst > 0 ? sm.fz.inc[n_wst - 1] : 0;
            for (int q0 = 0; q0 < wtotal; q0 += BUILD_CH) {
                const int nch = min(BUILD_CH, wtotal - q0);
                for (int t = threadIdx.x; t < nch; t += BUILD_THREADS) {
                    int k;
                    const int p = stencil_locate(sm.fz, q0 + t, k);
                    sm.cand[t] = stage(S.wcell_xyz[p], sm.fz.sx[k], sm.fz.sy[k], sm.fz.sz[k]);
                    sm.cand_layer[t] = sm.fz.layer[k];
                }
                __syncthreads();
                for (int ld0 = w; ld0 < nl; ld0 += 2 * BUILD_WAVES) {
                    const int ld1 = ld0 + BUILD_WAVES;
                    const bool two = ld1 < nl;
                    BuildLead A = load_lead(ld0, true);
                    BuildLead B = load_lead(two ? ld1 : ld0, true);
                    for (int s0 = 0; s0 < nch; s0 += 64) {
                        const int a = s0 + lane;
                        const bool ok = a < nch;
                        const float4 v = sm.cand[ok ? a : 0];
                        const int j = __float_as_int(v.w), lay = sm.cand_layer[ok ? a : 0];
                        const bool ha = test_fz(A, v, j, lay, ok, !fz_shift, wn);
                        append(A, ha, j, S.cap_mw);
                        if (two) {
                            const bool hb = test_fz(B, v, j, lay, ok, !fz_shift, wn);
                            append(B, hb, j, S.cap_mw);
                        }
                    }
                    if (lane == 0) {
                        sm.lead_cnt[ld0] = A.cnt;
                        if (two) sm.lead_cnt[ld1] = B.cnt;
                    }
                }
                __syncthreads();
            }
            if (threadIdx.x < nl) {
                const int cntw = sm.lead_cnt[threadIdx.x];
                S.cnt_mw[sm.lead_l[threadIdx.x]] = cntw;
                if (cntw > S.cap_mw) {
                    atomicMax(&S.flags->overflow_mw, cntw);
                    S.flags->overflow_step = S.flags->steps;
                }
            }
            __syncthreads();
        }
    }
}

__global__ __launch_bounds__(BUILD_THREADS) void k_build(DevSys S) {
    build_body(S, blockIdx.x, gridDim.x);
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
    cell_fill_body(r.S, r.lead_cell, r.items_tmp, r.xyz_tmp, m.y, r.lead_blocks);
}
__global__ __launch_bounds__(256) void k_cell_order_b(const RepArgs *__restrict__ RA, const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    const RepArgs &r = RA[m.x];
    cell_order_body(r.S, r.items_tmp, r.xyz_tmp, m.y, r.order_blocks);
}
__global__ __launch_bounds__(BUILD_THREADS) void k_build_b(const RepArgs *__restrict__ RA, const int2 *__restrict__ map) {
    const int2 m = map[blockIdx.x];
    build_body(RA[m.x].S, m.y, RA[m.x].build_blocks);
}
// every system's flags side by side, for one copy to the host
__global__ void k_flags_b(const RepArgs *__restrict__ RA, int n, DevFlags *__restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = *RA[i].S.flags;
}
}  // namespace mdf
