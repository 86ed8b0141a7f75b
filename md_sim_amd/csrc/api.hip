// C ABI of the force engine (include/mdforce.h) and its host-side runtime: constant
// derivation, k-set enumeration, tail-correction tables, device buffers and the launch
// sequence of one force evaluation.
//
// Reference correspondence:
//   mdf_create   force_init (src/forcem.f90:54-81): force_fouri_init (k-set, A_k, static
//                frozen-site S(k), self energy, :657-825), force_real_init (wall tail tables,
//                :157-231) and nnlist_init/nnlist_make (src/nnlist.f90:14-164)
//   mdf_compute  force_get (src/forcem.f90:93-129)
//   mdf_zero     force_zero (src/forcem.f90:84-90)
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mdforce.h"
#include "mdf_internal.h"

namespace mdf {
__global__ void k_skin(DevSys S, double *dip_part);
__global__ void k_decide(DevSys S, const double *dip_part, int n_dip, double *dipz_out, int force_rebuild);
__global__ void k_cell_scan_build(DevSys S);
__global__ void k_cell_count(DevSys S, int *lead_cell_out);
__global__ void k_cell_scan(int *cnt, int *start, int *cursor, int n);
__global__ void k_cell_fill(DevSys S, const int *lead_cell, int *items_tmp, double4 *xyz_tmp);
__global__ void k_cell_order(DevSys S, const int *items_tmp, const double4 *xyz_tmp);
__global__ void k_cell_fill_static(const int *item_cell, int n_items, int *cursor, int *items);
__global__ void k_cell_sort(const int *start, const int *items_in, int *items_out, int n_cells);
__global__ void k_cell_xyz(const int *items, int n, int offset, const double *x, const double *y, const double *z,
                           double4 *out);
__global__ void k_build(DevSys S);
__global__ void k_wall_cells(DevSys S, int *wcell_of, int ncz, double z0, double csz);
__global__ void k_count_cells(const int *item_cell, int n_items, int *cnt);
__global__ void k_zero_int(int *p, int n);
template <typename R, int G, bool VIR, bool ACC, bool STRIDED, int NMC>
__global__ void k_pair(DevSys S);
template <typename R, int NS>
__global__ void k_sf(const double *, const double *, const double *, const double *, int, double, double, double,
                     const KRow *, int, int, double2 *, int, const MtsState *);
__global__ void k_sf_reduce(const double2 *, int, int, const double2 *, const double *, const int *, double, double,
                            double, double, double2 *, double4 *, double *, const MtsState *, unsigned *, const int *,
                            int, int, const float2 *);
template <typename R, int NS>
__global__ void k_bp(const double *, const double *, const double *, const double *, int, double, double, double,
                     const KRow *, const int *, const double4 *, double, double, double, double *, const MtsState *);
__global__ void k_sg_control(MtsArgs a, int again);
template <int G>
__global__ void k_sg_features(DevSys S, MtsArgs a);
__global__ void k_sg_norm(MtsArgs a, int n);
__global__ void k_sg_pack(DevSys S, MtsArgs a);
__global__ void k_sg_infer(MtsArgs a, int n, int after_training, const DevFlags *flags);
__global__ void k_sg_train(MtsArgs a, int n, int it, const DevFlags *flags);
__global__ void k_sg_adam(MtsArgs a, int nb, int it, const DevFlags *flags);
__global__ void k_sg_eval_prep(MtsArgs a, const float *weights, int n, const double *target, const double *target_b,
                               int adam_t);
__global__ void k_sg_combine(DevSys S, MtsArgs a);
__global__ void k_sg_finish(DevSys S, MtsArgs a, Scalars *scal);
__global__ void k_sg_exact(DevSys S, MtsArgs a, const double *fpart, int n_groups);
__global__ void k_sg_judge(DevSys S, MtsArgs a);
__global__ void k_sg_delta(DevSys S, MtsArgs a);
__global__ void k_sg_post(DevSys S, MtsArgs a);
constexpr int SG_FEAT_G = 16;  // lanes per lead in the input kernel
constexpr int BP_NS = 4;   // sites per back-projection thread
__global__ void k_epilogue(DevSys S, const double *fpart, int n_groups, const WallLR *wl, int ew3dc, double crf0,
                           const double *dipz, double *epart, double *fcom, double *tq, const MtsState *mts,
                           const double *fmts, double *efi);
template <int NT>
__global__ void k_sf_mma(const double *x, const double *y, const double *z, const double *q, int n, double cux,
                         double cuy, double cuz, const KRow *rows, int n_rows, int knmax, int n_nt, float2 *partial,
                         int nk, int sb_sites, const MtsState *mts);
__global__ void k_sf_reduce_rows(const double2 *, const float2 *, int, int, const KRow *, const double2 *,
                                 const double *, double, double, double, double, double2 *, double4 *, double *,
                                 const MtsState *, unsigned *, int *, int, int);
template <int NT>
__global__ void k_sf64(const double *, const double *, const double *, const double *, int, double, double, double,
                       const int2 *, const int4 *, const int *, int, int, int, const int *, int, double2 *, int,
                       const MtsState *);
__global__ void k_sf64_reduce(const double2 *, int, int, const int *, const int *, const KRow *, const double2 *,
                              const double *, double, double, double, double, double2 *, double4 *, double *,
                              const int2 *, double *, const MtsState *);
template <int NKS, int NG>
__global__ void k_bp64(const double *, const double *, const double *, const double *, int, double, double, double,
                       const int4 *, int, const int2 *, const int2 *, const double *, double, double, double, double *,
                       const MtsState *);
template <int NT>
__global__ void k_bp_mma(const double *x, const double *y, const double *z, const double *q, int n, double cux,
                         double cuy, double cuz, const int2 *rowkk, const int2 *tile_ks, int n_rows, int knmax,
                         int n_nt, const unsigned *tpack, const int *tscale, double pipix, double pipiy, double pipiz,
                         double *fpart,
                         const MtsState *mts);
constexpr int BPM_WAVES_MAX = 8; // waves per k_bp_mma block (16 sites each), at most
constexpr int BPM_ROWS = 1024;   // rows k_bp_mma keeps in LDS
constexpr int SFM_SITES = 512;   // sites per k_sf_mma block
constexpr int SFM_ROWS_PER_BLOCK = 64;
constexpr size_t SFM_LDS_MAX = 64 * 1024 - 1024;   // dynamic LDS of a k_sf_mma block, at most
// the k_sf_mma instance for nt kn tiles of 16
static const void *sfm_kernel(int nt) {
    switch (nt) {
        case 1: return reinterpret_cast<const void *>(&k_sf_mma<1>);
        case 2: return reinterpret_cast<const void *>(&k_sf_mma<2>);
        case 3: return reinterpret_cast<const void *>(&k_sf_mma<3>);
        case 4: return reinterpret_cast<const void *>(&k_sf_mma<4>);
        case 5: return reinterpret_cast<const void *>(&k_sf_mma<5>);
        case 6: return reinterpret_cast<const void *>(&k_sf_mma<6>);
        case 7: return reinterpret_cast<const void *>(&k_sf_mma<7>);
        default: return reinterpret_cast<const void *>(&k_sf_mma<8>);
    }
}
constexpr int SFM_TILES_MAX = 8; // kn tiles of 16 (knmax <= 63)
__global__ void k_obs_ngc(DevSys S, ObsArgs o);
__global__ void k_obs_efipos(DevSys S, ObsArgs o);
__global__ void k_obs_ngang(DevSys S, const int *ang_type, int n_atypes, unsigned long long *hist);
__global__ void k_chem_scan(DevSys S, ChemArgs c);
__global__ void k_chem_first(DevSys S, ChemArgs c);
__global__ void k_chem_rot(DevSys S, ChemArgs c);
__global__ void k_chem_bsum(ChemArgs c);
__global__ void k_chem_accum(ChemArgs c);
__global__ void k_colsum(DevSys S, const int *lead_slot, int per_block, double *colpart);
__global__ void k_pair_totals(DevSys S);
__global__ void k_final(DevSys S, FinalArgs A, const double *colpart, const double *kpart, const double *eppart,
                        const double *dipz, double *e_vdw, double *e_coul, Scalars *out);

__global__ void k_skin_b(const RepArgs *, const int2 *);
__global__ void k_decide_b(const RepArgs *);
__global__ void k_cell_count_b(const RepArgs *, const int2 *);
__global__ void k_cell_scan_build_b(const RepArgs *);
__global__ void k_cell_fill_b(const RepArgs *, const int2 *);
__global__ void k_cell_order_b(const RepArgs *, const int2 *);
__global__ void k_build_b(const RepArgs *, const int2 *);
__global__ void k_flags_b(const RepArgs *, int, DevFlags *);
template <typename R, int G, bool VIR, bool ACC, bool STRIDED, int NMC>
__global__ void k_pair_b(const RepArgs *, const int2 *);
template <typename R>
__global__ void k_sf_b(const RepArgs *, const int2 *);
__global__ void k_sf_reduce_b(const RepArgs *, const int2 *);
template <int NT>
__global__ void k_sf_mma_b(const RepArgs *, const int2 *);
template <int NT>
__global__ void k_bp_mma_b(const RepArgs *, const int2 *);
__global__ void k_sf_reduce_rows_b(const RepArgs *, const int2 *);
template <int NT>
__global__ void k_sf64_b(const RepArgs *, const int2 *);
__global__ void k_sf64_reduce_b(const RepArgs *, const int2 *);
template <int NKS, int NG>
__global__ void k_bp64_b(const RepArgs *, const int2 *);
template <typename R>
__global__ void k_bp_b(const RepArgs *, const int2 *);
__global__ void k_epilogue_b(const RepArgs *, const int2 *);
__global__ void k_colsum_b(const RepArgs *, const int2 *);
__global__ void k_final_b(const RepArgs *);

// device-resident inputs -> the handle's buffers, one launch
__device__ __forceinline__ void gather_body(const DevSys &S, const double *__restrict__ x, const double *__restrict__ y,
                         const double *__restrict__ z, const double *__restrict__ cx, const double *__restrict__ cy,
                         const double *__restrict__ cz, int bid, int nblk) {
    int t = bid * blockDim.x + threadIdx.x;
    if (t < S.n_move) { S.x[t] = x[t]; S.y[t] = y[t]; S.z[t] = z[t]; }
    if (t < S.n_rig) {
        const_cast<double *>(S.cx)[t] = cx[t];
        const_cast<double *>(S.cy)[t] = cy[t];
        const_cast<double *>(S.cz)[t] = cz[t];
    }
}

__global__ void k_gather(DevSys S, const double *__restrict__ x, const double *__restrict__ y,
                         const double *__restrict__ z, const double *__restrict__ cx, const double *__restrict__ cy,
                         const double *__restrict__ cz) {
    gather_body(S, x, y, z, cx, cy, cz, blockIdx.x, gridDim.x);
}

// the same for every system of a batch: in[6 * system + c] holds its six input pointers
__global__ void k_gather_b(const RepArgs *__restrict__ RA, const int2 *__restrict__ map,
                           const double *const *__restrict__ in) {
    const int2 m = map[blockIdx.x];
    const double *const *p = in + 6 * m.x;
    gather_body(RA[m.x].S, p[0], p[1], p[2], p[3], p[4], p[5], m.y, RA[m.x].gather_blocks);
}
}  // namespace mdf

using namespace mdf;

static thread_local std::string g_err;

static int fail(int code, const char *fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(x)                                                                                   \
    do {                                                                                            \
        hipError_t e_ = (x);                                                                        \
        if (e_ != hipSuccess) return fail(MDF_ERR_HIP, "%s: %s (%s:%d)", #x, hipGetErrorString(e_), \
                                          __FILE__, __LINE__);                                      \
    } while (0)

// reference constants: default-real literals widened to REAL*8 (src/global.f90:18-24,
// src/forcem.f90:11-13,59,1081) or exact values
struct Consts {
    double pi, e, na, eps0, k4pi, a2m, a1, a2, a3, a4, a5, p, c309, e30, e5;
};
static Consts make_consts(int mode) {
    Consts c;
    if (mode == MDF_CONST_REFERENCE) {
        c.pi = (double)3.1415926535898f; c.e = (double)1.6021892E-19f; c.na = (double)6.022045E23f;
        c.eps0 = (double)8.85418782E-12f; c.k4pi = (double)8.987551784952e+9f; c.a2m = (double)1.0E-10f;
        c.a1 = (double)0.254829592f; c.a2 = (double)-0.284496736f; c.a3 = (double)1.421413741f;
        c.a4 = (double)-1.453152027f; c.a5 = (double)1.061405429f; c.p = (double)0.3275911f;
        c.c309 = (double)3.09f; c.e30 = (double)1e-30f; c.e5 = (double)1e-5f;
    } else {
        c.pi = 3.14159265358979323846; c.e = 1.6021892E-19; c.na = 6.022045E23;
        c.eps0 = 1.0 / (4.0 * 3.14159265358979323846 * 8.987551784952e+9); c.k4pi = 8.987551784952e+9;
        c.a2m = 1.0E-10; c.a1 = 0.254829592; c.a2 = -0.284496736; c.a3 = 1.421413741; c.a4 = -1.453152027;
        c.a5 = 1.061405429; c.p = 0.3275911; c.c309 = 3.09; c.e30 = 1e-30; c.e5 = 1e-5;
    }
    return c;
}

template <typename T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
    hipError_t alloc(size_t cnt) {
        n = cnt;
        return hipMalloc((void **)&p, sizeof(T) * (cnt > 0 ? cnt : 1));
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

// list: skin test and rebuild decision; build: cell binning and list build (early exits
// on the steps that keep the list)
static const char *STAGE_NAMES[] = {"list", "build", "pair", "surrogate", "sf", "sf_reduce", "backproj", "epilogue",
                                    "reduce", "step"};
constexpr int N_STAGES = 9;  // timed stages; entry N_STAGES of the names is the whole step
// kernels timed by their own dispatch (start and stop events of hipExtLaunchKernelGGL) in the
// steps with stage timing on: the pair kernel and the three reciprocal-sum kernels
static const char *KT_NAMES[] = {"k_pair", "k_sf", "k_sf_reduce", "k_bp"};
constexpr int N_KT = 4;
enum { KT_PAIR = 0, KT_SF = 1, KT_SFRED = 2, KT_BP = 3 };

struct mdf_system {
    mdf_topology top;
    mdf_params par;
    Consts c;
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = true;
    // the captured step graph forks the k-space chain (structure factor, reduction,
    // back-projection) onto a second stream beside the list and pair kernels, and joins it
    // before the epilogue; direct launches keep one stream
    hipStream_t stream2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    bool kspace_forked = false;   // set while a capture holds the forked chain
    int fork_env = -1;            // MDF_FORK_KSPACE: 1 always, 0 never, unset: large systems only
    int nijt = 0, n_lead = 0, n_wall = 0;
    DevSys S;
    // host copies
    std::vector<int> h_lead, h_lead_slot, h_itype;
    std::vector<int> slot_type, cls_type;
    std::vector<KChunk> chunks;
    std::vector<KRow> rows;
    std::vector<int> bp_row_start;
    std::vector<int> kvec;
    std::vector<double> ak, s_wall;
    double ewqWW = 0, uself = 0, crf0 = 0;
    WallLR wl;
    bool fixed_cap = false;
    // device buffers
    DBuf<int> itype, nsi, ijtype, vkind, cls, lead, lead_slot, con_ptr, con_idx, ang_ptr, ang_inc, ang_ijk;
    DBuf<int> pmeta, wcls;           // pair kernel: per moving site slot | nsi << 4, per frozen site class
    DBuf<double> prec, ptab;         // pair kernel: molecule records, vdW table per (slot, class)
    DBuf<double4> wrec;              // pair kernel: frozen sites (x, y, z, q)
    DBuf<double> q, eps4, sig6, aa, rhoi, cc, ang_kt, ang_th;
    DBuf<double> x, y, z, cx, cy, cz, ocx, ocy, ocz, ox, oy, oz, fx, fy, fz, part;
    DBuf<int> nbr_mm, cnt_mm, nbr_mw, cnt_mw, cnt_lo;
    DBuf<int> cell_cnt, cell_start, cell_items, cell_cursor, lead_cell, cell_items_tmp;
    DBuf<double4> cell_xyz, wcell_xyz, cell_xyz_tmp;
    DBuf<int> wcell_start, wcell_items;
    DBuf<DevFlags> flags;
    DBuf<double> dip_part, dipz;
    DBuf<KChunk> d_chunks;
    DBuf<KRow> d_rows;
    DBuf<int> d_kvec, d_bp_row_start;
    DBuf<double> d_ak;
    DBuf<double2> d_swall, d_s, sf_partial;
    DBuf<double4> d_t;
    DBuf<double> kpart, fpart, eppart, colpart, fcom, torque, e_vdw, e_coul;
    DBuf<WallLR> d_wl;
    DBuf<Scalars> d_scal;
    // launch geometry
    // the FP32 form's structure factor on the matrix cores (k_sf_mma)
    bool sf_mma = false, sf_mma_allowed = true, bp_mma = false, bp_mma_allowed = true;
    DBuf<int2> d_rowkk, d_tile_ks;   // k_bp_mma: (kl, km) per row, K-step range per 16-row tile
    DBuf<int> d_krow;                // row of each k-vector
    DBuf<unsigned> d_tpack;          // T split for k_bp_mma, [row][plane][kn + knmax]
    DBuf<int> d_tscale;              // each row's power-of-two scales of the split T
    int n_sb_mma = 1, n_rg_mma = 1, knmax = 0, n_nt = 0;
    // the FP64 form's reciprocal sum on the matrix cores (k_sf64, k_sf64_reduce, k_bp64)
    bool mma64 = false, mma64_allowed = true;
    int klmax = 0, kmmax = 0;
    int s64_nt = 1, s64_npass = 1, s64_nsb = 1, s64_sites = SF64_SUB, s64_pcols = 1, s64_sites_env = 0;
    int s64_pass_env = 0;   // MDF_SF64_PASSES: at least this many passes (fewer, larger site blocks)
    int b64_nks = 4, b64_ng = 2, b64_nrg = 1, b64_nmt = 1, b64_sites = 128, b64_rg_env = 0, b64_ng_env = 0;
    DBuf<int2> d_s64_rowkk, d_b64_tile, d_b64_rg;
    DBuf<int4> d_s64_waves;
    DBuf<int> d_s64_pncol, d_s64_pcoff, d_s64_sorted_of;
    DBuf<int4> d_b64_rinfo;
    DBuf<double2> d_s64_part;
    DBuf<double> d_b64_tpk;
    int sfm_sites = SFM_SITES, bpm_waves = BPM_WAVES_MAX;   // k_sf_mma sites per block, k_bp_mma waves
    int sfm_sites_env = 0, bpm_waves_env = 0;                // fixed by MDF_SFM_SITES / MDF_BPM_WAVES
    int n_cu = 256;                                          // compute units of the device
    int skin_blocks = 1, n_sb = 1, n_cg = 1, cpb = 1, n_kblk = 1, n_rg = 1, rpg = 1, n_epblk = 1, n_colblk = 1,
        colper = 1, ncells_xy = 1, n_cells_cap = 1;
    int build_blocks = 1, order_blocks = 1;   // list build: blocks (cells at a time), order pass
    // timing: a ring of event sets, one per in-flight step; stage durations are accumulated
    // when a set is reused or at the next synchronisation
    static constexpr int NRING = 32;
    hipEvent_t ring[NRING][N_STAGES + 1] = {};
    hipEvent_t kring[NRING][N_KT][2] = {};   // kernel start/stop events per slot
    bool kring_used[NRING][N_KT] = {};
    bool ring_used[NRING] = {};
    long long step_no = 0;
    hipEvent_t *ev = nullptr;       // events of the step being launched
    hipEvent_t (*kev)[2] = nullptr;  // kernel events of the step being launched (stage timing only)
    bool *kev_used = nullptr;
    // one captured step graph per ring slot (each records its slot's events)
    hipGraphExec_t graph[NRING] = {};
    bool use_graphs = true;
    hipEvent_t in_event = nullptr;   // orders a device-input step after the inputs' producer
    bool pair_strided = false;  // pair loop form: row-strided (true) or cut-off compacted (false)
    int pair_cell_order = 1;    // pair kernel groups take leads in the list build's cell order
    int pair_g = 16;            // lanes per lead in the pair kernel (8, 16 or 32)
    // launch-shape targets of the reciprocal kernels (blocks per launch; the row groups follow)
    int sf_target_blocks = 4096, bp_target_blocks = 1024;
    double stage_sum[N_STAGES] = {};
    double kt_sum[N_KT] = {};
    long long kt_n[N_KT] = {};
    long long stage_steps = 0;       // steps with per-stage events (direct launches)
    double step_sum = 0.0;
    long long step_steps = 0;        // steps with start/end events (any mode)
    bool slot_staged[NRING] = {};
    float stage_ms[N_STAGES] = {};   // last completed step
    int stage_timing = 0;            // per-stage events: direct launches instead of graphs;
                                     // 2: and a list build on every step
    long long timing_failures = 0;
    bool pending = false;
    int last_rebuilt = 0;
    DevFlags last_flags;
    int *h_mirror = nullptr;      // host-mapped flag mirror (DevSys::flag_mirror)
    bool rebuild_next = false;    // a fixed-capacity overflow left truncated rows: rebuild next step
    // reciprocal-force surrogate and multiple time steps (off unless mdf_mts_enable)
    bool mts_on = false;
    int mts_model = 1, mts_interval = 4, mts_train_frames = 0, mts_train_iters = 1;
    double mts_tol = 1e-3;
    float mts_lr = 1e-3f;
    int mts_again = 0;                   // the step being launched redoes an overflowed one
    int sg_n_sblk = 1, sg_n_tblk = 1, sg_n_iblk = 1, sg_train_stride = 1;
    DBuf<MtsState> mts_st;
    DBuf<float> sg_feat, sg_vec, sg_out, sg_params, sg_m, sg_v, sg_gsum, sg_norm, sg_gpart, sg_feat_prev,
        sg_vec_prev;
    DBuf<unsigned short> sg_wb;
    DBuf<double> sg_delta, sg_frec, sg_fmts, sg_rprev, sg_fprev, sg_jpart, sg_epart, sg_frec_prev, sg_delta_prev;
    DBuf<float4> sg_pos4;
    DBuf<float> sg_out_prev;
    // observables of the reference's hot loops (off unless mdf_observables_enable)
    bool obs_on = false;
    ObsArgs obs;
    DBuf<unsigned long long> obs_ngc;
    DBuf<double> obs_efi, obs_efipos;
    // force_intra's angle histogram (off unless mdf_angle_histogram_enable)
    bool ang_on = false;
    int n_atypes = 0;
    DBuf<int> ang_type;
    DBuf<unsigned long long> ang_hist;
    DBuf<int> sg_msize;
    // Widom insertions (off unless mdf_chempot_enable)
    bool chem_on = false;
    mdf_chempot_params chem_p;
    ChemArgs chem;
    double chem_seed = 0, chem_lwmin = 0, chem_lwmax = 0, chem_facbin = 0;
    double chem_aco = 0, chem_asi = 0, chem_lo = 0;
    size_t chem_lds = 0;
    DBuf<double> ch_ins, ch_rot, ch_uox, ch_nbr_d, ch_boltz, ch_bsum, ch_pave;
    DBuf<int> ch_il, ch_insrot, ch_rotins, ch_over, ch_nnbr, ch_nbr, ch_status;
    DBuf<unsigned long long> ch_inave;
    double *h_ins = nullptr, *h_rot = nullptr;   // pinned staging of one call's draws
    int *h_il = nullptr, *h_insrot = nullptr, *h_rotins = nullptr, *h_status = nullptr;
    std::vector<double> ch_seed_pos;             // generator state after each insertion's position
    double ch_seed_end = 0;
    bool ch_pending = false;
    hipEvent_t ch_done = nullptr;
};

static void chem_free(mdf_system *s) {
    DBuf<double> *db[] = {&s->ch_ins, &s->ch_rot, &s->ch_uox, &s->ch_nbr_d, &s->ch_boltz, &s->ch_bsum, &s->ch_pave};
    for (auto *b : db) b->release();
    DBuf<int> *ib[] = {&s->ch_il, &s->ch_insrot, &s->ch_rotins, &s->ch_over, &s->ch_nnbr, &s->ch_nbr, &s->ch_status};
    for (auto *b : ib) b->release();
    s->ch_inave.release();
    void *hb[] = {s->h_ins, s->h_rot, s->h_il, s->h_insrot, s->h_rotins, s->h_status};
    for (void *b : hb)
        if (b) (void)hipHostFree(b);
    s->h_ins = s->h_rot = nullptr;
    s->h_il = s->h_insrot = s->h_rotins = s->h_status = nullptr;
    if (s->ch_done) (void)hipEventDestroy(s->ch_done);
    s->ch_done = nullptr;
    s->ch_pending = false;
    s->chem_on = false;
}

static void mts_free(mdf_system *s) {
    s->mts_st.release();
    DBuf<float> *fb[] = {&s->sg_feat, &s->sg_vec, &s->sg_out, &s->sg_params, &s->sg_m, &s->sg_v, &s->sg_gsum,
                         &s->sg_norm, &s->sg_gpart, &s->sg_feat_prev, &s->sg_vec_prev};
    for (auto *b : fb) b->release();
    s->sg_wb.release();
    s->sg_pos4.release();
    s->sg_out_prev.release();
    s->sg_msize.release();
    DBuf<double> *db[] = {&s->sg_delta, &s->sg_frec, &s->sg_fmts, &s->sg_rprev, &s->sg_fprev, &s->sg_jpart,
                          &s->sg_epart, &s->sg_frec_prev, &s->sg_delta_prev};
    for (auto *b : db) b->release();
    s->mts_on = false;
}

static void free_all(mdf_system *s) {
    DBuf<int> *ib[] = {&s->itype, &s->nsi, &s->ijtype, &s->vkind, &s->cls, &s->lead, &s->lead_slot, &s->con_ptr,
                       &s->con_idx, &s->ang_ptr, &s->ang_inc, &s->ang_ijk, &s->nbr_mm, &s->cnt_mm, &s->nbr_mw, &s->cnt_mw, &s->cnt_lo, &s->cell_cnt, &s->cell_start,
                       &s->cell_items, &s->cell_cursor, &s->lead_cell, &s->cell_items_tmp, &s->wcell_start,
                       &s->wcell_items, &s->d_kvec};
    for (auto *b : ib) b->release();
    s->cell_xyz.release(); s->cell_xyz_tmp.release(); s->wcell_xyz.release();
    DBuf<double> *db[] = {&s->q, &s->eps4, &s->sig6, &s->aa, &s->rhoi, &s->cc, &s->x, &s->y, &s->z, &s->cx,
                          &s->cy, &s->cz, &s->ocx, &s->ocy, &s->ocz, &s->ox, &s->oy, &s->oz, &s->fx, &s->fy,
                          &s->fz, &s->part, &s->dip_part, &s->dipz, &s->d_ak, &s->kpart, &s->fpart, &s->eppart,
                          &s->colpart, &s->fcom, &s->torque, &s->e_vdw, &s->e_coul, &s->ang_kt, &s->ang_th};
    for (auto *b : db) b->release();
    s->d_rowkk.release(); s->d_tile_ks.release(); s->d_krow.release(); s->d_tpack.release(); s->d_tscale.release();
    s->d_s64_rowkk.release(); s->d_s64_waves.release(); s->d_b64_tile.release(); s->d_b64_rg.release();
    s->d_s64_pncol.release(); s->d_s64_pcoff.release(); s->d_s64_sorted_of.release(); s->d_b64_rinfo.release();
    s->d_s64_part.release(); s->d_b64_tpk.release();
    s->pmeta.release(); s->wcls.release(); s->prec.release(); s->ptab.release(); s->wrec.release();
    s->flags.release(); s->d_bp_row_start.release(); s->d_chunks.release(); s->d_rows.release(); s->d_swall.release(); s->d_s.release();
    s->d_t.release(); s->sf_partial.release(); s->d_wl.release(); s->d_scal.release();
    mts_free(s);
    s->obs_ngc.release(); s->obs_efi.release(); s->obs_efipos.release();
    s->obs_on = false;
    s->ang_type.release(); s->ang_hist.release();
    s->ang_on = false;
    chem_free(s);
    if (s->h_mirror) (void)hipHostFree(s->h_mirror);
    s->h_mirror = nullptr;
    s->S.flag_mirror = nullptr;
}

template <typename T>
static hipError_t upload(DBuf<T> &b, const T *h, size_t n) {
    hipError_t e = b.alloc(n);
    if (e != hipSuccess) return e;
    if (n) return hipMemcpy(b.p, h, sizeof(T) * n, hipMemcpyHostToDevice);
    return hipSuccess;
}

// ---- force_real_init wall tail tables (src/forcem.f90:157-231) --------------------------
static void wall_tables(mdf_system *s, const double *q_unused) {
    (void)q_unused;
    const mdf_topology &t = s->top;
    WallLR &w = s->wl;
    memset(&w, 0, sizeof w);
    int nit = t.n_types, nliq = nit - 2;
    double lx = s->par.box[0], ly = s->par.box[1], lw = t.wall_lw, rcut = s->par.rcut;
    double rholiq = (double)t.n_move / (lx * ly * lw);
    double rhowall = (double)(t.wall_nss[0] + t.wall_nss[1]) / (2.0 * t.wall_cell[0] * t.wall_cell[1] * t.wall_cell[2]);
    double xd[MAX_TYPES + 2] = {0};
    for (int i = 0; i < t.n_move; i++) xd[t.site_type[i]] += 1.0;
    for (int it = 1; it <= nliq; it++) xd[it] = xd[it] / (double)(float)t.n_move * rholiq;
    float tot = (float)(t.wall_nss[0] + t.wall_nss[1]);
    xd[nit - 1] = (double)((float)t.wall_nss[0] / tot) * rhowall;
    xd[nit] = (double)((float)t.wall_nss[1] / tot) * rhowall;
    for (int it = 0; it <= MAX_TYPES; it++) w.flr_bur[it] = 1.0;
    for (int it = 1; it <= nliq; it++) {
        for (int jt = 1; jt <= nit; jt++) {
            int ijt = t.ijtype[(it - 1) * nit + (jt - 1)] - 1;
            double sgn = jt <= nliq ? -1.0 : 1.0, flr = 0.0, ul = 0.0;
            if (t.vdw_kind[ijt] == 1) {
                double e = t.vdw_eps4[ijt], g = t.vdw_sig6[ijt];
                flr = e * g * (g / (10.0 * pow(rcut, 10)) - 1.0 / (4.0 * pow(rcut, 4)));
                ul = e * g * (g / (90.0 * pow(rcut, 9)) - 1.0 / (12.0 * pow(rcut, 3)));
                w.flr_lja[it] += sgn * xd[jt] * e * g * g / 10.0;
                w.flr_ljc[it] += sgn * xd[jt] * e * g / 4.0;
            } else if (t.vdw_kind[ijt] == 2) {
                double ro = 1.0 / t.vdw_rhoi[ijt];
                flr = ro * (ro + rcut) * t.vdw_a[ijt] * exp(-rcut / ro) - t.vdw_c[ijt] / (4.0 * pow(rcut, 4));
                ul = ro * ro * (2.0 * ro + rcut) * t.vdw_a[ijt] * exp(-rcut / ro) - t.vdw_c[ijt] / (12.0 * pow(rcut, 3));
                w.flr_bua[it] += sgn * xd[jt] * t.vdw_a[ijt] * ro;
                w.flr_buc[it] += sgn * xd[jt] * t.vdw_c[ijt] / 4.0;
                w.flr_bur[it] = ro;
            }
            if (sgn < 0.0) { w.flr_l[it] += flr * xd[jt]; w.ulr_l[it] += ul * xd[jt]; }
            else { w.flr_w[it] += flr * xd[jt]; w.ulr_w[it] += ul * xd[jt]; }
        }
    }
    double tp = 2.0 * s->c.pi;
    for (int it = 0; it <= MAX_TYPES; it++) {
        w.flr_lja[it] *= tp; w.flr_ljc[it] *= tp; w.flr_bua[it] *= tp; w.flr_buc[it] *= tp;
        w.flr_l[it] *= tp; w.flr_w[it] *= tp; w.ulr_l[it] *= tp; w.ulr_w[it] *= tp;
    }
    w.rcut = rcut;
    w.on = s->par.wall_lr;
}

// ---- k-set (src/forcem.f90:664-776): rows of contiguous kn, chunks of <= KC -------------
static void kset(mdf_system *s) {
    double lx = s->par.box[0], ly = s->par.box[1], lz = s->par.box[2];
    double lmin = std::min(lx, std::min(ly, lz));
    int kmax = s->par.kmax;
    double pipix = 2.0 * s->c.pi / lx, pipiy = 2.0 * s->c.pi / ly, pipiz = 2.0 * s->c.pi / lz;
    double alf4i = -1.0 / (4.0 * s->S.alfa * s->S.alfa);
    double fac = 2.0 * s->c.e * s->c.e * s->c.na / (s->c.eps0 * s->c.a2m * 1000.0 * lx * ly * lz);
    int klmax = (int)lround((double)(float)kmax * lx / lmin);
    int kmmax = (int)lround((double)(float)kmax * ly / lmin);
    int knmax = (int)lround((double)(float)kmax * lz / lmin);
    s->kvec.clear(); s->ak.clear(); s->rows.clear(); s->chunks.clear();
    int kmmin = 0, knmin = 1;
    for (int kl = 0; kl <= klmax; kl++) {
        for (int km = kmmin; km <= kmmax; km++) {
            bool found = false;
            int first = 0, last = 0;
            for (int kn = knmin; kn <= knmax; kn++) {
                float ksq = (float)(kl * kl + km * km) +
                            (float)((long long)kn * kn * kmax * kmax) / (float)((long long)knmax * knmax);
                if (ksq < (float)(kmax * kmax) + 2.01f) {
                    double rl = pipix * (double)(float)kl, rm = pipiy * (double)(float)km, rn = pipiz * (double)(float)kn;
                    double rksq = rl * rl + rm * rm + rn * rn;
                    if (!found) { first = kn; found = true; }
                    last = kn;
                    s->kvec.push_back(kl); s->kvec.push_back(km); s->kvec.push_back(kn);
                    s->ak.push_back(fac * exp(rksq * alf4i) / rksq);
                }
            }
            if (found) {
                int koff = (int)s->ak.size() - (last - first + 1);
                s->rows.push_back({kl, km, first, last, koff});
                for (int kn0 = first; kn0 <= last; kn0 += KC)
                    s->chunks.push_back({kl, km, kn0, std::min(KC, last - kn0 + 1), koff + (kn0 - first)});
            }
            knmin = -knmax;
        }
        kmmin = -kmmax;
    }
}

// sets up launch geometry for the current sizes
static void geometry(mdf_system *s) {
    int nm = s->top.n_move;
    int work = std::max(s->top.n_rig + (nm - s->top.n_rigid_sites), std::max(nm, s->n_lead));
    s->skin_blocks = std::max(1, std::min(1024, (work + 255) / 256));
    int per_sb = SF_BLOCK * 4;
    s->n_sb = std::max(1, (nm + per_sb - 1) / per_sb);
    int nrow = (int)s->rows.size();
    s->n_cg = std::max(1, std::min(nrow, (s->sf_target_blocks + s->n_sb - 1) / s->n_sb));   // row groups
    s->cpb = std::max(1, (nrow + s->n_cg - 1) / s->n_cg);                     // rows per group
    s->n_cg = std::max(1, (nrow + s->cpb - 1) / s->cpb);
    int nk = (int)s->ak.size();
    s->n_kblk = std::max(1, (nk + SFR_K - 1) / SFR_K);
    s->knmax = 0;
    for (size_t k = 2; k < s->kvec.size(); k += 3) s->knmax = std::max(s->knmax, std::abs(s->kvec[k]));
    s->n_nt = (2 * s->knmax + 1 + 15) / 16;
    s->klmax = 0; s->kmmax = 0;
    for (size_t k = 0; k + 2 < s->kvec.size(); k += 3) {
        s->klmax = std::max(s->klmax, std::abs(s->kvec[k]));
        s->kmmax = std::max(s->kmmax, std::abs(s->kvec[k + 1]));
    }
    // the FP64 form on the matrix cores: tables index kl, km and kn up to 63
    s->mma64 = s->mma64_allowed && s->par.precision == MDF_FP64 && nk > 0 && s->knmax <= 63 && s->klmax <= 63 &&
               s->kmmax <= 63 && sf64_lds_bytes(s->klmax, s->kmmax, (2 * s->knmax + 1 + 15) / 16) <= SF64_LDS_MAX;
    s->sf_mma = s->sf_mma_allowed && s->par.precision == MDF_FP32_ACC64 && nk > 0 && s->n_nt <= SFM_TILES_MAX;
    s->bp_mma = s->bp_mma_allowed && s->par.precision == MDF_FP32_ACC64 && nk > 0 && s->n_nt >= 1 &&
                s->n_nt <= SFM_TILES_MAX && nrow <= BPM_ROWS;
    // k_sf_mma: row groups of at most four 16-row tiles (spread evenly by the kernel); then as
    // many site blocks as fill the resident block slots in whole rounds, 32 sites a stage
    {
        const int n_mt = (nrow + 15) / 16;
        s->n_rg_mma = std::max(1, (n_mt + SFM_ROWS_PER_BLOCK / 16 - 1) / (SFM_ROWS_PER_BLOCK / 16));
        // sites per block: the least time on the busiest compute unit, its blocks running in
        // rounds of as many as stay resident (dynamic LDS sets that), a round of c blocks at a
        // throughput that grows slower than c (latency hiding)
        int sites = s->sfm_sites_env > 0 ? s->sfm_sites_env : SFM_SITES;
        while (sites > 32 && sfm_lds_bytes(sites, s->n_nt) > SFM_LDS_MAX) sites -= 32;
        if (s->sfm_sites_env <= 0 && nm > 0 && s->sf_mma) {
            auto thr = [](long c) { return c >= 4 ? 1.5 : c == 3 ? 1.3 : c == 2 ? 1.0 : 0.6; };
            double best = 1e300;
            for (int cand : {512, 384, 256, 128}) {
                if (sfm_lds_bytes(cand, s->n_nt) > SFM_LDS_MAX) continue;
                int occ = 0;
                if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, sfm_kernel(s->n_nt), 256,
                                                                 sfm_lds_bytes(cand, s->n_nt)) != hipSuccess ||
                    occ < 1) {
                    (void)hipGetLastError();
                    continue;
                }
                const long blocks = (long)((nm + cand - 1) / cand) * s->n_rg_mma;
                const long per_cu = (blocks + s->n_cu - 1) / s->n_cu;
                const long full = per_cu / occ, rem = per_cu % occ;
                const double cost = full * occ * cand / thr(occ) + (rem ? rem * cand / thr(rem) : 0.0);
                if (cost < best) { best = cost; sites = cand; }
            }
        }
        s->sfm_sites = sites;
        s->n_sb_mma = std::max(1, (nm + sites - 1) / sites);
        // k_bp_mma: many waves per block, over which the block's staging of each T tile spreads
        // (measured faster than balancing blocks over the compute units), two blocks resident
        // per unit: 4 waves per SIMD up to 7 kn tiles, 3 at 8 (its registers)
        s->bpm_waves = s->bpm_waves_env > 0 ? s->bpm_waves_env : (s->n_nt <= 7 ? BPM_WAVES_MAX : 6);
    }
    // back-projection row groups: about equal k counts, each at most BP_GROUP_K k-vectors (the
    // LDS stage), and enough groups that sites x groups fill the GPU
    int nsb_bp = std::max(1, (nm + 256 * BP_NS - 1) / (256 * BP_NS));
    int nrows = (int)s->rows.size();
    int nk_all = (int)s->ak.size();
    int want = std::max(1, std::min(std::min(nrows, BP_ROWGROUPS_MAX), (s->bp_target_blocks + nsb_bp - 1) / nsb_bp));
    want = std::max(want, (nk_all + BP_GROUP_K - 1) / BP_GROUP_K);
    s->bp_row_start.clear();
    s->bp_row_start.push_back(0);
    {
        const double per = (double)nk_all / want;
        int group_k = 0;
        for (int r = 0; r < nrows; r++) {
            const int len = s->rows[r].kn1 - s->rows[r].kn0 + 1;
            if (r > s->bp_row_start.back() && (group_k + len > BP_GROUP_K || group_k >= per)) {
                s->bp_row_start.push_back(r);
                group_k = 0;
            }
            group_k += len;
        }
        s->bp_row_start.push_back(nrows);
    }
    s->n_rg = (int)s->bp_row_start.size() - 1;
    s->rpg = 0;
    s->n_epblk = std::max(1, (nm + EPI_BLOCK - 1) / EPI_BLOCK);
    s->n_colblk = std::max(1, std::min(64, (s->n_lead + 1023) / 1024));
    s->colper = std::max(1, (s->n_lead + s->n_colblk - 1) / s->n_colblk);
}

static void drop_graphs(mdf_system *s);

// Tables and launch shapes of the FP64 matrix-core reciprocal kernels.
// Structure factor: rows sorted by kn extent E (widest first, stable), 16-row tiles with their
// column tiles ceil((2 E + 1) / 16), slots of (row tile, column tile) dealt to the waves of
// each pass, partial columns per row 2 E + 1 in sorted order.  Back-projection: M-tiles of 4 rows in the
// reference order, K-steps ceil((2 E_max + 1) / 4) each, row groups of about equal K-steps.
// the captured step graph forks the k-space chain onto a second stream from this many moving
// sites (MDF_FORK_KSPACE=1 / 0 force it on / off)
constexpr int FORK_MIN_SITES = 16384;
static bool fork_planned(const mdf_system *s) {
    return s->fork_env > 0 || (s->fork_env < 0 && s->top.n_move >= FORK_MIN_SITES);
}

static int setup_mma64(mdf_system *s) {
    const int nr = (int)s->rows.size(), nm = s->top.n_move;
    std::vector<int> E(nr);
    for (int r = 0; r < nr; r++) E[r] = std::max(std::abs(s->rows[r].kn0), std::abs(s->rows[r].kn1));
    // structure factor tables
    std::vector<int> sorted(nr), sorted_of(nr);
    for (int r = 0; r < nr; r++) sorted[r] = r;
    std::stable_sort(sorted.begin(), sorted.end(), [&](int a, int b) { return E[a] > E[b]; });
    for (int i = 0; i < nr; i++) sorted_of[sorted[i]] = i;
    const int n_rt = (nr + 15) / 16;
    std::vector<int2> rowkk((size_t)n_rt * 16, make_int2(0, 0));
    std::vector<int> tile_nt(n_rt), pcoff(nr + 1, 0);
    for (int i = 0; i < nr; i++) {
        const KRow &w = s->rows[sorted[i]];
        rowkk[i] = make_int2(w.kl, w.km);
        pcoff[i + 1] = pcoff[i] + 2 * E[sorted[i]] + 1;
    }
    for (int t = 0; t < n_rt; t++) tile_nt[t] = (2 * E[sorted[16 * t]] + 1 + 15) / 16;
    // work slots (row tile, column tile) in row-tile order, dealt to passes of SF64_WAVES waves
    // of up to SF64_SLOTS consecutive slots; within a pass wave w gets q or q + 1 slots, the
    // larger counts first, so that the SIMDs (wave w on SIMD w % 4) get about equal numbers
    std::vector<int2> slots;
    for (int t = 0; t < n_rt; t++)
        for (int c = 0; c < tile_nt[t]; c++) slots.push_back(make_int2(t, c));
    const int n_sl = (int)slots.size(), per_pass = SF64_WAVES * SF64_SLOTS;
    const int npass = std::min(n_sl, std::max(std::max(1, s->s64_pass_env), (n_sl + per_pass - 1) / per_pass));
    std::vector<int4> waves((size_t)npass * SF64_WAVES, make_int4(0, 0, 0, 0));
    std::vector<int> pass_ncol(npass, 16);
    for (int p = 0; p < npass; p++) {
        const int b0 = (int)((long)p * n_sl / npass), b1 = (int)((long)(p + 1) * n_sl / npass);
        const int q = (b1 - b0) / SF64_WAVES, rem = (b1 - b0) % SF64_WAVES;
        int k = b0;
        for (int w = 0; w < SF64_WAVES; w++) {
            const int nw = q + (w < rem ? 1 : 0);   // at most SF64_SLOTS
            int v[SF64_SLOTS] = {0, 0, 0};
            int code = nw;
            for (int u = 0; u < nw; u++, k++) {
                v[u] = (slots[k].x << 8) | slots[k].y;
                if (u > 0 && slots[k].x == slots[k - 1].x) code |= u == 1 ? 16 : 32;
                pass_ncol[p] = std::max(pass_ncol[p], 16 * (slots[k].y + 1));
            }
            waves[(size_t)p * SF64_WAVES + w] = make_int4(v[0], v[1], v[2], code);
        }
    }
    s->s64_nt = tile_nt[0];
    s->s64_npass = npass;
    s->s64_pcols = pcoff[nr];
    // sites per block: at most one block per compute unit (a block fills a unit's 16 wave slots
    // at 128 registers), at least four stages a block so that partial rows stay few.  When the
    // step graph forks, the blocks cover about 7/8 of the units: the skin test and the list
    // kernels of the other stream run on the rest instead of waiting for the whole structure
    // factor (slab16k: 220 blocks of 224 sites, 3657 -> 3872 steps/s, the kernel itself
    // 53 -> 60 us)
    {
        int sites = s->s64_sites_env;
        if (sites <= 0) {
            const long units = fork_planned(s) ? ((long)s->n_cu * 7) / 8 : (long)s->n_cu;
            const long nsb = std::max(1L, (units + npass - 1) / npass);
            sites = (int)((nm + nsb - 1) / nsb);
            sites = std::max(4 * SF64_SUB, SF64_SUB * ((sites + SF64_SUB - 1) / SF64_SUB));
        }
        s->s64_sites = sites;
        s->s64_nsb = std::max(1, (nm + sites - 1) / sites);
    }
    // back-projection tables
    const int n_mt = (nr + 3) / 4;
    std::vector<int2> bptile(n_mt);
    int koff = 0, nks_max = 1;
    for (int m = 0; m < n_mt; m++) {
        int emax = 0;
        for (int r = 4 * m; r < std::min(nr, 4 * m + 4); r++) emax = std::max(emax, E[r]);
        const int nks = (2 * emax + 1 + 3) / 4;
        bptile[m] = make_int2(koff, nks);
        koff += nks;
        nks_max = std::max(nks_max, nks);
    }
    std::vector<int4> rinfo(nr);
    for (int r = 0; r < nr; r++) {
        const KRow &w = s->rows[r];
        const bool cont = r >= 4 && s->rows[r - 4].kl == w.kl && s->rows[r - 4].km + 4 == w.km;
        rinfo[r] = make_int4(w.kl, w.km, cont ? 0 : 1, 0);
    }
    // site groups per wave: three (their operands at the exact K-step count, 9..20, fill the
    // registers: two waves per SIMD, and 2 n_cu blocks of 192 sites run in one round at
    // slab16k), else two up to 20 K-steps, else one
    const int ng_def = nks_max <= 20 ? (nks_max >= 9 ? 3 : 2) : 1;
    s->b64_ng = s->b64_ng_env > 0 ? s->b64_ng_env : ng_def;
    if (s->b64_ng == 3 && (nks_max < 9 || nks_max > 20)) s->b64_ng = ng_def;   // not instantiated
    if (s->b64_ng == 2 && nks_max > 32) s->b64_ng = 1;
    s->b64_nks = s->b64_ng == 3 ? nks_max : 4 * ((nks_max + 3) / 4);
    s->b64_sites = 4 * 16 * s->b64_ng;   // BP64_WAVES waves of 16 NG sites
    s->b64_nmt = n_mt;
    const int nsb_bp = std::max(1, (nm + s->b64_sites - 1) / s->b64_sites);
    int nrg = s->b64_rg_env > 0 ? s->b64_rg_env : std::max(1, std::min(8, (2 * s->n_cu + nsb_bp - 1) / nsb_bp));
    nrg = std::min(nrg, n_mt);
    std::vector<int2> rg;
    {
        int m = 0;
        for (int gi = 0; gi < nrg; gi++) {
            const long target = ((long)(gi + 1) * koff + nrg - 1) / nrg;
            int m1 = m;
            while (m1 < n_mt && (bptile[m1].x < target || m1 == m)) m1++;
            if (gi == nrg - 1) m1 = n_mt;
            rg.push_back(make_int2(m, m1));
            m = m1;
        }
    }
    s->b64_nrg = (int)rg.size();
    HIPCHK(upload(s->d_s64_rowkk, rowkk.data(), rowkk.size()));
    HIPCHK(upload(s->d_s64_waves, waves.data(), waves.size()));
    HIPCHK(upload(s->d_s64_pncol, pass_ncol.data(), pass_ncol.size()));
    HIPCHK(upload(s->d_s64_pcoff, pcoff.data(), pcoff.size()));
    HIPCHK(upload(s->d_s64_sorted_of, sorted_of.data(), sorted_of.size()));
    HIPCHK(upload(s->d_b64_tile, bptile.data(), bptile.size()));
    HIPCHK(upload(s->d_b64_rinfo, rinfo.data(), rinfo.size()));
    HIPCHK(upload(s->d_b64_rg, rg.data(), rg.size()));
    HIPCHK(s->d_s64_part.alloc((size_t)s->s64_nsb * s->s64_pcols));
    HIPCHK(s->d_b64_tpk.alloc((size_t)std::max(koff, 1) * 64));
    HIPCHK(hipMemset(s->d_b64_tpk.p, 0, sizeof(double) * s->d_b64_tpk.n));   // rows past the last stay zero
    return MDF_OK;
}

// launch shapes of the reciprocal kernels for the current targets: row groups, their device
// table and the partial-force buffer they need; captured graphs hold the old shapes
static int apply_geometry(mdf_system *s) {
    geometry(s);
    s->d_bp_row_start.release();
    HIPCHK(upload(s->d_bp_row_start, s->bp_row_start.data(), s->bp_row_start.size()));
    const size_t need = (size_t)s->n_rg * 3 * std::max(s->top.n_move, 1);
    if (s->fpart.n < need) {
        s->fpart.release();
        HIPCHK(s->fpart.alloc(need));
    }
    drop_graphs(s);
    return MDF_OK;
}

static void destroy_events(mdf_system *s) {
    if (s->ev_fork) { (void)hipEventDestroy(s->ev_fork); s->ev_fork = nullptr; }
    if (s->ev_join) { (void)hipEventDestroy(s->ev_join); s->ev_join = nullptr; }
    for (int r = 0; r < mdf_system::NRING; r++) {
        for (int k = 0; k <= N_STAGES; k++)
            if (s->ring[r][k]) (void)hipEventDestroy(s->ring[r][k]);
        for (int k = 0; k < N_KT; k++)
            for (int e = 0; e < 2; e++)
                if (s->kring[r][k][e]) (void)hipEventDestroy(s->kring[r][k][e]);
    }
}

static void drop_graphs(mdf_system *s) {
    for (int r = 0; r < mdf_system::NRING; r++)
        if (s->graph[r]) { (void)hipGraphExecDestroy(s->graph[r]); s->graph[r] = nullptr; }
}

static int alloc_lists(mdf_system *s, int cap_mm, int cap_mw) {
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    drop_graphs(s);
    s->nbr_mm.release(); s->nbr_mw.release();
    HIPCHK(s->nbr_mm.alloc((size_t)s->n_lead * cap_mm));
    HIPCHK(s->nbr_mw.alloc((size_t)s->n_lead * std::max(cap_mw, 1)));
    s->S.nbr_mm = s->nbr_mm.p; s->S.cap_mm = cap_mm;
    s->S.nbr_mw = s->nbr_mw.p; s->S.cap_mw = std::max(cap_mw, 1);
    return MDF_OK;
}

// list-building stage of the launch sequence (skin test, then a build when due)
static void launch_lists(mdf_system *s, int force_rebuild) {
    DevSys &S = s->S;
    hipStream_t st = s->stream;
    k_skin<<<s->skin_blocks, 256, 0, st>>>(S, s->dip_part.p);
    k_decide<<<1, 256, 0, st>>>(S, s->dip_part.p, s->skin_blocks, s->dipz.p, force_rebuild);
    if (s->ev) (void)hipEventRecord(s->ev[1], st);
    int lb = std::max(1, (s->n_lead + 255) / 256);
    k_cell_count<<<lb, 256, 0, st>>>(S, s->lead_cell.p);
    k_cell_scan_build<<<1, 1024, 0, st>>>(S);
    k_cell_fill<<<lb, 256, 0, st>>>(S, s->lead_cell.p, s->cell_items_tmp.p, s->cell_xyz_tmp.p);
    k_cell_order<<<s->order_blocks, 256, 0, st>>>(S, s->cell_items_tmp.p, s->cell_xyz_tmp.p);
    k_build<<<s->build_blocks, BUILD_THREADS, 0, st>>>(S);   // one block per cell
}

// phase factors of the reciprocal kernels: 2 pi / L (widened pi) and the same over pi
struct RecipArgs {
    double cux, cuy, cuz, pipix, pipiy, pipiz;
};
static RecipArgs recip_args(const mdf_system *s) {
    RecipArgs a;
    a.cux = 2.0 * s->c.pi / s->par.box[0] / M_PI;
    a.cuy = 2.0 * s->c.pi / s->par.box[1] / M_PI;
    a.cuz = 2.0 * s->c.pi / s->par.box[2] / M_PI;
    a.pipix = 2.0 * s->c.pi / s->par.box[0];
    a.pipiy = 2.0 * s->c.pi / s->par.box[1];
    a.pipiz = 2.0 * s->c.pi / s->par.box[2];
    return a;
}

// launch of a kernel timed by its own dispatch: with stage timing on, hipExtLaunchKernelGGL
// records the slot's start and stop events with the kernel itself (no launch gaps in the
// interval); otherwise an ordinary launch.  KERN is parenthesised when it has template
// arguments.
#define MDF_TLAUNCH(s_, slot, KERN, grid, block, shm, st, ...)                                                  \
    do {                                                                                                        \
        if ((s_)->kev) {                                                                                        \
            (s_)->kev_used[slot] = true;                                                                        \
            hipExtLaunchKernelGGL(KERN, dim3(grid), dim3(block), (uint32_t)(shm), st, (s_)->kev[slot][0],      \
                                  (s_)->kev[slot][1], 0u, __VA_ARGS__);                                         \
        } else {                                                                                                \
            KERN<<<grid, block, shm, st>>>(__VA_ARGS__);                                                        \
        }                                                                                                       \
    } while (0)

template <int G, bool ST, int NMC>
static void launch_pair_g(mdf_system *s, bool f32, bool vir, bool acc) {
    const int blocks = std::max(1, (s->n_lead + (256 / G) - 1) / (256 / G));
    hipStream_t st = s->stream;
    DevSys &S = s->S;
#define MDF_PAIR(R, V, A) MDF_TLAUNCH(s, KT_PAIR, (k_pair<R, G, V, A, ST, NMC>), blocks, 256, 0, st, S)
    if (f32) {
        if (vir) { if (acc) MDF_PAIR(float, true, true); else MDF_PAIR(float, true, false); }
        else { if (acc) MDF_PAIR(float, false, true); else MDF_PAIR(float, false, false); }
    } else {
        if (vir) { if (acc) MDF_PAIR(double, true, true); else MDF_PAIR(double, true, false); }
        else { if (acc) MDF_PAIR(double, false, true); else MDF_PAIR(double, false, false); }
    }
#undef MDF_PAIR
}

// real-space pair kernel with the handle's lanes per lead, loop form and erfc form
static void launch_pair(mdf_system *s) {
    const bool f32 = s->par.precision == MDF_FP32_ACC64, vir = s->par.virial != 0;
    const bool acc = s->par.erfc_mode == MDF_ERFC_ACCURATE;
    const bool st = s->pair_strided;
    // one moving lead type: a single energy class of moving partners
    if (s->slot_type.size() <= 1) {
        switch (s->pair_g) {
            case 8: if (st) launch_pair_g<8, true, 1>(s, f32, vir, acc); else launch_pair_g<8, false, 1>(s, f32, vir, acc); break;
            case 32: if (st) launch_pair_g<32, true, 1>(s, f32, vir, acc); else launch_pair_g<32, false, 1>(s, f32, vir, acc); break;
            default: if (st) launch_pair_g<16, true, 1>(s, f32, vir, acc); else launch_pair_g<16, false, 1>(s, f32, vir, acc); break;
        }
        return;
    }
    constexpr int M = MAX_LEAD_TYPES;
    switch (s->pair_g) {
        case 8: if (st) launch_pair_g<8, true, M>(s, f32, vir, acc); else launch_pair_g<8, false, M>(s, f32, vir, acc); break;
        case 32: if (st) launch_pair_g<32, true, M>(s, f32, vir, acc); else launch_pair_g<32, false, M>(s, f32, vir, acc); break;
        default: if (st) launch_pair_g<16, true, M>(s, f32, vir, acc); else launch_pair_g<16, false, M>(s, f32, vir, acc); break;
    }
}

// the scheme's device state when multiple time steps are on (the k-sum kernels skip the
// steps it serves without the full sum), else null
static const MtsState *mts_state(const mdf_system *s) { return s->mts_on ? s->mts_st.p : nullptr; }

static MtsArgs mts_args(mdf_system *s) {
    MtsArgs a;
    memset(&a, 0, sizeof a);
    a.st = s->mts_st.p;
    a.interval = s->mts_interval;
    a.train_frames = s->mts_train_frames;
    a.model = s->mts_model;
    a.n_sblk = s->sg_n_sblk;
    a.n_tblk = s->sg_n_tblk;
    a.train_iters = s->mts_train_iters;
    a.train_stride = s->sg_train_stride;
    a.tol = s->mts_tol;
    a.lr = s->mts_lr;
    a.feat = s->sg_feat.p; a.vec = s->sg_vec.p; a.sg = s->sg_out.p; a.sg_prev = s->sg_out_prev.p;
    a.feat_prev = s->sg_feat_prev.p; a.vec_prev = s->sg_vec_prev.p; a.frec_prev = s->sg_frec_prev.p;
    a.delta_prev = s->sg_delta_prev.p;
    a.pos4 = s->sg_pos4.p;
    a.msize = s->sg_msize.p;
    a.delta = s->sg_delta.p; a.frec = s->sg_frec.p; a.fmts = s->sg_fmts.p;
    a.rprev = s->sg_rprev.p; a.fprev = s->sg_fprev.p;
    a.params = s->sg_params.p; a.adam_m = s->sg_m.p; a.adam_v = s->sg_v.p; a.gsum = s->sg_gsum.p;
    a.wb = s->sg_wb.p; a.norm = s->sg_norm.p; a.gpart = s->sg_gpart.p;
    a.jpart = s->sg_jpart.p; a.epart = s->sg_epart.p;
    return a;
}

static void launch_sf(mdf_system *s) {
    const int nk = (int)s->ak.size(), nrows = (int)s->rows.size(), nm = s->top.n_move;
    if (nk == 0 || nm == 0) return;
    const RecipArgs a = recip_args(s);
    DevSys &S = s->S;
    if (s->mma64) {
        const dim3 grid(s->s64_nsb, s->s64_npass);
        const size_t lds = sf64_lds_bytes(s->klmax, s->kmmax, s->s64_nt);
#define SF64_LAUNCH(NT)                                                                                         \
    case NT:                                                                                                    \
        MDF_TLAUNCH(s, KT_SF, k_sf64<NT>, grid, 64 * SF64_WAVES, lds, s->stream, S.x, S.y, S.z, S.q, nm, a.cux,  \
                    a.cuy, a.cuz, s->d_s64_rowkk.p, s->d_s64_waves.p, s->d_s64_pncol.p, nrows, s->klmax, s->kmmax, \
                    s->d_s64_pcoff.p, s->s64_pcols, s->d_s64_part.p, s->s64_sites, mts_state(s));                \
        break;
        switch (s->s64_nt) {
            SF64_LAUNCH(1) SF64_LAUNCH(2) SF64_LAUNCH(3) SF64_LAUNCH(4) SF64_LAUNCH(5) SF64_LAUNCH(6) SF64_LAUNCH(7)
            SF64_LAUNCH(8)
        }
#undef SF64_LAUNCH
        return;
    }
    if (s->sf_mma) {
        const dim3 grid(s->n_sb_mma, s->n_rg_mma);
        const size_t lds = sfm_lds_bytes(s->sfm_sites, s->n_nt);
        float2 *part = reinterpret_cast<float2 *>(s->sf_partial.p);
#define SFM_LAUNCH(NT)                                                                                          \
    case NT:                                                                                                    \
        MDF_TLAUNCH(s, KT_SF, k_sf_mma<NT>, grid, 256, lds, s->stream, S.x, S.y, S.z, S.q, nm, a.cux, a.cuy,     \
                    a.cuz, s->d_rows.p, nrows, s->knmax, s->n_nt, part, nk, s->sfm_sites, mts_state(s));          \
        break;
        switch (s->n_nt) {
            SFM_LAUNCH(1) SFM_LAUNCH(2) SFM_LAUNCH(3) SFM_LAUNCH(4) SFM_LAUNCH(5) SFM_LAUNCH(6) SFM_LAUNCH(7)
            SFM_LAUNCH(8)
        }
#undef SFM_LAUNCH
        return;
    }
    dim3 g1(s->n_sb, s->n_cg);
    if (s->par.precision == MDF_FP32_ACC64)
        MDF_TLAUNCH(s, KT_SF, (k_sf<float, 4>), g1, SF_BLOCK, 0, s->stream, S.x, S.y, S.z, S.q, nm, a.cux, a.cuy,
                    a.cuz, s->d_rows.p, nrows, s->cpb, s->sf_partial.p, nk, mts_state(s));
    else
        MDF_TLAUNCH(s, KT_SF, (k_sf<double, 4>), g1, SF_BLOCK, 0, s->stream, S.x, S.y, S.z, S.q, nm, a.cux, a.cuy,
                    a.cuz, s->d_rows.p, nrows, s->cpb, s->sf_partial.p, nk, mts_state(s));
}

static void launch_bp(mdf_system *s) {
    const int nk = (int)s->ak.size(), nm = s->top.n_move;
    if (nk == 0 || nm == 0) return;
    const RecipArgs a = recip_args(s);
    DevSys &S = s->S;
    if (s->mma64) {
        const dim3 grid((nm + s->b64_sites - 1) / s->b64_sites, s->b64_nrg);
#define BP64_LAUNCH(NKS, NG)                                                                                    \
    case NKS * 4 + NG:                                                                                          \
        MDF_TLAUNCH(s, KT_BP, (k_bp64<NKS, NG>), grid, 256, 0, s->stream, S.x, S.y, S.z, S.q, nm, a.cux, a.cuy,  \
                    a.cuz, s->d_b64_rinfo.p, (int)s->rows.size(), s->d_b64_tile.p, s->d_b64_rg.p, s->d_b64_tpk.p, \
                    a.pipix, a.pipiy, a.pipiz, s->fpart.p, mts_state(s));                                        \
        break;
        switch (s->b64_nks * 4 + s->b64_ng) {
            BP64_LAUNCH(4, 1) BP64_LAUNCH(8, 1) BP64_LAUNCH(12, 1) BP64_LAUNCH(16, 1) BP64_LAUNCH(20, 1)
            BP64_LAUNCH(24, 1) BP64_LAUNCH(28, 1) BP64_LAUNCH(32, 1) BP64_LAUNCH(4, 2) BP64_LAUNCH(8, 2)
            BP64_LAUNCH(12, 2) BP64_LAUNCH(16, 2) BP64_LAUNCH(20, 2) BP64_LAUNCH(24, 2) BP64_LAUNCH(28, 2)
            BP64_LAUNCH(32, 2) BP64_LAUNCH(9, 3) BP64_LAUNCH(10, 3) BP64_LAUNCH(11, 3) BP64_LAUNCH(12, 3)
            BP64_LAUNCH(13, 3) BP64_LAUNCH(14, 3) BP64_LAUNCH(15, 3) BP64_LAUNCH(16, 3) BP64_LAUNCH(17, 3)
            BP64_LAUNCH(18, 3) BP64_LAUNCH(19, 3) BP64_LAUNCH(20, 3)
        }
#undef BP64_LAUNCH
        return;
    }
    if (s->bp_mma) {
        const int ns = 16 * s->bpm_waves;
        const dim3 grid((nm + ns - 1) / ns), block(64 * s->bpm_waves);
#define BPM_LAUNCH(NT)                                                                                          \
    case NT:                                                                                                    \
        MDF_TLAUNCH(s, KT_BP, k_bp_mma<NT>, grid, block, 0, s->stream, S.x, S.y, S.z, S.q, nm, a.cux, a.cuy,     \
                    a.cuz, s->d_rowkk.p, s->d_tile_ks.p, (int)s->rows.size(), s->knmax, s->n_nt, s->d_tpack.p,   \
                    s->d_tscale.p, a.pipix, a.pipiy, a.pipiz, s->fpart.p, mts_state(s));                         \
        break;
        switch (s->n_nt) {
            BPM_LAUNCH(1) BPM_LAUNCH(2) BPM_LAUNCH(3) BPM_LAUNCH(4) BPM_LAUNCH(5) BPM_LAUNCH(6) BPM_LAUNCH(7)
            BPM_LAUNCH(8)
        }
#undef BPM_LAUNCH
        return;
    }
    dim3 g2(std::max(1, (nm + 256 * BP_NS - 1) / (256 * BP_NS)), s->n_rg);
    if (s->par.precision == MDF_FP32_ACC64)
        MDF_TLAUNCH(s, KT_BP, (k_bp<float, BP_NS>), g2, 256, 0, s->stream, S.x, S.y, S.z, S.q, nm, a.cux, a.cuy,
                    a.cuz, s->d_rows.p, s->d_bp_row_start.p, s->d_t.p, a.pipix, a.pipiy, a.pipiz, s->fpart.p,
                    mts_state(s));
    else
        MDF_TLAUNCH(s, KT_BP, (k_bp<double, BP_NS>), g2, 256, 0, s->stream, S.x, S.y, S.z, S.q, nm, a.cux, a.cuy,
                    a.cuz, s->d_rows.p, s->d_bp_row_start.p, s->d_t.p, a.pipix, a.pipiy, a.pipiz, s->fpart.p,
                    mts_state(s));
}

// the scalars the final reduction needs
static FinalArgs final_args(const mdf_system *s) {
    FinalArgs A;
    memset(&A, 0, sizeof A);
    A.n_colblk = s->n_colblk;
    A.n_slots = (int)s->slot_type.size();
    for (int k = 0; k < A.n_slots; k++) A.slot_type[k] = s->slot_type[k];
    for (int c = 0; c < NCLS; c++) A.cls_type[c] = s->cls_type[c];
    A.n_kblk = s->ak.size() > 0 ? ((s->bp_mma || s->mma64) ? (int)s->rows.size() : s->n_kblk) : 0;
    A.n_epblk = s->n_epblk;
    A.e_self = s->uself;
    A.e_recip_ww = s->ewqWW;
    A.crf0 = s->crf0;
    A.ew3dc = s->par.ew3dc;
    A.lx = s->par.box[0];
    A.ly = s->par.box[1];
    A.na_scale = s->c.na * s->c.e30;
    A.bar = s->c.e5;
    return A;
}

// the k-space chain on the handle's current stream: structure factor, its reduction (S, T,
// energies, the back-projection's operand) and the back-projection
static void launch_kspace(mdf_system *s) {
    DevSys &S = s->S;
    hipStream_t st = s->stream;
    const int nm = s->top.n_move;
    const RecipArgs ra = recip_args(s);
    const int nk = (int)s->ak.size();
    launch_sf(s);
    if (s->ev) (void)hipEventRecord(s->ev[5], st);
    if (nk > 0) {
        double lamb = 1.0 / (4.0 * S.alfa * S.alfa);
        const int n_sb = nm > 0 ? (s->sf_mma ? s->n_sb_mma : s->n_sb) : 0;
        const float2 *p32 = s->sf_mma && nm > 0 ? reinterpret_cast<const float2 *>(s->sf_partial.p) : nullptr;
        if (s->mma64)    // one block per row: S, T, energies and the back-projection operand
            MDF_TLAUNCH(s, KT_SFRED, k_sf64_reduce, (int)s->rows.size(), SFR_THREADS, 0, st, s->d_s64_part.p,
                        nm > 0 ? s->s64_nsb : 0, s->s64_pcols, s->d_s64_pcoff.p, s->d_s64_sorted_of.p, s->d_rows.p,
                        s->d_swall.p, s->d_ak.p, ra.pipix, ra.pipiy, ra.pipiz, lamb, s->d_s.p, s->d_t.p,
                        s->kpart.p, s->d_b64_tile.p, s->d_b64_tpk.p, mts_state(s));
        else if (s->bp_mma)   // one block per row: T split with per-row scales for k_bp_mma
            MDF_TLAUNCH(s, KT_SFRED, k_sf_reduce_rows, (int)s->rows.size(), SFR_THREADS, 0, st, s->sf_partial.p,
                        p32, n_sb, nk, s->d_rows.p, s->d_swall.p, s->d_ak.p, ra.pipix, ra.pipiy, ra.pipiz, lamb,
                        s->d_s.p, s->d_t.p, s->kpart.p, mts_state(s), s->d_tpack.p, s->d_tscale.p, 16 * s->n_nt,
                        s->knmax);
        else
            MDF_TLAUNCH(s, KT_SFRED, k_sf_reduce, s->n_kblk, SFR_THREADS, 0, st, s->sf_partial.p, n_sb, nk,
                        s->d_swall.p, s->d_ak.p, s->d_kvec.p, ra.pipix, ra.pipiy, ra.pipiz, lamb, s->d_s.p, s->d_t.p,
                        s->kpart.p, mts_state(s), nullptr, nullptr, 0, 0, p32);
    }
    if (s->ev) (void)hipEventRecord(s->ev[6], st);
    launch_bp(s);
}

static void launch_forces(mdf_system *s) {
    DevSys &S = s->S;
    hipStream_t st = s->stream;
    const int nm = s->top.n_move;
    const RecipArgs ra = recip_args(s);
    const bool mts = s->mts_on;
    const MtsArgs ma = mts ? mts_args(s) : MtsArgs{};
    if (s->ev) (void)hipEventRecord(s->ev[2], st);
    if (mts) k_sg_control<<<1, 64, 0, st>>>(ma, s->mts_again);
    launch_pair(s);
    if (s->ev) (void)hipEventRecord(s->ev[3], st);
    if (mts) {
        if (s->mts_model) {
            k_sg_pack<<<std::max(1, (s->top.n_total + 255) / 256), 256, 0, st>>>(S, ma);
            k_sg_features<SG_FEAT_G><<<std::max(1, (s->n_lead + 256 / SG_FEAT_G - 1) / (256 / SG_FEAT_G)), 256, 0,
                                       st>>>(S, ma);
            k_sg_norm<<<1, 1024, 0, st>>>(ma, nm);
            k_sg_infer<<<s->sg_n_iblk, 256, 0, st>>>(ma, nm, 0, s->flags.p);
        }
        k_sg_combine<<<s->sg_n_sblk, 256, 0, st>>>(S, ma);
    }
    if (s->ev) (void)hipEventRecord(s->ev[4], st);
    const int nk = (int)s->ak.size();
    const dim3 gc(NCOL, s->n_colblk);
    if (s->kspace_forked) {
        // the pair kernel's per-lead columns are summed while the k-space chain still runs
        k_colsum<<<gc, 256, 0, st>>>(S, s->lead_slot.p, s->colper, s->colpart.p);
        (void)hipStreamWaitEvent(st, s->ev_join, 0);
    } else {
        launch_kspace(s);
    }
    if (s->ev) (void)hipEventRecord(s->ev[7], st);
    const int n_groups = (nk > 0 && nm > 0) ? (s->mma64 ? s->b64_nrg : s->bp_mma ? 1 : s->n_rg) : 0;
    k_epilogue<<<s->n_epblk, EPI_BLOCK, 0, st>>>(S, s->fpart.p, n_groups, s->d_wl.p, s->par.ew3dc, s->crf0,
                                                 s->dipz.p, s->eppart.p, s->fcom.p, s->torque.p, mts_state(s),
                                                 s->sg_fmts.p, s->obs_on ? s->obs_efi.p : nullptr);
    if (s->obs_on) {
        k_obs_ngc<<<std::max(1, (s->n_lead + 15) / 16), 256, 0, st>>>(S, s->obs);
        k_obs_efipos<<<1, 64, 0, st>>>(S, s->obs);
    }
    if (s->ang_on)
        k_obs_ngang<<<(S.n_angles + 255) / 256, 256, 0, st>>>(S, s->ang_type.p, s->n_atypes, s->ang_hist.p);
    if (s->ev) (void)hipEventRecord(s->ev[8], st);
    {
        const FinalArgs A = final_args(s);
        if (!s->kspace_forked) k_colsum<<<gc, 256, 0, st>>>(S, s->lead_slot.p, s->colper, s->colpart.p);
        k_final<<<1, 256, 0, st>>>(S, A, s->colpart.p, s->kpart.p, s->eppart.p, s->dipz.p, s->e_vdw.p, s->e_coul.p,
                                   s->d_scal.p);
    }
    if (mts) {
        // energies of a surrogate step; on a full-sum step: target, judge, training, correction
        k_sg_finish<<<1, 64, 0, st>>>(S, ma, s->d_scal.p);
        k_sg_exact<<<s->sg_n_sblk, 256, 0, st>>>(S, ma, s->fpart.p, n_groups);
        k_sg_judge<<<1, 256, 0, st>>>(S, ma);
        if (s->mts_model) {
            for (int it = 0; it < s->mts_train_iters; it++) {
                k_sg_train<<<s->sg_n_tblk, 256, 0, st>>>(ma, nm, it, s->flags.p);
                k_sg_adam<<<(SG_GP + 255) / 256, 256, 0, st>>>(ma, s->sg_n_tblk, it, s->flags.p);
            }
            k_sg_infer<<<s->sg_n_iblk, 256, 0, st>>>(ma, nm, 1, s->flags.p);
            k_sg_infer<<<s->sg_n_iblk, 256, 0, st>>>(ma, nm, 2, s->flags.p);
        }
        k_sg_delta<<<s->sg_n_sblk, 256, 0, st>>>(S, ma);
        k_sg_post<<<1, 64, 0, st>>>(S, ma);
    }
    if (s->ev) (void)hipEventRecord(s->ev[9], st);
}

static void drain_slot(mdf_system *s, int slot) {
    if (!s->ring_used[slot]) return;
    hipEvent_t *e = s->ring[slot];
    bool ok = hipEventSynchronize(e[N_STAGES]) == hipSuccess;
    float tot = 0.f;
    ok = ok && hipEventElapsedTime(&tot, e[0], e[N_STAGES]) == hipSuccess;
    float ms[N_STAGES];
    bool staged = ok && s->slot_staged[slot];
    for (int k = 0; k < N_STAGES && staged; k++) {
        ms[k] = 0.f;
        staged = hipEventElapsedTime(&ms[k], e[k], e[k + 1]) == hipSuccess;
    }
    if (ok) { s->step_sum += tot; s->step_steps++; }
    for (int k = 0; k < N_KT; k++) {
        if (!s->kring_used[slot][k]) continue;
        s->kring_used[slot][k] = false;
        float km = 0.f;
        if (ok && hipEventElapsedTime(&km, s->kring[slot][k][0], s->kring[slot][k][1]) == hipSuccess) {
            s->kt_sum[k] += km;
            s->kt_n[k]++;
        } else {
            (void)hipGetLastError();
        }
    }
    if (staged) {
        for (int k = 0; k < N_STAGES; k++) { s->stage_sum[k] += ms[k]; s->stage_ms[k] = ms[k]; }
        s->stage_steps++;
    }
    if (!ok || (s->slot_staged[slot] && !staged)) {
        s->timing_failures++;
        (void)hipGetLastError();  // timing is best effort: do not leave a sticky error behind
    }
    s->ring_used[slot] = false;
}

static void drain_all(mdf_system *s) {
    // oldest first, so stage_ms ends as the last step's
    for (int d = 0; d < mdf_system::NRING; d++) {
        long long n = s->step_no - mdf_system::NRING + d;
        if (n < 0) continue;
        drain_slot(s, (int)(n % mdf_system::NRING));
    }
}

// ---- Widom insertions -------------------------------------------------------------------

// the reference's generator: seed = mod(8192 seed, 67099547), r = seed 1.4903230270690e-8
// (src/rndmod.f90:28-35); the product stays below 2^53, so fmod is exact
struct ChemRng {
    double seed;
    double next() {
        seed = std::fmod(8192.0 * seed, 67099547.0);
        return seed * 1.4903230270690e-8;
    }
};

// one call's geometry in the reference's draw order (src/forcem.f90:1258-1272,1338-1369);
// the generator state after each insertion's position is kept, since an overlap ends the
// call right there
static int chem_draw(mdf_system *s, int *n_rot) {
    const int n = s->chem_p.n_insert;
    const double lx = s->par.box[0], ly = s->par.box[1];
    const double lwmin = s->chem_lwmin, lwmax = s->chem_lwmax, facbin = s->chem_facbin;
    const double aco = s->chem_aco, asi = s->chem_asi, lo = s->chem_lo, pi = s->c.pi;
    ChemRng g{s->chem_seed};
    int nr = 0;
    for (int ic = 0; ic < n; ic++) {
        const double ox = (g.next() - 0.5) * lx;
        const double oy = (g.next() - 0.5) * ly;
        double oz;
        long guard = 0;
        do {
            oz = ((g.next() - 0.5) * lwmax) * 2.0;
            if (++guard > 100000000L) return fail(MDF_ERR_ARG, "chempot: no insertion height found");
        } while (!(std::fabs(oz) > lwmin));
        s->ch_seed_pos[ic] = g.seed;
        const int il = (int)((std::fabs(oz) - lwmin) * facbin) + 1;
        const int irot = il > s->chem_p.bin_mid ? CHEM_ROT_MAX : 1;
        s->h_ins[3 * ic] = ox; s->h_ins[3 * ic + 1] = oy; s->h_ins[3 * ic + 2] = oz;
        s->h_il[ic] = il;
        s->h_insrot[ic] = nr;
        const double o[3] = {ox, oy, oz};
        for (int ir = 0; ir < irot; ir++) {
            double h1[3], ri;
            do {
                h1[0] = g.next() - 0.5;
                h1[1] = g.next() - 0.5;
                h1[2] = g.next() - 0.5;
                ri = h1[0] * h1[0] + h1[1] * h1[1] + h1[2] * h1[2];
            } while (!(ri < 0.25 && ri > lo));
            ri = 1.0 / std::sqrt(ri);
            for (int d = 0; d < 3; d++) h1[d] = h1[d] * ri;
            ri = 1.0 / std::sqrt(h1[0] * h1[0] + h1[1] * h1[1]);
            const double h2[3] = {-h1[1] * ri, h1[0] * ri, 0.0};
            const double mr[3] = {h2[1] * h1[2] - h2[2] * h1[1], h2[2] * h1[0] - h2[0] * h1[2],
                                  h2[0] * h1[1] - h2[1] * h1[0]};
            const double alph = 2.0 * pi * g.next();
            const double ca = std::cos(alph), sa = std::sin(alph);
            for (int d = 0; d < 3; d++) {
                s->h_rot[6 * nr + d] = (o[d] + aco * h1[d]) + asi * (ca * h2[d] + sa * mr[d]);
                s->h_rot[6 * nr + 3 + d] = o[d] + h1[d];
            }
            s->h_rotins[nr] = ic;
            nr++;
        }
    }
    s->h_insrot[n] = nr;
    s->ch_seed_end = g.seed;
    *n_rot = nr;
    return MDF_OK;
}

// the last call's outcome: the generator continues after the first overlapping insertion's
// position, or after the whole call
static int chem_resolve(mdf_system *s) {
    if (!s->ch_pending) return MDF_OK;
    HIPCHK(hipEventSynchronize(s->ch_done));
    s->ch_pending = false;
    const int *st = s->h_status;
    if (st[2]) return MDF_OK;   // dropped with its step, whose list overflowed: the redo inserts again
    if (st[1]) {
        const int had = s->chem.cap;
        int rc = MDF_OK;
        const int cap = std::max(had, (int)(st[3] * 1.25) + 16);
        DBuf<int> nb;
        DBuf<double> nd;
        if (nb.alloc((size_t)s->chem_p.n_insert * cap) != hipSuccess ||
            nd.alloc((size_t)3 * s->chem_p.n_insert * cap) != hipSuccess) {
            nb.release(); nd.release();
            return fail(MDF_ERR_HIP, "chempot: partner lists of %d", cap);
        }
        s->ch_nbr.release(); s->ch_nbr_d.release();
        s->ch_nbr = nb; s->ch_nbr_d = nd;
        s->chem.nbr = nb.p; s->chem.nbr_d = nd.p; s->chem.cap = cap;
        rc = fail(MDF_ERR_LIST_OVERFLOW, "chempot: an insertion had %d partners within rcut, more than %d: that "
                  "step's insertions were dropped (capacity now %d)", st[3], had, cap);
        return rc;
    }
    const int first = st[0];
    s->chem_seed = first < s->chem_p.n_insert ? s->ch_seed_pos[first] : s->ch_seed_end;
    return MDF_OK;
}

static int chem_enqueue(mdf_system *s) {
    int rc = chem_resolve(s);
    if (rc) return rc;
    int nr = 0;
    rc = chem_draw(s, &nr);
    if (rc) return rc;
    const int n = s->chem_p.n_insert;
    hipStream_t st = s->stream;
    HIPCHK(hipMemcpyAsync(s->ch_ins.p, s->h_ins, sizeof(double) * 3 * n, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(s->ch_il.p, s->h_il, sizeof(int) * n, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(s->ch_insrot.p, s->h_insrot, sizeof(int) * (n + 1), hipMemcpyHostToDevice, st));
    if (nr > 0) {
        HIPCHK(hipMemcpyAsync(s->ch_rotins.p, s->h_rotins, sizeof(int) * nr, hipMemcpyHostToDevice, st));
        HIPCHK(hipMemcpyAsync(s->ch_rot.p, s->h_rot, sizeof(double) * 6 * nr, hipMemcpyHostToDevice, st));
    }
    ChemArgs c = s->chem;
    c.n_rot = nr;
    k_chem_scan<<<n, 256, 0, st>>>(s->S, c);
    k_chem_first<<<1, 256, 0, st>>>(s->S, c);
    if (nr > 0) k_chem_rot<<<nr, 256, s->chem_lds, st>>>(s->S, c);
    k_chem_bsum<<<(n + 255) / 256, 256, 0, st>>>(c);
    k_chem_accum<<<std::min(s->chem_p.n_bins, 1024), 256, 0, st>>>(c);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(s->h_status, s->ch_status.p, sizeof(int) * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipEventRecord(s->ch_done, st));
    s->ch_pending = true;
    return MDF_OK;
}

// one step: the fixed launch sequence, replayed from a captured graph when possible (a
// forced rebuild goes through direct launches)
static int run_step(mdf_system *s, int force_rebuild) {
    if (s->rebuild_next) { force_rebuild = 1; s->rebuild_next = false; }
    int slot = (int)(s->step_no % mdf_system::NRING);
    drain_slot(s, slot);
    s->ev = s->ring[slot];
    s->ring_used[slot] = true;
    s->step_no++;
    hipEvent_t *ev = s->ev;
    s->mts_again = force_rebuild;   // a forced rebuild only ever redoes an overflowed step
    if (s->use_graphs && !force_rebuild && !s->stage_timing) {
        // the graph holds no events (HIP cannot time events recorded during capture); the
        // step is bracketed by events recorded outside it
        s->slot_staged[slot] = false;
        if (!s->graph[0]) {
            hipGraph_t g = nullptr;
            s->ev = nullptr;
            // the k-space chain needs only the gathered positions: forked onto the second
            // stream first, it runs beside the skin test, list build and pair kernel
            // (surrogate steps keep one stream: their kernels interleave with the k-sum).
            // Forked by default only from FORK_MIN_SITES moving sites: below, the chains are
            // short and many handles stepping side by side (each a stream pair on the few
            // hardware queues) measured slower forked.
            bool fork = fork_planned(s) && !s->mts_on && !s->ak.empty() && s->top.n_move > 0;
            // the second stream is made on first use, before the capture (handles that never
            // fork keep one stream: streams are spread over the hardware queues in the order
            // they are made)
            if (fork && !s->stream2) {
                if (hipStreamCreateWithFlags(&s->stream2, hipStreamNonBlocking) != hipSuccess ||
                    hipEventCreateWithFlags(&s->ev_fork, hipEventDisableTiming) != hipSuccess ||
                    hipEventCreateWithFlags(&s->ev_join, hipEventDisableTiming) != hipSuccess) {
                    (void)hipGetLastError();
                    fork = false;
                }
            }
            HIPCHK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeRelaxed));
            if (fork && s->stream2 && s->ev_fork && s->ev_join) {
                hipStream_t main = s->stream;
                (void)hipEventRecord(s->ev_fork, main);
                (void)hipStreamWaitEvent(s->stream2, s->ev_fork, 0);
                s->stream = s->stream2;
                launch_kspace(s);
                s->stream = main;
                (void)hipEventRecord(s->ev_join, s->stream2);
                s->kspace_forked = true;
            }
            launch_lists(s, 0);
            launch_forces(s);
            s->kspace_forked = false;
            hipError_t ec = hipStreamEndCapture(s->stream, &g);
            s->ev = ev;
            if (ec != hipSuccess) return fail(MDF_ERR_HIP, "graph capture: %s", hipGetErrorString(ec));
            ec = hipGraphInstantiate(&s->graph[0], g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ec != hipSuccess) return fail(MDF_ERR_HIP, "graph instantiate: %s", hipGetErrorString(ec));
        }
        (void)hipEventRecord(ev[0], s->stream);
        HIPCHK(hipGraphLaunch(s->graph[0], s->stream));
        (void)hipEventRecord(ev[N_STAGES], s->stream);
    } else {
        s->slot_staged[slot] = true;
        if (s->stage_timing) {
            s->kev = s->kring[slot];
            s->kev_used = s->kring_used[slot];
        }
        (void)hipEventRecord(ev[0], s->stream);
        launch_lists(s, force_rebuild || s->stage_timing == 2);
        launch_forces(s);
        s->kev = nullptr;
        s->kev_used = nullptr;
    }
    HIPCHK(hipGetLastError());
    s->pending = true;
    if (s->chem_on) {
        // after the step, outside its graph: the host draws each call's geometry
        int rc = chem_enqueue(s);
        if (rc) return rc;
    }
    return MDF_OK;
}

// the overflow flags are sticky on the device: the host clears them once it has read them
// (no kernel of this handle is in flight then)
static int clear_overflow(mdf_system *s) {
    const int zero[3] = {0, 0, 0};
    HIPCHK(hipMemcpy(&s->flags.p->overflow_mm, zero, sizeof zero, hipMemcpyHostToDevice));
    if (s->h_mirror) s->h_mirror[0] = 0;
    return MDF_OK;
}

// wait, read flags, grow lists and redo the step on overflow (unless the capacity was fixed).
// redo: 0 none, 1 redo the last step, 2 the same and the overflow came from an earlier step
// whose successors (enqueued with mdf_compute_device) also ran on the truncated rows.
static int finish_step(mdf_system *s, int *redo) {
    *redo = 0;
    HIPCHK(hipStreamSynchronize(s->stream));
    s->pending = false;
    HIPCHK(hipMemcpy(&s->last_flags, s->flags.p, sizeof(DevFlags), hipMemcpyDeviceToHost));
    const DevFlags f = s->last_flags;
    s->last_rebuilt = f.rebuild;
    if (f.overflow_mm || f.overflow_mw) {
        int rc = clear_overflow(s);
        if (rc) return rc;
        const int late = f.steps - f.overflow_step;   // steps run on the truncated rows after it
        if (s->fixed_cap) {
            s->rebuild_next = true;   // the rows stay truncated: the next step builds them again
            return fail(MDF_ERR_LIST_OVERFLOW, "nnlist_make(): nnlist too big! need %d moving / %d frozen partners per "
                        "lead, capacity %d / %d%s", f.overflow_mm, f.overflow_mw, s->S.cap_mm, s->S.cap_mw,
                        late > 0 ? " (and later steps ran on the truncated rows)" : "");
        }
        int cmm = std::max(s->S.cap_mm, (int)(f.overflow_mm * 1.25) + 16);
        int cmw = std::max(s->S.cap_mw, (int)(f.overflow_mw * 1.25) + 16);
        rc = alloc_lists(s, cmm, cmw);
        if (rc) return rc;
        *redo = late > 0 ? 2 : 1;
    }
    drain_all(s);
    return MDF_OK;
}

static int late_overflow_error(const mdf_system *s) {
    return fail(MDF_ERR_LIST_OVERFLOW, "a pair-list row overflowed %d step(s) before the last one; the steps since ran "
                "on truncated rows and their device results are wrong (capacity grown to %d / %d, the last step "
                "redone: the outputs are valid)", s->last_flags.steps - s->last_flags.overflow_step, s->S.cap_mm,
                s->S.cap_mw);
}

static int copy_outputs(mdf_system *s, mdf_outputs *o) {
    const int nm = s->top.n_move, nr = s->top.n_rig;
    if (o->f_site) {
        HIPCHK(hipMemcpy(o->f_site, s->fx.p, sizeof(double) * nm, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(o->f_site + nm, s->fy.p, sizeof(double) * nm, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(o->f_site + 2 * nm, s->fz.p, sizeof(double) * nm, hipMemcpyDeviceToHost));
    }
    if (o->f_com && nr) HIPCHK(hipMemcpy(o->f_com, s->fcom.p, sizeof(double) * 3 * nr, hipMemcpyDeviceToHost));
    if (o->torque && nr) HIPCHK(hipMemcpy(o->torque, s->torque.p, sizeof(double) * 3 * nr, hipMemcpyDeviceToHost));
    std::vector<double> ev(s->nijt + 1), ec(s->nijt + 1);
    HIPCHK(hipMemcpy(ev.data(), s->e_vdw.p, sizeof(double) * (s->nijt + 1), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(ec.data(), s->e_coul.p, sizeof(double) * (s->nijt + 1), hipMemcpyDeviceToHost));
    if (o->e_vdw) memcpy(o->e_vdw, ev.data() + 1, sizeof(double) * s->nijt);
    if (o->e_coul_real) memcpy(o->e_coul_real, ec.data() + 1, sizeof(double) * s->nijt);
    Scalars sc;
    HIPCHK(hipMemcpy(&sc, s->d_scal.p, sizeof(Scalars), hipMemcpyDeviceToHost));
    o->e_recip = sc.e_recip; o->e_recip_ll = sc.e_recip_ll; o->e_recip_wl = sc.e_recip_wl;
    o->e_recip_ww = sc.e_recip_ww; o->e_self = sc.e_self; o->e_slab = sc.e_slab; o->e_wall_lr = sc.e_wall_lr;
    o->e_conf = sc.e_conf; o->e_avex = sc.e_avex; o->press_wall = sc.press;
    memcpy(o->f_wall, sc.f_wall, sizeof sc.f_wall);
    memcpy(o->virial, sc.virial, sizeof sc.virial);
    o->list_rebuilt = s->last_rebuilt;
    o->n_pairs_listed = s->last_flags.n_pairs_mm;
    o->n_wall_pairs_listed = s->last_flags.n_pairs_mw;
    o->n_kvec = (int)s->ak.size();
    o->n_list_builds = s->last_flags.n_builds;
    if (sc.nonfinite) return fail(MDF_ERR_NONFINITE, "non-finite configurational energy");
    return MDF_OK;
}

extern "C" {

const char *mdf_last_error(void) { return g_err.c_str(); }

int mdf_create(const mdf_topology *top, const mdf_params *par, const double *x, const double *y, const double *z,
               const double *cx, const double *cy, const double *cz, mdf_system **out) {
    if (!top || !par || !out || !x || !y || !z) return fail(MDF_ERR_ARG, "null argument");
    *out = nullptr;
    const mdf_topology &t = *top;
    if (t.n_types < 3 || t.n_types > MAX_TYPES) return fail(MDF_ERR_ARG, "n_types %d outside 3..%d", t.n_types, MAX_TYPES);
    if (t.n_move < 0 || t.n_total < t.n_move || t.n_rigid_sites != 3 * t.n_rig || t.n_liquid_sites > t.n_move ||
        t.n_rigid_sites > t.n_liquid_sites)
        return fail(MDF_ERR_ARG, "inconsistent site counts");
    if (t.n_rig > 0 && (!cx || !cy || !cz)) return fail(MDF_ERR_ARG, "COM arrays required");
    if (t.n_constraints < 0 || (t.n_constraints > 0 && (!t.con_a || !t.con_b || !t.con_len)))
        return fail(MDF_ERR_ARG, "constraint arrays required");
    if (t.n_angles < 0 || (t.n_angles > 0 && (!t.ang_i || !t.ang_j || !t.ang_k || !t.ang_kt || !t.ang_theta)))
        return fail(MDF_ERR_ARG, "angle arrays required");
    if (par->rcut <= 0 || par->skin < 0 || par->kmax < 0) return fail(MDF_ERR_ARG, "bad cut-off/skin/kmax");
    if (par->box[0] <= 0 || par->box[1] <= 0 || par->box[2] <= 0) return fail(MDF_ERR_ARG, "bad box");
    int dev_count = 0;
    if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count == 0)
        return fail(MDF_ERR_NO_DEVICE, "no HIP device visible");
    if (par->device < 0 || par->device >= dev_count) return fail(MDF_ERR_ARG, "device %d not present", par->device);
    const int nit = t.n_types, nijt = nit * (nit + 1) / 2;
    for (int i = 0; i < t.n_total; i++)
        if (t.site_type[i] < 1 || t.site_type[i] > nit) return fail(MDF_ERR_ARG, "site %d has type %d", i, t.site_type[i]);
    for (int k = 0; k < nijt; k++)
        if (t.vdw_kind[k] < 0 || t.vdw_kind[k] > 2)
            return fail(MDF_ERR_VDW_TYPE, "force_real(): unknown interaction type %d for pair %d", t.vdw_kind[k], k + 1);
    for (int it = 0; it < nit; it++)
        if (t.nsi[it] < 0 || t.nsi[it] > 3) return fail(MDF_ERR_ARG, "nsi(%d) = %d: at most 3 sites per molecule", it + 1, t.nsi[it]);
    for (int m = 0; m < t.n_rig; m++)
        if (t.nsi[t.site_type[3 * m] - 1] != 3 || t.nsi[t.site_type[3 * m + 1] - 1] != 0 ||
            t.nsi[t.site_type[3 * m + 2] - 1] != 0)
            return fail(MDF_ERR_ARG, "rigid molecule %d is not a 3-site lead + 2 carried sites", m);

    (void)hipGetLastError();  // clear an error left by an earlier, unrelated call
    mdf_system *s = new mdf_system();
    s->top = t;
    s->par = *par;
    s->c = make_consts(par->const_mode);
    s->device = par->device;
    s->nijt = nijt;
    s->n_wall = t.n_total - t.n_move;
    {
        const char *g = getenv("MDF_GRAPHS");
        if (g && g[0] == '0') s->use_graphs = false;
        const char *pr = getenv("MDF_PAIR_STRIDED");
        if (pr && pr[0] == '1') s->pair_strided = true;
        const char *pco = getenv("MDF_PAIR_CELL_ORDER");
        if (pco && pco[0] == '0') s->pair_cell_order = 0;
        const char *fk = getenv("MDF_FORK_KSPACE");
        if (fk) s->fork_env = fk[0] == '0' ? 0 : 1;
        const char *sm = getenv("MDF_SF_MMA");
        if (sm && sm[0] == '0') s->sf_mma_allowed = false;
        const char *bm = getenv("MDF_BP_MMA");
        if (bm && bm[0] == '0') s->bp_mma_allowed = false;
        const char *m64 = getenv("MDF_MMA64");
        if (m64 && m64[0] == '0') s->mma64_allowed = false;
        const char *s64 = getenv("MDF_SF64_SITES");
        if (s64 && atoi(s64) > 0) s->s64_sites_env = SF64_SUB * ((atoi(s64) + SF64_SUB - 1) / SF64_SUB);
        const char *p64 = getenv("MDF_SF64_PASSES");
        if (p64 && atoi(p64) > 0) s->s64_pass_env = std::min(16, atoi(p64));
        const char *g64 = getenv("MDF_BP64_NG");
        if (g64 && atoi(g64) >= 1 && atoi(g64) <= 3) s->b64_ng_env = atoi(g64);
        const char *r64 = getenv("MDF_BP64_GROUPS");
        if (r64 && atoi(r64) > 0) s->b64_rg_env = std::min(64, atoi(r64));
        // k_sf_mma splits q exp(i phi) of every moving site into FP16 hi/lo parts, which hold
        // magnitudes below 65504: far larger charges than any physical one (|q| > 2^14 e) take
        // the vector-ALU structure factor instead
        for (int i = 0; i < t.n_move; i++)
            if (!(std::fabs(t.charge[i]) < 16384.0)) { s->sf_mma_allowed = false; break; }
        const char *sb = getenv("MDF_SF_BLOCKS");
        if (sb && atoi(sb) > 0) s->sf_target_blocks = atoi(sb);
        const char *bb = getenv("MDF_BP_BLOCKS");
        if (bb && atoi(bb) > 0) s->bp_target_blocks = atoi(bb);
        const char *ss = getenv("MDF_SFM_SITES");
        if (ss && atoi(ss) > 0) s->sfm_sites_env = std::min(SFM_SITES, 32 * ((atoi(ss) + 31) / 32));
        const char *bw = getenv("MDF_BPM_WAVES");
        if (bw && atoi(bw) > 0) s->bpm_waves_env = std::min(BPM_WAVES_MAX, atoi(bw));
    }
    hipError_t he = hipSetDevice(par->device);
    if (he != hipSuccess) { delete s; return fail(MDF_ERR_HIP, "hipSetDevice: %s", hipGetErrorString(he)); }
    {   // compute units size the matrix-core launches
        int cu = 0;
        if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, par->device) == hipSuccess && cu > 0)
            s->n_cu = cu;
        (void)hipGetLastError();
    }
    int rc = MDF_OK;
#define TRY(x)                        \
    do {                              \
        rc = (x);                     \
        if (rc) { goto fail_out; }    \
    } while (0)
#define HTRY(x)                                                                                        \
    do {                                                                                               \
        hipError_t e_ = (x);                                                                           \
        if (e_ != hipSuccess) { rc = fail(MDF_ERR_HIP, "%s: %s", #x, hipGetErrorString(e_)); goto fail_out; } \
    } while (0)
    {
        HTRY(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
        HTRY(hipEventCreateWithFlags(&s->in_event, hipEventDisableTiming));
        for (int r = 0; r < mdf_system::NRING; r++)
            for (int k = 0; k <= N_STAGES; k++) HTRY(hipEventCreate(&s->ring[r][k]));
        for (int r = 0; r < mdf_system::NRING; r++)
            for (int k = 0; k < N_KT; k++)
                for (int e = 0; e < 2; e++) HTRY(hipEventCreate(&s->kring[r][k][e]));
        // leads, lead-type slots, energy classes
        s->h_itype.assign(t.site_type, t.site_type + t.n_total);
        std::vector<int> lead_types;
        for (int i = 0; i < t.n_move; i++) {
            int ti = t.site_type[i];
            if (t.nsi[ti - 1] > 0) {
                s->h_lead.push_back(i);
                int slot = -1;
                for (size_t k = 0; k < lead_types.size(); k++)
                    if (lead_types[k] == ti) slot = (int)k;
                if (slot < 0) {
                    if ((int)lead_types.size() >= MAX_LEAD_TYPES) { rc = fail(MDF_ERR_ARG, "more than %d moving lead types", MAX_LEAD_TYPES); goto fail_out; }
                    lead_types.push_back(ti);
                    slot = (int)lead_types.size() - 1;
                }
                s->h_lead_slot.push_back(slot);
            }
        }
        s->n_lead = (int)s->h_lead.size();
        s->slot_type = lead_types;
        std::vector<int> cls(nit + 1, 0);
        s->cls_type.assign(NCLS, 0);
        for (size_t k = 0; k < lead_types.size(); k++) { cls[lead_types[k]] = (int)k; s->cls_type[k] = lead_types[k]; }
        // crystal types: the last two (src/sysdef.f90:418); classes 4 and 5
        cls[nit - 1] = MAX_LEAD_TYPES; cls[nit] = MAX_LEAD_TYPES + 1;
        s->cls_type[MAX_LEAD_TYPES] = nit - 1; s->cls_type[MAX_LEAD_TYPES + 1] = nit;
        for (int w = t.n_move; w < t.n_total; w++)
            if (t.site_type[w] < nit - 1) { rc = fail(MDF_ERR_ARG, "frozen site %d has a non-crystal type", w); goto fail_out; }
        // 1-based tables
        std::vector<int> nsi1(nit + 1, 0), ij1((nit + 1) * (nit + 1), 0), vk1(nijt + 1, 0);
        std::vector<double> e41(nijt + 1, 0), s61(nijt + 1, 0), a1(nijt + 1, 0), r1(nijt + 1, 0), c1(nijt + 1, 0);
        for (int it = 1; it <= nit; it++) nsi1[it] = t.nsi[it - 1];
        for (int it = 1; it <= nit; it++)
            for (int jt = 1; jt <= nit; jt++) ij1[it * (nit + 1) + jt] = t.ijtype[(it - 1) * nit + (jt - 1)];
        for (int k = 0; k < nijt; k++) {
            vk1[k + 1] = t.vdw_kind[k]; e41[k + 1] = t.vdw_eps4[k]; s61[k + 1] = t.vdw_sig6[k];
            a1[k + 1] = t.vdw_a[k]; r1[k + 1] = t.vdw_rhoi[k]; c1[k + 1] = t.vdw_c[k];
        }
        // constraint partners (CSR over moving sites)
        std::vector<int> cptr(t.n_move + 1, 0), cidx;
        {
            std::vector<std::vector<int>> adj(t.n_move);
            for (int k = 0; k < t.n_constraints; k++) {
                int a = t.con_a[k], b = t.con_b[k];
                if (a < 0 || b < 0 || a >= t.n_total || b >= t.n_total) { rc = fail(MDF_ERR_ARG, "constraint %d out of range", k); goto fail_out; }
                if (a < t.n_move) adj[a].push_back(b);
                if (b < t.n_move) adj[b].push_back(a);
            }
            for (int i = 0; i < t.n_move; i++) { cptr[i + 1] = cptr[i] + (int)adj[i].size(); cidx.insert(cidx.end(), adj[i].begin(), adj[i].end()); }
        }
        DevSys &S = s->S;
        memset(&S, 0, sizeof S);
        S.n_total = t.n_total; S.n_move = t.n_move; S.n_rig = t.n_rig; S.n_rigid_sites = t.n_rigid_sites;
        S.n_liquid_sites = t.n_liquid_sites; S.n_lead = s->n_lead; S.n_wall = s->n_wall; S.n_types = nit;
        S.n_ltype = t.n_ltype; S.nijt = nijt;
        HTRY(upload(s->itype, t.site_type, t.n_total));
        HTRY(upload(s->q, t.charge, t.n_total));
        HTRY(upload(s->nsi, nsi1.data(), nsi1.size()));
        HTRY(upload(s->ijtype, ij1.data(), ij1.size()));
        HTRY(upload(s->vkind, vk1.data(), vk1.size()));
        HTRY(upload(s->eps4, e41.data(), e41.size()));
        HTRY(upload(s->sig6, s61.data(), s61.size()));
        HTRY(upload(s->aa, a1.data(), a1.size()));
        HTRY(upload(s->rhoi, r1.data(), r1.size()));
        HTRY(upload(s->cc, c1.data(), c1.size()));
        HTRY(upload(s->cls, cls.data(), cls.size()));
        HTRY(upload(s->lead, s->h_lead.data(), s->h_lead.size()));
        HTRY(upload(s->lead_slot, s->h_lead_slot.data(), s->h_lead_slot.size()));
        // the pair kernel's static inputs: lead slot and sites per lead of every moving site,
        // crystal class of every frozen site, vdW parameters per (lead slot, partner class)
        {
            std::vector<int> pm(std::max(t.n_move, 1), 0), wc(std::max(s->n_wall, 1), 0);
            for (int l = 0; l < s->n_lead; l++) {
                const int i = s->h_lead[l];
                const int ns = t.nsi[t.site_type[i] - 1];
                if (ns > 3 || i + ns > t.n_move) { rc = fail(MDF_ERR_ARG, "lead %d carries %d sites (at most 3 moving)", i, ns); goto fail_out; }
                pm[i] = s->h_lead_slot[l] | (ns << 4);
            }
            for (int w = 0; w < s->n_wall; w++) wc[w] = t.site_type[t.n_move + w] - (nit - 1);
            std::vector<double> tab((size_t)MAX_LEAD_TYPES * NCLS * PTAB_W, 0.0);
            for (int si = 0; si < (int)s->slot_type.size(); si++)
                for (int c = 0; c < NCLS; c++) {
                    const int tj = s->cls_type[c];
                    if (tj <= 0) continue;
                    const int ijt = ij1[s->slot_type[si] * (nit + 1) + tj];
                    double *e = &tab[((size_t)si * NCLS + c) * PTAB_W];
                    e[0] = vk1[ijt]; e[1] = e41[ijt]; e[2] = s61[ijt]; e[3] = a1[ijt]; e[4] = r1[ijt]; e[5] = c1[ijt];
                }
            HTRY(upload(s->pmeta, pm.data(), pm.size()));
            HTRY(upload(s->wcls, wc.data(), wc.size()));
            HTRY(upload(s->ptab, tab.data(), tab.size()));
            HTRY(s->prec.alloc((size_t)std::max(t.n_move, 1) * PREC));
            HTRY(hipMemset(s->prec.p, 0, sizeof(double) * s->prec.n));
        }
        HTRY(upload(s->con_ptr, cptr.data(), cptr.size()));
        HTRY(upload(s->con_idx, cidx.data(), cidx.size()));
        // angle terms: the reference adds each angle's force to its sites j and k only
        // (src/forcem.f90:603-614); a CSR per moving site lets one thread per site sum its
        // own terms without atomics
        {
            std::vector<int> aptr(t.n_move + 1, 0), ainc, aijk(3 * (size_t)std::max(t.n_angles, 0));
            std::vector<std::vector<int>> inc(t.n_move);
            for (int a = 0; a < t.n_angles; a++) {
                const int ii = t.ang_i[a], jj = t.ang_j[a], kk = t.ang_k[a];
                if (ii < 0 || jj < 0 || kk < 0 || ii >= t.n_total || jj >= t.n_total || kk >= t.n_total ||
                    !std::isfinite(t.ang_kt[a]) || !std::isfinite(t.ang_theta[a])) {
                    rc = fail(MDF_ERR_ARG, "angle %d invalid", a);
                    goto fail_out;
                }
                aijk[3 * a] = ii; aijk[3 * a + 1] = jj; aijk[3 * a + 2] = kk;
                if (jj < t.n_move) inc[jj].push_back(2 * a);
                if (kk < t.n_move) inc[kk].push_back(2 * a + 1);
            }
            for (int i = 0; i < t.n_move; i++) {
                aptr[i + 1] = aptr[i] + (int)inc[i].size();
                ainc.insert(ainc.end(), inc[i].begin(), inc[i].end());
            }
            HTRY(upload(s->ang_ptr, aptr.data(), aptr.size()));
            HTRY(upload(s->ang_inc, ainc.data(), ainc.size()));
            HTRY(upload(s->ang_ijk, aijk.data(), aijk.size()));
            HTRY(upload(s->ang_kt, t.ang_kt, (size_t)std::max(t.n_angles, 0)));
            HTRY(upload(s->ang_th, t.ang_theta, (size_t)std::max(t.n_angles, 0)));
            S.n_angles = std::max(t.n_angles, 0);
            S.ang_ptr = s->ang_ptr.p; S.ang_inc = s->ang_inc.p; S.ang_ijk = s->ang_ijk.p;
            S.ang_kt = s->ang_kt.p; S.ang_th = s->ang_th.p;
        }
        S.itype = s->itype.p; S.q = s->q.p; S.nsi = s->nsi.p; S.ijtype = s->ijtype.p; S.vkind = s->vkind.p;
        S.eps4 = s->eps4.p; S.sig6 = s->sig6.p; S.aa = s->aa.p; S.rhoi = s->rhoi.p; S.cc = s->cc.p; S.cls = s->cls.p;
        S.pmeta = s->pmeta.p; S.wcls = s->wcls.p; S.ptab = s->ptab.p; S.prec = s->prec.p;
        S.pair_cell_order = s->pair_cell_order;
        S.lead = s->lead.p; S.con_ptr = s->con_ptr.p; S.con_idx = s->con_idx.p;
        // positions (all sites once; moving part per call)
        HTRY(upload(s->x, x, t.n_total)); HTRY(upload(s->y, y, t.n_total)); HTRY(upload(s->z, z, t.n_total));
        int nr = std::max(t.n_rig, 1);
        std::vector<double> zero(nr, 0.0);
        HTRY(upload(s->cx, t.n_rig ? cx : zero.data(), nr)); HTRY(upload(s->cy, t.n_rig ? cy : zero.data(), nr));
        HTRY(upload(s->cz, t.n_rig ? cz : zero.data(), nr));
        HTRY(s->ocx.alloc(nr)); HTRY(s->ocy.alloc(nr)); HTRY(s->ocz.alloc(nr));
        HTRY(s->ox.alloc(std::max(t.n_move, 1))); HTRY(s->oy.alloc(std::max(t.n_move, 1))); HTRY(s->oz.alloc(std::max(t.n_move, 1)));
        HTRY(s->fx.alloc(std::max(t.n_move, 1))); HTRY(s->fy.alloc(std::max(t.n_move, 1))); HTRY(s->fz.alloc(std::max(t.n_move, 1)));
        HTRY(s->part.alloc((size_t)NCOL * std::max(s->n_lead, 1)));
        S.x = s->x.p; S.y = s->y.p; S.z = s->z.p; S.cx = s->cx.p; S.cy = s->cy.p; S.cz = s->cz.p;
        {   // frozen sites for the pair kernel, once (their positions do not change)
            std::vector<double4> wr(std::max(s->n_wall, 1), make_double4(0.0, 0.0, 0.0, 0.0));
            for (int w = 0; w < s->n_wall; w++)
                wr[w] = make_double4(x[t.n_move + w], y[t.n_move + w], z[t.n_move + w], t.charge[t.n_move + w]);
            HTRY(upload(s->wrec, wr.data(), wr.size()));
            S.wrec = s->wrec.p;
        }
        S.ocx = s->ocx.p; S.ocy = s->ocy.p; S.ocz = s->ocz.p; S.ox = s->ox.p; S.oy = s->oy.p; S.oz = s->oz.p;
        S.fx = s->fx.p; S.fy = s->fy.p; S.fz = s->fz.p; S.part = s->part.p;
        // box and constants (force_init, src/forcem.f90:58-60; config_init src/config.f90:67-69)
        S.lx = par->box[0]; S.ly = par->box[1]; S.lz = par->box[2];
        S.lxi = 1.0 / S.lx; S.lyi = 1.0 / S.ly; S.lzi = 1.0 / S.lz;
        S.lwh = 0.5 * t.wall_lw;
        S.zper_list = par->z_mode == MDF_Z_PERIODIC;
        S.zper_force = par->z_mode != MDF_Z_OPEN;
        S.rcutsq = par->rcut * par->rcut;
        S.skin2 = par->skin * par->skin;
        S.spii = 2.0 / sqrt(s->c.pi);
        S.alfa = par->alpha_mode == MDF_ALPHA_REFERENCE ? s->c.c309 / par->rcut : par->alpha;
        S.facq = s->c.k4pi * s->c.e * s->c.e * s->c.na / (1000.0 * s->c.a2m);
        S.a1 = s->c.a1; S.a2 = s->c.a2; S.a3 = s->c.a3; S.a4 = s->c.a4; S.a5 = s->c.a5; S.p = s->c.p;
        S.erfc_mode = par->erfc_mode;
        {
            double off = 0.0;
            if (t.n_rig > 0)
                off = sqrt((x[0] - cx[0]) * (x[0] - cx[0]) + (y[0] - cy[0]) * (y[0] - cy[0]) + (z[0] - cz[0]) * (z[0] - cz[0]));
            double rs = par->rcut + par->skin;
            rs = rs + 2.0 * off;
            S.rskinsq = rs * rs;
        }
        // cell grid: cells at least the list radius wide
        {
            double rs = sqrt(S.rskinsq) * (1.0 + 1e-9);
            S.g.ncx = std::max(1, (int)floor(S.lx / rs));
            S.g.ncy = std::max(1, (int)floor(S.ly / rs));
            S.g.csx = S.lx / S.g.ncx; S.g.csy = S.ly / S.g.ncy;
            S.g.zper = S.zper_list;
            S.g.ncz_fixed = std::max(1, (int)floor(S.lz / rs));
            S.g.csz_fixed = S.lz / S.g.ncz_fixed;
            S.g.ncz_max = S.g.zper ? S.g.ncz_fixed : 1024;
            s->ncells_xy = S.g.ncx * S.g.ncy;
            s->n_cells_cap = s->ncells_xy * S.g.ncz_max;
            S.n_cells_cap = s->n_cells_cap;
            HTRY(s->cell_cnt.alloc(s->n_cells_cap + 1)); HTRY(s->cell_start.alloc(s->n_cells_cap + 1));
            HTRY(s->cell_cursor.alloc(s->n_cells_cap + 1));
            HTRY(s->cell_items.alloc(std::max(s->n_lead, 1))); HTRY(s->cell_xyz.alloc(std::max(s->n_lead, 1)));
            HTRY(s->cell_items_tmp.alloc(std::max(s->n_lead, 1)));
            HTRY(s->cell_xyz_tmp.alloc(std::max(s->n_lead, 1)));
            HTRY(s->lead_cell.alloc(std::max(s->n_lead, 1)));
            S.cell_cnt = s->cell_cnt.p; S.cell_start = s->cell_start.p; S.cell_items = s->cell_items.p;
            S.cell_xyz = s->cell_xyz.p;
            // blocks of the list build and the order pass: the cells of the initial lead z
            // range (the blocks stride over the cells a later build has)
            int ncz0 = S.g.ncz_fixed;
            if (!S.g.zper && s->n_lead > 0) {
                double lo = z[s->h_lead[0]], hi = lo;
                for (int l = 0; l < s->n_lead; l++) { lo = std::min(lo, z[s->h_lead[l]]); hi = std::max(hi, z[s->h_lead[l]]); }
                ncz0 = std::max(1, std::min(S.g.ncz_max, (int)floor((hi - lo) / rs)));
            }
            s->build_blocks = std::max(1, s->ncells_xy * ncz0 * BUILD_ITEMS_PER_CELL);
            s->order_blocks = std::max(1, (s->build_blocks + 3) / 4);
            S.cell_cursor = s->cell_cursor.p;
        }
        HTRY(s->flags.alloc(1));
        {
            DevFlags f0;
            memset(&f0, 0, sizeof f0);
            HTRY(hipMemcpy(s->flags.p, &f0, sizeof f0, hipMemcpyHostToDevice));
        }
        S.flags = s->flags.p;
        // host-mapped mirror of a pending list overflow, polled by mdf_compute_device
        HTRY(hipHostMalloc((void **)&s->h_mirror, sizeof(int) * 4, hipHostMallocMapped));
        memset(s->h_mirror, 0, sizeof(int) * 4);
        HTRY(hipHostGetDevicePointer((void **)&S.flag_mirror, s->h_mirror, 0));
        HTRY(hipMemset(s->cell_cnt.p, 0, sizeof(int) * (s->n_cells_cap + 1)));
        // frozen-site cells, built once
        {
            int nw = s->n_wall;
            double wz0 = 0.0, wcsz = 1.0;
            int wncz = 1;
            if (S.g.zper) {
                wncz = S.g.ncz_fixed; wcsz = S.g.csz_fixed;
            } else if (nw > 0) {
                double lo = z[t.n_move], hi = z[t.n_move];
                for (int w = t.n_move; w < t.n_total; w++) { lo = std::min(lo, z[w]); hi = std::max(hi, z[w]); }
                double rs = sqrt(S.rskinsq) * (1.0 + 1e-9);
                wncz = std::max(1, std::min(1024, (int)floor((hi - lo) / rs)));
                wz0 = lo;
                // cells at least the list radius high, so the three-cell stencil of a lead
                // reaches every frozen site within it, also for crystals thinner than that
                wcsz = std::max(hi > lo ? (hi - lo) / wncz : 0.0, rs);
            }
            int nwc = s->ncells_xy * wncz;
            DBuf<int> wcell_of, wcnt, wcur, witems2;
            HTRY(wcell_of.alloc(std::max(nw, 1))); HTRY(wcnt.alloc(nwc + 1)); HTRY(wcur.alloc(nwc + 1));
            HTRY(witems2.alloc(std::max(nw, 1)));
            HTRY(s->wcell_start.alloc(nwc + 1)); HTRY(s->wcell_items.alloc(std::max(nw, 1)));
            HTRY(s->wcell_xyz.alloc(std::max(nw, 1)));
            k_zero_int<<<64, 256, 0, s->stream>>>(wcnt.p, nwc + 1);
            if (nw > 0) {
                k_wall_cells<<<(nw + 255) / 256, 256, 0, s->stream>>>(S, wcell_of.p, wncz, wz0, wcsz);
                k_count_cells<<<(nw + 255) / 256, 256, 0, s->stream>>>(wcell_of.p, nw, wcnt.p);
            }
            k_cell_scan<<<1, 1024, 0, s->stream>>>(wcnt.p, s->wcell_start.p, wcur.p, s->ncells_xy * wncz);
            if (nw > 0) {
                k_cell_fill_static<<<(nw + 255) / 256, 256, 0, s->stream>>>(wcell_of.p, nw, wcur.p, witems2.p);
                k_cell_sort<<<512, 64, 0, s->stream>>>(s->wcell_start.p, witems2.p, s->wcell_items.p, nwc);
                k_cell_xyz<<<(nw + 255) / 256, 256, 0, s->stream>>>(s->wcell_items.p, nw, t.n_move, S.x, S.y, S.z,
                                                                    s->wcell_xyz.p);
            }
            HTRY(hipStreamSynchronize(s->stream));
            wcell_of.release(); wcnt.release(); wcur.release(); witems2.release();
            S.wcell_start = s->wcell_start.p; S.wcell_items = s->wcell_items.p; S.wcell_xyz = s->wcell_xyz.p;
            S.wncz = wncz; S.wz0 = wz0; S.wcsz = wcsz;
        }
        // lists
        HTRY(s->cnt_mm.alloc(std::max(s->n_lead, 1))); HTRY(s->cnt_mw.alloc(std::max(s->n_lead, 1)));
        HTRY(s->cnt_lo.alloc(std::max(s->n_lead, 1)));
        S.cnt_mm = s->cnt_mm.p; S.cnt_mw = s->cnt_mw.p; S.cnt_lo = s->cnt_lo.p;
        {
            int cap_mm = 256, cap_mw = s->n_wall > 0 ? 256 : 1;
            if (par->max_neighbors > 0) { cap_mm = par->max_neighbors; cap_mw = par->max_neighbors; s->fixed_cap = true; }
            TRY(alloc_lists(s, cap_mm, cap_mw));
        }
        // k-set, A_k, static frozen S(k), self energy
        kset(s);
        {
            // rows and chunks must tile the flat k list exactly
            size_t nk = s->ak.size(), covered = 0, ccov = 0;
            for (const KRow &r : s->rows) {
                for (int kn = r.kn0; kn <= r.kn1; kn++) {
                    size_t k = (size_t)r.koff + (kn - r.kn0);
                    if (k >= nk || s->kvec[3 * k] != r.kl || s->kvec[3 * k + 1] != r.km || s->kvec[3 * k + 2] != kn) {
                        rc = fail(MDF_ERR_KSET, "k-set row table inconsistent at row (%d,%d)", r.kl, r.km);
                        goto fail_out;
                    }
                }
                covered += (size_t)(r.kn1 - r.kn0 + 1);
                if (r.kn1 - r.kn0 + 1 > BP_GROUP_K || (r.kn0 > 1) || (r.kn0 < 0 && r.kn1 < 0)) {
                    // the back-projection stages a whole row in LDS and walks kn from 0/1 up and -1 down
                    rc = fail(MDF_ERR_KSET, "k-set row (%d,%d) spans kn %d..%d, unsupported", r.kl, r.km, r.kn0,
                              r.kn1);
                    goto fail_out;
                }
            }
            for (const KChunk &c : s->chunks) ccov += (size_t)c.len;
            if (covered != nk || ccov != nk) {
                rc = fail(MDF_ERR_KSET, "k-set rows cover %zu, chunks %zu of %zu k-vectors", covered, ccov, nk);
                goto fail_out;
            }
        }
        geometry(s);
        {
            int nk = (int)s->ak.size();
            HTRY(upload(s->d_chunks, s->chunks.data(), s->chunks.size()));
            HTRY(upload(s->d_rows, s->rows.data(), s->rows.size()));
            HTRY(upload(s->d_bp_row_start, s->bp_row_start.data(), s->bp_row_start.size()));
            HTRY(upload(s->d_kvec, s->kvec.data(), s->kvec.size()));
            HTRY(upload(s->d_ak, s->ak.data(), s->ak.size()));
            HTRY(s->d_s.alloc(std::max(nk, 1)));
            HTRY(s->d_t.alloc(nk + 4));   // padded: the back-projection prefetches one entry past a row
            if (s->bp_mma) {
                // k_bp_mma's tables: rows' (kl, km), each 16-row tile's K-step range, each k's row,
                // and the zero-padded split T that k_sf_reduce fills inside the kn sphere
                const int nr = (int)s->rows.size(), kn_pad = 16 * s->n_nt;
                std::vector<int2> rkk(std::max(nr, 1)), tks(std::max((nr + 15) / 16, 1));
                std::vector<int> krow(std::max(nk, 1), 0);
                for (int r = 0; r < nr; r++) {
                    const KRow &w = s->rows[r];
                    rkk[r] = make_int2(w.kl, w.km);
                    for (int kn = w.kn0; kn <= w.kn1; kn++) krow[w.koff + kn - w.kn0] = r;
                }
                for (int m = 0; m * 16 < nr; m++) {
                    int lo = 1 << 30, hi = -(1 << 30);
                    for (int r = 16 * m; r < std::min(nr, 16 * m + 16); r++) {
                        lo = std::min(lo, s->rows[r].kn0);
                        hi = std::max(hi, s->rows[r].kn1);
                    }
                    tks[m] = make_int2((lo + s->knmax) >> 4, std::min(s->n_nt - 1, (hi + s->knmax) >> 4));
                }
                HTRY(upload(s->d_rowkk, rkk.data(), rkk.size()));
                HTRY(upload(s->d_tile_ks, tks.data(), tks.size()));
                HTRY(upload(s->d_krow, krow.data(), krow.size()));
                HTRY(s->d_tpack.alloc((size_t)std::max(nr, 1) * 4 * kn_pad));
                HTRY(hipMemset(s->d_tpack.p, 0, sizeof(unsigned) * s->d_tpack.n));
                HTRY(s->d_tscale.alloc(std::max(nr, 1)));
                HTRY(hipMemset(s->d_tscale.p, 0, sizeof(int) * s->d_tscale.n));
            }
            if (s->mma64) TRY(setup_mma64(s));
            HTRY(s->kpart.alloc((size_t)RE_N * std::max(s->n_kblk, (int)s->rows.size())));
            std::vector<double2> zero2(std::max(nk, 1), make_double2(0.0, 0.0));
            HTRY(upload(s->d_swall, zero2.data(), zero2.size()));
            int nw = s->n_wall;
            int per_sb = SF_BLOCK * 4;
            int n_sb_w = std::max(1, (nw + per_sb - 1) / per_sb);
            HTRY(s->sf_partial.alloc((size_t)std::max(std::max(s->n_sb, s->n_sb_mma), n_sb_w) * std::max(nk, 1)));
            HTRY(s->fpart.alloc((size_t)std::max(s->n_rg, s->b64_nrg) * 3 * std::max(t.n_move, 1)));
            s->s_wall.assign(2 * std::max(nk, 1), 0.0);
            s->ewqWW = 0.0;
            if (nw > 0 && nk > 0) {
                const double cux = 2.0 * s->c.pi / S.lx / M_PI, cuy = 2.0 * s->c.pi / S.ly / M_PI,
                             cuz = 2.0 * s->c.pi / S.lz / M_PI;
                int nrow = (int)s->rows.size();
                int ncg = std::max(1, std::min(nrow, (2048 + n_sb_w - 1) / n_sb_w));
                int rpb = (nrow + ncg - 1) / ncg;
                ncg = (nrow + rpb - 1) / rpb;
                k_sf<double, 4><<<dim3(n_sb_w, ncg), SF_BLOCK, 0, s->stream>>>(
                    S.x + t.n_move, S.y + t.n_move, S.z + t.n_move, S.q + t.n_move, nw, cux, cuy, cuz,
                    s->d_rows.p, nrow, rpb, s->sf_partial.p, nk, nullptr);
                k_sf_reduce<<<s->n_kblk, SFR_THREADS, 0, s->stream>>>(s->sf_partial.p, n_sb_w, nk, s->d_swall.p, s->d_ak.p,
                                                              s->d_kvec.p, 1.0, 1.0, 1.0, 0.0, s->d_s.p, s->d_t.p,
                                                              s->kpart.p, nullptr, nullptr, nullptr, 0, 0, nullptr);
                HTRY(hipStreamSynchronize(s->stream));
                HTRY(hipMemcpy(s->s_wall.data(), s->d_s.p, sizeof(double2) * nk, hipMemcpyDeviceToHost));
                HTRY(hipMemcpy(s->d_swall.p, s->d_s.p, sizeof(double2) * nk, hipMemcpyDeviceToDevice));
                for (int k = 0; k < nk; k++)
                    s->ewqWW += s->ak[k] * (s->s_wall[2 * k] * s->s_wall[2 * k] + s->s_wall[2 * k + 1] * s->s_wall[2 * k + 1]);
            }
            s->ewqWW *= 0.5;
            // self energy (src/forcem.f90:794-822)
            double q2 = 0.0;
            for (int i = 0; i < t.n_total; i++) q2 += t.charge[i] * t.charge[i];
            s->uself = -0.5 * S.spii * S.facq * S.alfa * q2;
            auto erfc_as = [&](double ar) {
                double ex = exp(-ar * ar);
                if (par->erfc_mode == MDF_ERFC_AS7126) {
                    double tt = 1.0 / (1.0 + s->c.p * ar);
                    return ((((s->c.a5 * tt + s->c.a4) * tt + s->c.a3) * tt + s->c.a2) * tt + s->c.a1) * tt * ex;
                }
                return erfc(ar);
            };
            if (t.n_rig > 0)
                for (int a = 0; a < 3; a++)
                    for (int b = a + 1; b < 3; b++) {
                        double rij = sqrt((x[a] - x[b]) * (x[a] - x[b]) + (y[a] - y[b]) * (y[a] - y[b]) + (z[a] - z[b]) * (z[a] - z[b]));
                        s->uself = s->uself - (double)t.n_rig * S.facq * t.charge[a] * t.charge[b] * (1.0 - erfc_as(S.alfa * rij)) / rij;
                    }
            for (int k = 0; k < t.n_constraints; k++) {
                double rij = t.con_len[k];
                s->uself = s->uself - S.facq * t.charge[t.con_a[k]] * t.charge[t.con_b[k]] * (1.0 - erfc_as(S.alfa * rij)) / rij;
            }
            s->crf0 = -2.0 * s->c.pi * S.lzi * S.lxi * S.lyi;
        }
        // tables and reduction buffers
        wall_tables(s, t.charge);
        HTRY(upload(s->d_wl, &s->wl, 1));
        HTRY(s->dip_part.alloc(4 * (size_t)s->skin_blocks)); HTRY(s->dipz.alloc(1));
        HTRY(s->eppart.alloc((size_t)6 * s->n_epblk));
        HTRY(s->colpart.alloc((size_t)NCOL * s->n_colblk * MAX_LEAD_TYPES));
        HTRY(s->fcom.alloc(3 * nr)); HTRY(s->torque.alloc(3 * nr));
        HTRY(s->e_vdw.alloc(nijt + 1)); HTRY(s->e_coul.alloc(nijt + 1));
        HTRY(s->d_scal.alloc(1));
        // first list (nnlist_init -> nnlist_make at force_init)
        for (int attempt = 0; attempt < 4; attempt++) {
            launch_lists(s, 1);
            k_pair_totals<<<1, 256, 0, s->stream>>>(S);
            HTRY(hipGetLastError());
            int redo = 0;
            HTRY(hipStreamSynchronize(s->stream));
            HTRY(hipMemcpy(&s->last_flags, s->flags.p, sizeof(DevFlags), hipMemcpyDeviceToHost));
            if (s->last_flags.overflow_mm || s->last_flags.overflow_mw) {
                TRY(clear_overflow(s));
                if (s->fixed_cap) {
                    rc = fail(MDF_ERR_LIST_OVERFLOW, "nnlist_make(): nnlist too big! need %d / %d partners per lead",
                              s->last_flags.overflow_mm, s->last_flags.overflow_mw);
                    goto fail_out;
                }
                TRY(alloc_lists(s, std::max(s->S.cap_mm, (int)(s->last_flags.overflow_mm * 1.25) + 16),
                                std::max(s->S.cap_mw, (int)(s->last_flags.overflow_mw * 1.25) + 16)));
                redo = 1;
            }
            if (!redo) break;
        }
    }
    *out = s;
    return MDF_OK;
fail_out:
    free_all(s);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    if (s->stream2) (void)hipStreamDestroy(s->stream2);
    destroy_events(s);
    delete s;
    return rc;
#undef TRY
#undef HTRY
}

static int set_inputs(mdf_system *s, const double *x, const double *y, const double *z, const double *cx,
                      const double *cy, const double *cz, hipMemcpyKind kind) {
    const int nm = s->top.n_move, nr = s->top.n_rig;
    hipStream_t st = s->stream;
    HIPCHK(hipMemcpyAsync(s->x.p, x, sizeof(double) * nm, kind, st));
    HIPCHK(hipMemcpyAsync(s->y.p, y, sizeof(double) * nm, kind, st));
    HIPCHK(hipMemcpyAsync(s->z.p, z, sizeof(double) * nm, kind, st));
    if (nr > 0) {
        HIPCHK(hipMemcpyAsync(s->cx.p, cx, sizeof(double) * nr, kind, st));
        HIPCHK(hipMemcpyAsync(s->cy.p, cy, sizeof(double) * nr, kind, st));
        HIPCHK(hipMemcpyAsync(s->cz.p, cz, sizeof(double) * nr, kind, st));
    }
    return MDF_OK;
}

int mdf_compute(mdf_system *s, const double *x, const double *y, const double *z, const double *cx, const double *cy,
                const double *cz, mdf_outputs *out) {
    if (!s || !x || !y || !z || !out) return fail(MDF_ERR_ARG, "null argument");
    if (s->top.n_rig > 0 && (!cx || !cy || !cz)) return fail(MDF_ERR_ARG, "COM arrays required");
    HIPCHK(hipSetDevice(s->device));
    bool late = false;   // a step enqueued earlier by mdf_compute_device overflowed
    for (int attempt = 0; attempt < 4; attempt++) {
        int rc = set_inputs(s, x, y, z, cx, cy, cz, hipMemcpyHostToDevice);
        if (rc) return rc;
        rc = run_step(s, attempt > 0);
        if (rc) return rc;
        int redo = 0;
        rc = finish_step(s, &redo);
        if (rc) return rc;
        if (redo == 2) late = true;
        if (!redo) {
            rc = copy_outputs(s, out);
            return (rc == MDF_OK && late) ? late_overflow_error(s) : rc;
        }
    }
    return fail(MDF_ERR_LIST_OVERFLOW, "pair list kept overflowing");
}

int mdf_compute_device(mdf_system *s, const double *d_x, const double *d_y, const double *d_z, const double *d_cx,
                       const double *d_cy, const double *d_cz, void *stream) {
    if (!s || !d_x || !d_y || !d_z) return fail(MDF_ERR_ARG, "null argument");
    if (s->top.n_rig > 0 && (!d_cx || !d_cy || !d_cz)) return fail(MDF_ERR_ARG, "COM arrays required");
    HIPCHK(hipSetDevice(s->device));
    if (s->h_mirror && *(volatile int *)s->h_mirror)
        return fail(MDF_ERR_LIST_OVERFLOW, "an earlier step's pair list overflowed its rows (capacity %d / %d); the "
                    "step was not enqueued: call mdf_read_outputs", s->S.cap_mm, s->S.cap_mw);
    // start after the producer of the inputs (the handle's stream does not block on others)
    HIPCHK(hipEventRecord(s->in_event, (hipStream_t)stream));
    HIPCHK(hipStreamWaitEvent(s->stream, s->in_event, 0));
    int n = std::max(s->top.n_move, s->top.n_rig);
    k_gather<<<std::max(1, (n + 255) / 256), 256, 0, s->stream>>>(s->S, d_x, d_y, d_z, d_cx, d_cy, d_cz);
    return run_step(s, 0);
}

int mdf_read_outputs(mdf_system *s, mdf_outputs *out) {
    if (!s || !out) return fail(MDF_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(s->device));
    int redo = 0;
    int rc = finish_step(s, &redo);
    if (rc) return rc;
    const bool late = redo == 2;
    if (redo) {
        // the inputs are still on the device: run the step again with the grown lists
        rc = run_step(s, 1);
        if (rc) return rc;
        rc = finish_step(s, &redo);
        if (rc) return rc;
        if (redo) return fail(MDF_ERR_LIST_OVERFLOW, "pair list kept overflowing");
    }
    rc = copy_outputs(s, out);
    return (rc == MDF_OK && late) ? late_overflow_error(s) : rc;
}

int mdf_device_results(mdf_system *s, const double **d_f_site, const double **d_f_com, const double **d_torque) {
    if (!s) return fail(MDF_ERR_ARG, "null handle");
    if (d_f_site) *d_f_site = s->fx.p;  // fx, fy, fz are separate allocations: see f_site layout note
    if (d_f_com) *d_f_com = s->fcom.p;
    if (d_torque) *d_torque = s->torque.p;
    return MDF_OK;
}

int mdf_zero(mdf_system *s) { return s ? MDF_OK : fail(MDF_ERR_ARG, "null handle"); }

int mdf_get_pairs(mdf_system *s, int32_t *i_mm, int32_t *j_mm, int32_t cap_mm, int32_t *n_mm, int32_t *i_mw,
                  int32_t *j_mw, int32_t cap_mw, int32_t *n_mw) {
    if (!s || !n_mm || !n_mw) return fail(MDF_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(s->device));
    HIPCHK(hipStreamSynchronize(s->stream));
    const int L = s->n_lead, cm = s->S.cap_mm, cw = s->S.cap_mw;
    std::vector<int> cnt(L), cntw(L), nb((size_t)L * cm), nw((size_t)L * cw);
    HIPCHK(hipMemcpy(cnt.data(), s->cnt_mm.p, sizeof(int) * L, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(cntw.data(), s->cnt_mw.p, sizeof(int) * L, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(nb.data(), s->nbr_mm.p, sizeof(int) * nb.size(), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(nw.data(), s->nbr_mw.p, sizeof(int) * nw.size(), hipMemcpyDeviceToHost));
    std::vector<std::pair<int, int>> pm, pw;
    for (int l = 0; l < L; l++) {
        int i = s->h_lead[l];
        for (int k = 0; k < std::min(cnt[l], cm); k++) {
            int j = nb[(size_t)l * cm + k];
            if (j > i) pm.push_back({i, j});
        }
        for (int k = 0; k < std::min(cntw[l], cw); k++) pw.push_back({i, nw[(size_t)l * cw + k]});
    }
    std::sort(pm.begin(), pm.end());
    std::sort(pw.begin(), pw.end());
    *n_mm = (int32_t)pm.size();
    *n_mw = (int32_t)pw.size();
    if ((int)pm.size() > cap_mm || (int)pw.size() > cap_mw) return MDF_ERR_ARG;
    for (size_t k = 0; k < pm.size(); k++) { i_mm[k] = pm[k].first; j_mm[k] = pm[k].second; }
    for (size_t k = 0; k < pw.size(); k++) { i_mw[k] = pw[k].first; j_mw[k] = pw[k].second; }
    return MDF_OK;
}

int mdf_get_kvectors(mdf_system *s, int32_t *kvec, double *ak, int32_t cap, int32_t *n) {
    if (!s || !n) return fail(MDF_ERR_ARG, "null argument");
    *n = (int32_t)s->ak.size();
    if (*n > cap) return MDF_ERR_ARG;
    if (kvec) memcpy(kvec, s->kvec.data(), sizeof(int32_t) * s->kvec.size());
    if (ak) memcpy(ak, s->ak.data(), sizeof(double) * s->ak.size());
    return MDF_OK;
}

int mdf_get_structure_factor(mdf_system *s, double *sf, int32_t cap) {
    if (!s || !sf) return fail(MDF_ERR_ARG, "null argument");
    int nk = (int)s->ak.size();
    if (nk > cap) return MDF_ERR_ARG;
    HIPCHK(hipSetDevice(s->device));
    HIPCHK(hipStreamSynchronize(s->stream));
    if (nk) HIPCHK(hipMemcpy(sf, s->d_s.p, sizeof(double2) * nk, hipMemcpyDeviceToHost));
    return MDF_OK;
}

int mdf_stage_times(mdf_system *s, const char **names, float *ms, int32_t cap, int32_t *n) {
    if (!s || !n) return fail(MDF_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(s->device));
    HIPCHK(hipStreamSynchronize(s->stream));
    drain_all(s);
    *n = N_STAGES + 1 + N_KT;
    for (int k = 0; k < N_STAGES && k < cap; k++) {
        if (names) names[k] = STAGE_NAMES[k];
        if (ms) ms[k] = s->stage_steps > 0 ? (float)(s->stage_sum[k] / (double)s->stage_steps) : 0.f;
    }
    if (N_STAGES < cap) {
        if (names) names[N_STAGES] = STAGE_NAMES[N_STAGES];
        if (ms) ms[N_STAGES] = s->step_steps > 0 ? (float)(s->step_sum / (double)s->step_steps) : 0.f;
    }
    for (int k = 0; k < N_KT && N_STAGES + 1 + k < cap; k++) {
        if (names) names[N_STAGES + 1 + k] = KT_NAMES[k];
        if (ms) ms[N_STAGES + 1 + k] = s->kt_n[k] > 0 ? (float)(s->kt_sum[k] / (double)s->kt_n[k]) : 0.f;
    }
    return MDF_OK;
}

int mdf_autotune(mdf_system *s, int32_t reps, int32_t *chosen) {
    if (!s) return fail(MDF_ERR_ARG, "null handle");
    reps = std::max(reps, 1);
    HIPCHK(hipSetDevice(s->device));
    HIPCHK(hipStreamSynchronize(s->stream));
    drain_all(s);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    int rc = MDF_OK;
    // mean time of one launch of the given stage, after one untimed launch
    auto timed = [&](auto launch, float *ms) -> int {
        launch();
        HIPCHK(hipEventRecord(e0, s->stream));
        for (int r = 0; r < reps; r++) launch();
        HIPCHK(hipEventRecord(e1, s->stream));
        HIPCHK(hipEventSynchronize(e1));
        float t = 0.0f;
        HIPCHK(hipEventElapsedTime(&t, e0, e1));
        HIPCHK(hipGetLastError());
        *ms = t / reps;
        return MDF_OK;
    };
    float best = 1e30f, t = 0.0f;
    int best_g = s->pair_g;
    bool best_strided = s->pair_strided;
    for (int g : {8, 16, 32})
        for (int ro = 0; ro < 2; ro++) {
            s->pair_g = g;
            s->pair_strided = ro != 0;
            if ((rc = timed([&] { launch_pair(s); }, &t)) != MDF_OK) goto done;
            if (t < best) { best = t; best_g = g; best_strided = ro != 0; }
        }
    s->pair_g = best_g;
    s->pair_strided = best_strided;
    {
        best = 1e30f;
        int best_sf = s->sf_target_blocks;
        for (int tb : {1024, 2048, 4096, 8192}) {
            s->sf_target_blocks = tb;
            geometry(s);
            if ((rc = timed([&] { launch_sf(s); }, &t)) != MDF_OK) goto done;
            if (t < best) { best = t; best_sf = tb; }
        }
        s->sf_target_blocks = best_sf;
        best = 1e30f;
        int best_bp = s->bp_target_blocks;
        for (int tb : {512, 1024, 2048}) {
            s->bp_target_blocks = tb;
            if ((rc = apply_geometry(s)) != MDF_OK) goto done;
            if ((rc = timed([&] { launch_bp(s); }, &t)) != MDF_OK) goto done;
            if (t < best) { best = t; best_bp = tb; }
        }
        s->bp_target_blocks = best_bp;
    }
    rc = apply_geometry(s);
    if (chosen) {
        chosen[0] = s->pair_g;
        chosen[1] = s->pair_strided ? 1 : 0;
        chosen[2] = s->sf_target_blocks;
        chosen[3] = s->bp_target_blocks;
    }
done:
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

int mdf_set_launch_shapes(mdf_system *s, int32_t pair_lanes, int32_t pair_strided, int32_t sf_blocks,
                          int32_t bp_blocks) {
    if (!s) return fail(MDF_ERR_ARG, "null handle");
    if (pair_lanes != 8 && pair_lanes != 16 && pair_lanes != 32)
        return fail(MDF_ERR_ARG, "pair lanes per lead must be 8, 16 or 32 (got %d)", pair_lanes);
    if (sf_blocks <= 0 || bp_blocks <= 0) return fail(MDF_ERR_ARG, "block targets must be positive");
    HIPCHK(hipSetDevice(s->device));
    HIPCHK(hipStreamSynchronize(s->stream));
    drain_all(s);
    s->pair_g = pair_lanes;
    s->pair_strided = pair_strided != 0;
    s->sf_target_blocks = sf_blocks;
    s->bp_target_blocks = bp_blocks;
    return apply_geometry(s);
}

int mdf_set_stage_timing(mdf_system *s, int32_t on) {
    if (!s) return fail(MDF_ERR_ARG, "null handle");
    HIPCHK(hipSetDevice(s->device));
    HIPCHK(hipStreamSynchronize(s->stream));
    if (on < 0 || on > 2) return fail(MDF_ERR_ARG, "stage timing mode %d (0 off, 1 on, 2 on with a list build every step)", on);
    drain_all(s);
    s->stage_timing = on;
    return MDF_OK;
}

int mdf_reset_stage_times(mdf_system *s) {
    if (!s) return fail(MDF_ERR_ARG, "null handle");
    HIPCHK(hipSetDevice(s->device));
    HIPCHK(hipStreamSynchronize(s->stream));
    drain_all(s);
    for (int k = 0; k < N_STAGES; k++) s->stage_sum[k] = 0.0;
    for (int k = 0; k < N_KT; k++) { s->kt_sum[k] = 0.0; s->kt_n[k] = 0; }
    s->stage_steps = 0;
    s->step_sum = 0.0;
    s->step_steps = 0;
    return MDF_OK;
}

void mdf_destroy(mdf_system *s) {
    if (!s) return;
    (void)hipSetDevice(s->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    drop_graphs(s);
    free_all(s);
    destroy_events(s);
    if (s->in_event) (void)hipEventDestroy(s->in_event);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    if (s->stream2) (void)hipStreamDestroy(s->stream2);
    delete s;
}

}  // extern "C"

// ---- reciprocal-force surrogate and multiple time steps: host side ----------------------
static unsigned short host_f2bf(float f) {   // round to nearest even, as the kernels
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

// bfloat16 copies (and the transposes the backward pass uses) of the weights
static void host_weight_copies(const float *p, unsigned short *wb) {
    for (int k = 0; k < SG_H * SG_D; k++) wb[SG_WB_W1 + k] = host_f2bf(p[SG_OFF_W1 + k]);
    for (int u2 = 0; u2 < SG_H; u2++)
        for (int u1 = 0; u1 < SG_H; u1++) {
            const unsigned short v = host_f2bf(p[SG_OFF_W2 + u2 * SG_H + u1]);
            wb[SG_WB_W2 + u2 * SG_H + u1] = v;
            wb[SG_WB_W2T + u1 * SG_H + u2] = v;
        }
    for (int m = 0; m < SG_M; m++)
        for (int u = 0; u < SG_H; u++) {
            const unsigned short v = host_f2bf(p[SG_OFF_W3 + m * SG_H + u]);
            wb[SG_WB_W3 + m * SG_H + u] = v;
            wb[SG_WB_W3T + u * SG_M + m] = v;
        }
}

// initial weights: W1, W2 normal with variance 1/fan-in (splitmix64 + Box-Muller), W3 and
// the biases zero; the normalisation kernel then sets c_0 = su_0 (the physics prior)
static void host_init_params(std::vector<float> &p, uint32_t seed) {
    p.assign(SG_NP, 0.0f);
    uint64_t st = 0x9E3779B97F4A7C15ull ^ seed;
    auto next = [&]() {
        uint64_t z = (st += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return ((z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    };
    auto normal = [&]() { return sqrt(-2.0 * log(next())) * cos(2.0 * M_PI * next()); };
    for (int k = 0; k < SG_H * SG_D; k++) p[SG_OFF_W1 + k] = (float)(normal() / sqrt((double)SG_D));
    for (int k = 0; k < SG_H * SG_H; k++) p[SG_OFF_W2 + k] = (float)(normal() / sqrt((double)SG_H));
}

static void mts_geometry(int n, int &n_sblk, int &n_iblk, int &n_tblk) {
    const int ntiles = (n + 15) / 16;
    n_sblk = std::max(1, (n + 255) / 256);
    n_iblk = std::min(256, std::max(1, (ntiles + 3) / 4));
    n_tblk = std::min(SG_TRAIN_BLOCKS, std::max(1, (ntiles + 3) / 4));
}

// an Adam step of the scheme trains on about SG_TRAIN_TILES tiles of 16 sites
constexpr int SG_TRAIN_TILES = 1024;
static int mts_train_stride(int n) { return std::max(1, ((n + 15) / 16 + SG_TRAIN_TILES - 1) / SG_TRAIN_TILES); }

static int mts_alloc(mdf_system *s) {
    const size_t nm = (size_t)std::max(s->top.n_move, 1);
    mts_geometry(s->top.n_move, s->sg_n_sblk, s->sg_n_iblk, s->sg_n_tblk);
    s->sg_train_stride = mts_train_stride(s->top.n_move);
    HIPCHK(s->mts_st.alloc(1));
    HIPCHK(s->sg_feat.alloc(nm * SG_D));
    HIPCHK(s->sg_vec.alloc(nm * SG_M * 3));
    HIPCHK(s->sg_feat_prev.alloc(nm * SG_D));
    HIPCHK(s->sg_pos4.alloc((size_t)std::max(s->top.n_total, 1)));
    {
        std::vector<int> nsi(s->top.n_types + 1), msize(std::max(s->top.n_total, 1), 0);
        HIPCHK(hipMemcpy(nsi.data(), s->nsi.p, sizeof(int) * nsi.size(), hipMemcpyDeviceToHost));
        for (int l = 0; l < s->n_lead; l++) msize[s->h_lead[l]] = nsi[s->h_itype[s->h_lead[l]]];
        HIPCHK(upload(s->sg_msize, msize.data(), msize.size()));
    }
    HIPCHK(s->sg_vec_prev.alloc(nm * SG_M * 3));
    HIPCHK(s->sg_out.alloc(3 * nm));
    HIPCHK(s->sg_out_prev.alloc(3 * nm));
    HIPCHK(s->sg_params.alloc(SG_NP));
    HIPCHK(s->sg_m.alloc(SG_NP));
    HIPCHK(s->sg_v.alloc(SG_NP));
    HIPCHK(s->sg_gsum.alloc(SG_GP));
    HIPCHK(s->sg_norm.alloc(SG_NORM_N));
    HIPCHK(s->sg_gpart.alloc((size_t)SG_TRAIN_BLOCKS * SG_GP));
    HIPCHK(s->sg_wb.alloc(SG_WB_N));
    DBuf<double> *db[] = {&s->sg_delta, &s->sg_frec, &s->sg_fmts, &s->sg_rprev, &s->sg_fprev, &s->sg_frec_prev,
                          &s->sg_delta_prev};
    for (auto *b : db) {
        HIPCHK(b->alloc(3 * nm));
        HIPCHK(hipMemset(b->p, 0, sizeof(double) * 3 * nm));
    }
    HIPCHK(hipMemset(s->sg_out.p, 0, sizeof(float) * 3 * nm));
    HIPCHK(s->sg_jpart.alloc(4 * (size_t)s->sg_n_sblk));
    HIPCHK(s->sg_epart.alloc((size_t)s->sg_n_sblk));
    return MDF_OK;
}

// parameters, moments, copies and normalisation to the device; the state restarts collecting
static int mts_load(mdf_system *s, const float *params, const float *norm, bool norm_ready) {
    std::vector<unsigned short> wb(SG_WB_N);
    host_weight_copies(params, wb.data());
    HIPCHK(hipMemcpy(s->sg_params.p, params, sizeof(float) * SG_NP, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(s->sg_wb.p, wb.data(), sizeof(unsigned short) * SG_WB_N, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(s->sg_norm.p, norm, sizeof(float) * SG_NORM_N, hipMemcpyHostToDevice));
    HIPCHK(hipMemset(s->sg_m.p, 0, sizeof(float) * SG_NP));
    HIPCHK(hipMemset(s->sg_v.p, 0, sizeof(float) * SG_NP));
    MtsState st;
    memset(&st, 0, sizeof st);
    st.exact = 1;
    st.norm_ready = norm_ready ? 1 : 0;
    st.prior = norm_ready ? 0 : 1;
    HIPCHK(hipMemcpy(s->mts_st.p, &st, sizeof st, hipMemcpyHostToDevice));
    return MDF_OK;
}

extern "C" {

int mdf_mts_enable(mdf_system *s, const mdf_mts_params *p) {
    if (!s) return fail(MDF_ERR_ARG, "null handle");
    HIPCHK(hipSetDevice(s->device));
    HIPCHK(hipStreamSynchronize(s->stream));
    drain_all(s);
    drop_graphs(s);
    if (!p) {
        mts_free(s);
        return MDF_OK;
    }
    if (s->chem_on) return fail(MDF_ERR_ARG, "Widom insertions read every step's S(k): switch them off first");
    if (p->interval < 1) return fail(MDF_ERR_ARG, "interval %d < 1", p->interval);
    if (p->model != MDF_MTS_HELD && p->model != MDF_MTS_NETWORK) return fail(MDF_ERR_ARG, "unknown model %d", p->model);
    if (p->train_frames < 0 || p->train_iters < 0 || p->train_iters > 16)
        return fail(MDF_ERR_ARG, "train_frames %d / train_iters %d out of range", p->train_frames, p->train_iters);
    if (!(p->tol > 0.0) || !(p->lr >= 0.0f)) return fail(MDF_ERR_ARG, "tol must be > 0 and lr >= 0");
    const int nm = s->top.n_move;
    if (nm <= 0 || s->ak.empty()) return fail(MDF_ERR_ARG, "multiple time steps need moving sites and a k-set");
    if (p->model == MDF_MTS_NETWORK) {
        // every moving site in a lead's molecule, and the lists reach every site within 6 A
        std::vector<int> nsi(s->top.n_types + 1);
        HIPCHK(hipMemcpy(nsi.data(), s->nsi.p, sizeof(int) * nsi.size(), hipMemcpyDeviceToHost));
        std::vector<double> hx(nm), hy(nm), hz(nm);
        HIPCHK(hipMemcpy(hx.data(), s->S.x, sizeof(double) * nm, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(hy.data(), s->S.y, sizeof(double) * nm, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(hz.data(), s->S.z, sizeof(double) * nm, hipMemcpyDeviceToHost));
        std::vector<char> cover(nm, 0);
        double off = 0.0;
        for (int l = 0; l < s->n_lead; l++) {
            const int i = s->h_lead[l];
            const int ns = nsi[s->h_itype[i]];
            if (i + ns > nm) return fail(MDF_ERR_ARG, "surrogate: the molecule of lead site %d is not all moving", i);
            for (int a = 0; a < ns; a++) {
                cover[i + a] = 1;
                const double dx = hx[i + a] - hx[i], dy = hy[i + a] - hy[i], dz = hz[i + a] - hz[i];
                off = std::max(off, sqrt(dx * dx + dy * dy + dz * dz));
            }
        }
        for (int i = 0; i < nm; i++)
            if (!cover[i]) return fail(MDF_ERR_ARG, "surrogate: moving site %d is in no molecule with a lead site", i);
        if (SG_RC + 2.0 * off >= sqrt(s->S.rcutsq))
            return fail(MDF_ERR_ARG, "surrogate: 6 A + 2 x %.3f A (site-to-lead) must be below rc %.3f", off,
                        sqrt(s->S.rcutsq));
        const double lmin = std::min(s->par.box[0], s->par.box[1]);
        if (2.0 * SG_RC >= lmin || (s->S.zper_force && 2.0 * SG_RC >= s->par.box[2]))
            return fail(MDF_ERR_ARG, "surrogate: box shorter than 12 A");
    }
    mts_free(s);
    int rc = mts_alloc(s);
    if (rc) { mts_free(s); return rc; }
    std::vector<float> params;
    host_init_params(params, p->seed);
    std::vector<float> norm(SG_NORM_N, 1.0f);
    for (int k = 0; k < SG_D; k++) norm[k] = 0.0f;
    rc = mts_load(s, params.data(), norm.data(), false);
    if (rc) { mts_free(s); return rc; }
    s->mts_interval = p->interval;
    s->mts_model = p->model;
    s->mts_train_frames = p->train_frames;
    s->mts_train_iters = p->model == MDF_MTS_NETWORK ? p->train_iters : 0;
    s->mts_tol = p->tol;
    s->mts_lr = p->lr;
    s->mts_on = true;
    return MDF_OK;
}

int mdf_reciprocal_kernels(const mdf_system *s, int32_t *sf_matrix, int32_t *bp_matrix) {
    if (!s) return fail(MDF_ERR_ARG, "null handle");
    if (sf_matrix) *sf_matrix = (s->sf_mma || s->mma64) ? 1 : 0;
    if (bp_matrix) *bp_matrix = (s->bp_mma || s->mma64) ? 1 : 0;
    return MDF_OK;
}

int mdf_observables_enable(mdf_system *s, double rdel) {
    if (!s) return fail(MDF_ERR_ARG, "null handle");
    HIPCHK(hipSetDevice(s->device));
    HIPCHK(hipStreamSynchronize(s->stream));
    drain_all(s);
    drop_graphs(s);
    s->obs_ngc.release(); s->obs_efi.release(); s->obs_efipos.release();
    s->obs_on = false;
    if (!(rdel > 0.0)) return MDF_OK;
    // bins 0..lmax, lmax = int(lw rdeli + 0.5) + 1 (src/measure.f90:72,80)
    const double rdeli = 1.0 / rdel;
    const int nbins = (int)(s->top.wall_lw * rdeli + 0.5) + 2;
    if (nbins < 1 || nbins > (1 << 24)) return fail(MDF_ERR_ARG, "rdel %g gives %d bins", rdel, nbins);
    HIPCHK(s->obs_ngc.alloc((size_t)nbins * std::max(s->nijt, 1)));
    HIPCHK(s->obs_efipos.alloc(2 * (size_t)nbins));
    HIPCHK(s->obs_efi.alloc((size_t)std::max(s->top.n_move, 1)));
    HIPCHK(hipMemset(s->obs_ngc.p, 0, sizeof(unsigned long long) * s->obs_ngc.n));
    HIPCHK(hipMemset(s->obs_efipos.p, 0, sizeof(double) * s->obs_efipos.n));
    HIPCHK(hipMemset(s->obs_efi.p, 0, sizeof(double) * s->obs_efi.n));
    memset(&s->obs, 0, sizeof s->obs);
    s->obs.ngc = s->obs_ngc.p;
    s->obs.efi = s->obs_efi.p;
    s->obs.efipos = s->obs_efipos.p;
    s->obs.rdeli = rdeli;
    s->obs.lwh = 0.5 * s->top.wall_lw;
    s->obs.nbins = nbins;
    s->obs.count_mm = s->S.zper_force ? 1 : 0;
    s->obs_on = true;
    return MDF_OK;
}

int mdf_get_observables(mdf_system *s, int64_t *ngc, double *efipos, int32_t *nbins, int32_t reset) {
    if (!s) return fail(MDF_ERR_ARG, "null handle");
    if (!s->obs_on) return fail(MDF_ERR_ARG, "observables are off");
    HIPCHK(hipSetDevice(s->device));
    HIPCHK(hipStreamSynchronize(s->stream));
    if (nbins) *nbins = s->obs.nbins;
    if (ngc) HIPCHK(hipMemcpy(ngc, s->obs_ngc.p, sizeof(int64_t) * s->obs_ngc.n, hipMemcpyDeviceToHost));
    if (efipos) HIPCHK(hipMemcpy(efipos, s->obs_efipos.p, sizeof(double) * s->obs_efipos.n, hipMemcpyDeviceToHost));
    if (reset) {
        HIPCHK(hipMemset(s->obs_ngc.p, 0, sizeof(unsigned long long) * s->obs_ngc.n));
        HIPCHK(hipMemset(s->obs_efipos.p, 0, sizeof(double) * s->obs_efipos.n));
    }
    return MDF_OK;
}

int mdf_angle_histogram_enable(mdf_system *s, const int32_t *ang_type, int32_t n_atypes) {
    if (!s) return fail(MDF_ERR_ARG, "null handle");
    HIPCHK(hipSetDevice(s->device));
    HIPCHK(hipStreamSynchronize(s->stream));
    drain_all(s);
    drop_graphs(s);
    s->ang_type.release(); s->ang_hist.release();
    s->ang_on = false;
    s->n_atypes = 0;
    if (!ang_type) return MDF_OK;
    const int na = s->top.n_angles;
    if (na <= 0) return fail(MDF_ERR_ARG, "the system has no angle terms");
    if (n_atypes < 1 || n_atypes > 4096) return fail(MDF_ERR_ARG, "n_atypes %d", n_atypes);
    for (int a = 0; a < na; a++)
        if (ang_type[a] < 1 || ang_type[a] > n_atypes)
            return fail(MDF_ERR_ARG, "angle %d has type %d outside 1..%d", a, ang_type[a], n_atypes);
    HIPCHK(s->ang_type.alloc(na));
    HIPCHK(hipMemcpy(s->ang_type.p, ang_type, sizeof(int) * na, hipMemcpyHostToDevice));
    HIPCHK(s->ang_hist.alloc((size_t)MDF_NGANG_BINS * n_atypes));
    HIPCHK(hipMemset(s->ang_hist.p, 0, sizeof(unsigned long long) * s->ang_hist.n));
    s->n_atypes = n_atypes;
    s->ang_on = true;
    return MDF_OK;
}

int mdf_get_angle_histogram(mdf_system *s, int64_t *ngang, int32_t reset) {
    if (!s) return fail(MDF_ERR_ARG, "null handle");
    if (!s->ang_on) return fail(MDF_ERR_ARG, "the angle histogram is off");
    HIPCHK(hipSetDevice(s->device));
    HIPCHK(hipStreamSynchronize(s->stream));
    if (ngang) HIPCHK(hipMemcpy(ngang, s->ang_hist.p, sizeof(int64_t) * s->ang_hist.n, hipMemcpyDeviceToHost));
    if (reset) HIPCHK(hipMemset(s->ang_hist.p, 0, sizeof(unsigned long long) * s->ang_hist.n));
    return MDF_OK;
}

int mdf_chempot_enable(mdf_system *s, const mdf_chempot_params *p) {
    if (!s) return fail(MDF_ERR_ARG, "null handle");
    HIPCHK(hipSetDevice(s->device));
    HIPCHK(hipStreamSynchronize(s->stream));
    chem_free(s);
    if (!p) return MDF_OK;
    if (s->mts_on) return fail(MDF_ERR_ARG, "insertions read every step's S(k): switch the k-space surrogate off first");
    if (p->n_insert < 1 || p->n_insert > (1 << 22)) return fail(MDF_ERR_ARG, "n_insert %d outside 1..2^22", p->n_insert);
    if (p->n_bins < 1 || p->n_bins > (1 << 24)) return fail(MDF_ERR_ARG, "n_bins %d outside 1..2^24", p->n_bins);
    if (p->ew3dc != 0 && p->ew3dc != 1) return fail(MDF_ERR_ARG, "ew3dc must be 0 or 1");
    if (p->max_partners < 0) return fail(MDF_ERR_ARG, "max_partners < 0");
    // wall distances to center distances (src/measure.f90:151-154)
    const double lwh = 0.5 * s->top.wall_lw;
    const double lwmin = lwh - p->wall_max, lwmax = lwh - p->wall_min;
    if (!(lwmax > lwmin) || !(lwmax > 0.0))
        return fail(MDF_ERR_ARG, "insertion window: wall_min %g must be below wall_max %g and below lw/2 = %g",
                    p->wall_min, p->wall_max, lwh);
    int klmax = 0, kmmax = 0, knmax = 0;
    for (size_t k = 0; k + 2 < s->kvec.size(); k += 3) {
        klmax = std::max(klmax, std::abs(s->kvec[k]));
        kmmax = std::max(kmmax, std::abs(s->kvec[k + 1]));
        knmax = std::max(knmax, std::abs(s->kvec[k + 2]));
    }
    const int per = (klmax + 1) + (2 * kmmax + 1) + (2 * knmax + 1);
    if (per > CHEM_TAB_MAX) return fail(MDF_ERR_ARG, "k-set too wide for the insertion tables (%d > %d)", per, CHEM_TAB_MAX);
    int n_nonh = 0;
    for (int t : s->h_itype) n_nonh += t != 2;
    const int cap = std::max(1, p->max_partners > 0 ? p->max_partners : std::min(n_nonh, 4096));
    const int n = p->n_insert, nrmax = CHEM_ROT_MAX * n;
    const bool ref = s->par.const_mode == MDF_CONST_REFERENCE;
    auto lit = [&](double v) { return ref ? (double)(float)v : v; };
    s->chem_p = *p;
    s->chem_seed = p->seed > 0.0 ? p->seed : 14135621769.0 + 3.0;   // rnd_init on rank 0 (src/measure.f90:150)
    s->chem_lwmin = lwmin;
    s->chem_lwmax = lwmax;
    s->chem_facbin = 1.0 / (lwmax - lwmin) * (double)(float)p->n_bins;
    s->chem_aco = lit(-0.333333);
    s->chem_asi = lit(0.942809);
    s->chem_lo = lit(0.001);
    s->chem_lds = sizeof(double2) * 3 * per;
#define CTRY(x)                          \
    do {                                 \
        hipError_t e_ = (x);             \
        if (e_ != hipSuccess) {          \
            chem_free(s);                \
            return fail(MDF_ERR_HIP, "chempot buffers: %s", hipGetErrorString(e_)); \
        }                                \
    } while (0)
    CTRY(s->ch_ins.alloc(3 * (size_t)n));
    CTRY(s->ch_il.alloc(n));
    CTRY(s->ch_insrot.alloc((size_t)n + 1));
    CTRY(s->ch_rotins.alloc(nrmax));
    CTRY(s->ch_rot.alloc(6 * (size_t)nrmax));
    CTRY(s->ch_over.alloc(n));
    CTRY(s->ch_nnbr.alloc(n));
    CTRY(s->ch_uox.alloc(n));
    CTRY(s->ch_nbr.alloc((size_t)n * cap));
    CTRY(s->ch_nbr_d.alloc(3 * (size_t)n * cap));
    CTRY(s->ch_boltz.alloc(nrmax));
    CTRY(s->ch_bsum.alloc(n));
    CTRY(s->ch_status.alloc(4));
    CTRY(s->ch_pave.alloc(p->n_bins));
    CTRY(s->ch_inave.alloc(p->n_bins));
    CTRY(hipMemset(s->ch_pave.p, 0, sizeof(double) * p->n_bins));
    CTRY(hipMemset(s->ch_inave.p, 0, sizeof(unsigned long long) * p->n_bins));
    CTRY(hipMemset(s->ch_status.p, 0, sizeof(int) * 4));
    CTRY(hipHostMalloc((void **)&s->h_ins, sizeof(double) * 3 * n, hipHostMallocDefault));
    CTRY(hipHostMalloc((void **)&s->h_rot, sizeof(double) * 6 * nrmax, hipHostMallocDefault));
    CTRY(hipHostMalloc((void **)&s->h_il, sizeof(int) * n, hipHostMallocDefault));
    CTRY(hipHostMalloc((void **)&s->h_insrot, sizeof(int) * (n + 1), hipHostMallocDefault));
    CTRY(hipHostMalloc((void **)&s->h_rotins, sizeof(int) * nrmax, hipHostMallocDefault));
    CTRY(hipHostMalloc((void **)&s->h_status, sizeof(int) * 4, hipHostMallocDefault));
    CTRY(hipEventCreateWithFlags(&s->ch_done, hipEventDisableTiming));
#undef CTRY
    s->ch_seed_pos.assign(n, 0.0);
    ChemArgs &c = s->chem;
    memset(&c, 0, sizeof c);
    c.ins = s->ch_ins.p; c.ins_il = s->ch_il.p; c.ins_rot = s->ch_insrot.p; c.rot_ins = s->ch_rotins.p;
    c.rot = s->ch_rot.p; c.overlap = s->ch_over.p; c.n_nbr = s->ch_nnbr.p; c.u_ox = s->ch_uox.p;
    c.nbr = s->ch_nbr.p; c.nbr_d = s->ch_nbr_d.p; c.boltz = s->ch_boltz.p; c.bsum = s->ch_bsum.p; c.status = s->ch_status.p;
    c.pave = s->ch_pave.p; c.inave = s->ch_inave.p; c.dipz = s->dipz.p;
    c.kvec = s->d_kvec.p; c.ak = s->d_ak.p; c.sk = s->d_s.p;
    c.n_ins = n; c.cap = cap; c.n_bins = p->n_bins; c.nk = (int)s->ak.size();
    c.klmax = klmax; c.kmmax = kmmax; c.knmax = knmax;
    const double chtemper = p->temperature > 0.0 ? p->temperature : 300.0;
    c.beta = 1000.0 / (lit(8.31441) * chtemper);
    c.rcut = s->par.rcut;
    c.rinsmin = 4.0;
    c.rwinsmin = ref ? (double)(2.35f * 2.35f) : 2.35 * 2.35;
    c.qsw[0] = lit(-0.8476); c.qsw[1] = lit(0.4238); c.qsw[2] = lit(0.4238);
    c.crf_scale = p->ew3dc ? 2.0 * s->crf0 * s->S.facq : 0.0;
    c.lwh = lwh;
    c.pipix = 2.0 * s->c.pi / s->par.box[0];
    c.pipiy = 2.0 * s->c.pi / s->par.box[1];
    c.pipiz = 2.0 * s->c.pi / s->par.box[2];
    c.ulr_w = s->wl.ulr_w[1]; c.ulr_l = s->wl.ulr_l[1]; c.lja = s->wl.flr_lja[1]; c.ljc = s->wl.flr_ljc[1];
    c.bua = s->wl.flr_bua[1]; c.buc = s->wl.flr_buc[1]; c.bur = s->wl.flr_bur[1];
    s->chem_on = true;
    return MDF_OK;
}

int mdf_get_chempot(mdf_system *s, double *pave, int64_t *inave, double *seed, int32_t reset) {
    if (!s) return fail(MDF_ERR_ARG, "null handle");
    if (!s->chem_on) return fail(MDF_ERR_ARG, "insertions are off");
    HIPCHK(hipSetDevice(s->device));
    HIPCHK(hipStreamSynchronize(s->stream));
    int rc = chem_resolve(s);
    if (rc) return rc;
    const int nb = s->chem_p.n_bins;
    if (pave) HIPCHK(hipMemcpy(pave, s->ch_pave.p, sizeof(double) * nb, hipMemcpyDeviceToHost));
    if (inave) HIPCHK(hipMemcpy(inave, s->ch_inave.p, sizeof(int64_t) * nb, hipMemcpyDeviceToHost));
    if (seed) *seed = s->chem_seed;
    if (reset) {
        HIPCHK(hipMemset(s->ch_pave.p, 0, sizeof(double) * nb));
        HIPCHK(hipMemset(s->ch_inave.p, 0, sizeof(unsigned long long) * nb));
    }
    return MDF_OK;
}

int mdf_mts_get_stats(mdf_system *s, mdf_mts_stats *out) {
    if (!s || !out) return fail(MDF_ERR_ARG, "null argument");
    if (!s->mts_on) return fail(MDF_ERR_ARG, "multiple time steps are off");
    HIPCHK(hipSetDevice(s->device));
    HIPCHK(hipStreamSynchronize(s->stream));
    MtsState st;
    HIPCHK(hipMemcpy(&st, s->mts_st.p, sizeof st, hipMemcpyDeviceToHost));
    out->state = st.state;
    out->last_exact = st.exact;
    out->n_exact = st.n_exact;
    out->n_surrogate = st.n_surrogate;
    out->n_fallback = st.n_fallback;
    out->n_judged = st.n_judged;
    out->adam_steps = st.adam_t;
    out->err_last = st.err_last;
    out->err_max = st.err_max;
    out->err_mean = st.n_judged ? st.err_sum / st.n_judged : 0.0;
    out->loss_last = st.loss_last;
    return MDF_OK;
}

int mdf_mts_get_weights(mdf_system *s, float *w, int32_t cap) {
    if (!s || !w) return fail(MDF_ERR_ARG, "null argument");
    if (!s->mts_on) return fail(MDF_ERR_ARG, "multiple time steps are off");
    if (cap < MDF_MTS_WEIGHTS) return fail(MDF_ERR_ARG, "capacity %d < %d", cap, MDF_MTS_WEIGHTS);
    HIPCHK(hipSetDevice(s->device));
    HIPCHK(hipStreamSynchronize(s->stream));
    HIPCHK(hipMemcpy(w, s->sg_params.p, sizeof(float) * SG_NP, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(w + SG_NP, s->sg_norm.p, sizeof(float) * SG_NORM_N, hipMemcpyDeviceToHost));
    return MDF_OK;
}

static int check_weights(const float *w) {
    for (int k = 0; k < MDF_MTS_WEIGHTS; k++)
        if (!std::isfinite(w[k])) return fail(MDF_ERR_ARG, "weights: entry %d is not finite", k);
    for (int k = SG_D; k < SG_NORM_N; k++)
        if (!(w[SG_NP + k] > 0.0f)) return fail(MDF_ERR_ARG, "weights: scale %d must be positive", k);
    return MDF_OK;
}

int mdf_mts_set_weights(mdf_system *s, const float *w, int32_t n) {
    if (!s || !w) return fail(MDF_ERR_ARG, "null argument");
    if (!s->mts_on) return fail(MDF_ERR_ARG, "multiple time steps are off");
    if (n != MDF_MTS_WEIGHTS) return fail(MDF_ERR_ARG, "expected %d floats, got %d", MDF_MTS_WEIGHTS, n);
    int rc = check_weights(w);
    if (rc) return rc;
    HIPCHK(hipSetDevice(s->device));
    HIPCHK(hipStreamSynchronize(s->stream));
    drain_all(s);
    return mts_load(s, w, w + SG_NP, true);
}

int mdf_mts_features(mdf_system *s, float *feat, float *vec) {
    if (!s || !feat || !vec) return fail(MDF_ERR_ARG, "null argument");
    if (!s->mts_on || s->mts_model != MDF_MTS_NETWORK)
        return fail(MDF_ERR_ARG, "the surrogate's inputs need the scheme on with the network model");
    HIPCHK(hipSetDevice(s->device));
    HIPCHK(hipStreamSynchronize(s->stream));
    const MtsArgs ma = mts_args(s);
    k_sg_pack<<<std::max(1, (s->top.n_total + 255) / 256), 256, 0, s->stream>>>(s->S, ma);
    k_sg_features<SG_FEAT_G><<<std::max(1, (s->n_lead + 256 / SG_FEAT_G - 1) / (256 / SG_FEAT_G)), 256, 0,
                               s->stream>>>(s->S, ma);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s->stream));
    const size_t nm = s->top.n_move;
    HIPCHK(hipMemcpy(feat, s->sg_feat.p, sizeof(float) * nm * SG_D, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(vec, s->sg_vec.p, sizeof(float) * nm * SG_M * 3, hipMemcpyDeviceToHost));
    return MDF_OK;
}

int mdf_mts_eval(const float *weights, int32_t n, const float *feat, const float *vec, const double *target,
                 const float *feat_b, const float *vec_b, const double *target_b, float lr, int32_t adam_t,
                 float *f_out, float *grad_out, float *params_out) {
    if (!weights || !feat || !vec || n <= 0) return fail(MDF_ERR_ARG, "null argument or n <= 0");
    if (adam_t < 0 || !(lr >= 0.0f)) return fail(MDF_ERR_ARG, "adam_t >= 0 and lr >= 0 required");
    const bool train = target != nullptr;
    if (train && (!feat_b || !vec_b || !target_b)) return fail(MDF_ERR_ARG, "training needs the second frame");
    int rc = check_weights(weights);
    if (rc) return rc;
    double t2 = 0.0;
    if (train) {
        for (size_t k = 0; k < 3 * (size_t)n; k++) t2 += (target[k] - target_b[k]) * (target[k] - target_b[k]);
        t2 /= n;
        if (!(t2 > 0.0)) return fail(MDF_ERR_ARG, "the two targets are equal");
    }
    int n_sblk, n_iblk, n_tblk;
    mts_geometry(n, n_sblk, n_iblk, n_tblk);
    DBuf<MtsState> st;
    DBuf<float> d_feat, d_vec, d_feat_b, d_vec_b, d_sg, d_params, d_m, d_v, d_gsum, d_norm, d_gpart;
    DBuf<unsigned short> d_wb;
    DBuf<double> d_frec, d_frec_b;
    hipStream_t stream = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&stream, hipStreamNonBlocking);
    std::vector<unsigned short> wb(SG_WB_N);
    host_weight_copies(weights, wb.data());
    MtsState hs;
    memset(&hs, 0, sizeof hs);
    hs.norm_ready = 1;
    hs.have_prev = train ? 1 : 0;
    hs.train_now = train ? 1 : 0;
    hs.t2 = t2;
    hs.adam_base = adam_t;
    hs.adam_t = adam_t + (train ? 1 : 0);
    const size_t nn = (size_t)n;
#define MTRY(x) do { if (e == hipSuccess) e = (x); } while (0)
    MTRY(st.alloc(1));
    MTRY(hipMemcpy(st.p, &hs, sizeof hs, hipMemcpyHostToDevice));
    MTRY(upload(d_feat, feat, nn * SG_D));
    MTRY(upload(d_vec, vec, nn * SG_M * 3));
    MTRY(d_sg.alloc(3 * nn));
    MTRY(upload(d_params, weights, SG_NP));
    MTRY(upload(d_norm, weights + SG_NP, SG_NORM_N));
    MTRY(upload(d_wb, wb.data(), SG_WB_N));
    MTRY(d_m.alloc(SG_NP));
    MTRY(d_v.alloc(SG_NP));
    MTRY(hipMemset(d_m.p, 0, sizeof(float) * SG_NP));
    MTRY(hipMemset(d_v.p, 0, sizeof(float) * SG_NP));
    MTRY(d_gsum.alloc(SG_GP));
    MTRY(d_gpart.alloc((size_t)SG_TRAIN_BLOCKS * SG_GP));
    if (train) {
        MTRY(upload(d_frec, target, 3 * nn));
        MTRY(upload(d_frec_b, target_b, 3 * nn));
        MTRY(upload(d_feat_b, feat_b, nn * SG_D));
        MTRY(upload(d_vec_b, vec_b, nn * SG_M * 3));
    }
    if (e == hipSuccess) {
        MtsArgs a;
        memset(&a, 0, sizeof a);
        a.st = st.p; a.model = 1; a.n_sblk = n_sblk; a.n_tblk = n_tblk; a.lr = lr; a.train_iters = 1;
        a.train_stride = 1;
        a.feat = d_feat.p; a.vec = d_vec.p; a.sg = d_sg.p; a.frec = d_frec.p;
        a.feat_prev = d_feat_b.p; a.vec_prev = d_vec_b.p; a.frec_prev = d_frec_b.p;
        a.params = d_params.p; a.adam_m = d_m.p; a.adam_v = d_v.p; a.gsum = d_gsum.p;
        a.wb = d_wb.p; a.norm = d_norm.p; a.gpart = d_gpart.p;
        k_sg_infer<<<n_iblk, 256, 0, stream>>>(a, n, 0, nullptr);
        if (train) {
            k_sg_train<<<n_tblk, 256, 0, stream>>>(a, n, 0, nullptr);
            k_sg_adam<<<(SG_GP + 255) / 256, 256, 0, stream>>>(a, n_tblk, 0, nullptr);
        }
        MTRY(hipGetLastError());
        MTRY(hipStreamSynchronize(stream));
        if (f_out) MTRY(hipMemcpy(f_out, d_sg.p, sizeof(float) * 3 * nn, hipMemcpyDeviceToHost));
        if (train && grad_out) MTRY(hipMemcpy(grad_out, d_gsum.p, sizeof(float) * SG_GP, hipMemcpyDeviceToHost));
        if (train && params_out) MTRY(hipMemcpy(params_out, d_params.p, sizeof(float) * SG_NP, hipMemcpyDeviceToHost));
    }
#undef MTRY
    st.release(); d_feat.release(); d_vec.release(); d_feat_b.release(); d_vec_b.release(); d_sg.release();
    d_params.release(); d_m.release(); d_v.release(); d_gsum.release(); d_norm.release(); d_gpart.release();
    d_wb.release(); d_frec.release(); d_frec_b.release();
    if (stream) (void)hipStreamDestroy(stream);
    if (e != hipSuccess) return fail(MDF_ERR_HIP, "surrogate evaluation: %s", hipGetErrorString(e));
    return MDF_OK;
}

// ---- the same evaluation allocating nothing (mdf_mts_eval_device) ------------------------
// Its scratch - the scheme state, the FP32 parameters being trained, input normalisation,
// bfloat16 weight copies, Adam moments, the gradient sum and the training kernel's per-block
// gradient partials - is one static device array.  Each device holds its own copy from the
// load of this library to the end of the process; every call on a device shares it.
static constexpr size_t ev_align(size_t b) { return (b + 255) & ~(size_t)255; }
constexpr size_t EV_OFF_ST = 0;
constexpr size_t EV_OFF_PARAMS = EV_OFF_ST + ev_align(sizeof(MtsState));
constexpr size_t EV_OFF_NORM = EV_OFF_PARAMS + ev_align(sizeof(float) * SG_NP);
constexpr size_t EV_OFF_WB = EV_OFF_NORM + ev_align(sizeof(float) * SG_NORM_N);
constexpr size_t EV_OFF_M = EV_OFF_WB + ev_align(sizeof(unsigned short) * SG_WB_N);
constexpr size_t EV_OFF_V = EV_OFF_M + ev_align(sizeof(float) * SG_NP);
constexpr size_t EV_OFF_GSUM = EV_OFF_V + ev_align(sizeof(float) * SG_NP);
constexpr size_t EV_OFF_GPART = EV_OFF_GSUM + ev_align(sizeof(float) * SG_GP);
constexpr size_t EV_SCRATCH_BYTES = EV_OFF_GPART + ev_align(sizeof(float) * SG_TRAIN_BLOCKS * SG_GP);
static_assert(EV_SCRATCH_BYTES == MDF_MTS_EVAL_SCRATCH_BYTES, "the header states the scratch size");

static __device__ __align__(256) unsigned char scratch[EV_SCRATCH_BYTES];

constexpr int EV_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static unsigned char *g_scratch_dev[EV_MAX_DEVICES] = {};

// the address of this device's copy of `scratch` (the host-side name is no device address),
// looked up once per device with that device current
static int scratch_on_device(int dev, unsigned char **out) {
    if (dev < 0 || dev >= EV_MAX_DEVICES) return fail(MDF_ERR_ARG, "device %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(MDF_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
        g_scratch_dev[dev] = static_cast<unsigned char *>(p);
    }
    *out = g_scratch_dev[dev];
    return MDF_OK;
}

int mdf_mts_eval_device(const float *weights, int32_t n, const float *feat, const float *vec, const double *target,
                        const float *feat_b, const float *vec_b, const double *target_b, float lr, int32_t adam_t,
                        float *f_out, float *grad_out, float *params_out, void *stream) {
    if (!weights || !feat || !vec || !f_out) return fail(MDF_ERR_ARG, "null argument");
    if (n <= 0 || n > MDF_MTS_EVAL_MAX_SITES)
        return fail(MDF_ERR_ARG, "n = %d sites: 1 .. %d supported", n, MDF_MTS_EVAL_MAX_SITES);
    if (adam_t < 0 || !(lr >= 0.0f)) return fail(MDF_ERR_ARG, "adam_t >= 0 and lr >= 0 required");
    const bool train = target != nullptr;
    if (train && (!feat_b || !vec_b || !target_b)) return fail(MDF_ERR_ARG, "training needs the second frame");
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    unsigned char *sp = nullptr;
    int rc = scratch_on_device(dev, &sp);
    if (rc) return rc;
    int n_sblk, n_iblk, n_tblk;
    mts_geometry(n, n_sblk, n_iblk, n_tblk);
    MtsArgs a;
    memset(&a, 0, sizeof a);
    a.st = reinterpret_cast<MtsState *>(sp + EV_OFF_ST);
    a.model = 1; a.n_sblk = n_sblk; a.n_tblk = n_tblk; a.lr = lr; a.train_iters = 1; a.train_stride = 1;
    a.feat = const_cast<float *>(feat); a.vec = const_cast<float *>(vec); a.sg = f_out;
    a.frec = const_cast<double *>(target);
    a.feat_prev = const_cast<float *>(feat_b); a.vec_prev = const_cast<float *>(vec_b);
    a.frec_prev = const_cast<double *>(target_b);
    // trained parameters go straight to params_out; a forward pass leaves it alone
    a.params = (train && params_out) ? params_out : reinterpret_cast<float *>(sp + EV_OFF_PARAMS);
    a.norm = reinterpret_cast<float *>(sp + EV_OFF_NORM);
    a.wb = reinterpret_cast<unsigned short *>(sp + EV_OFF_WB);
    a.adam_m = reinterpret_cast<float *>(sp + EV_OFF_M);
    a.adam_v = reinterpret_cast<float *>(sp + EV_OFF_V);
    a.gsum = (train && grad_out) ? grad_out : reinterpret_cast<float *>(sp + EV_OFF_GSUM);
    a.gpart = reinterpret_cast<float *>(sp + EV_OFF_GPART);
    hipStream_t st = (hipStream_t)stream;
    k_sg_eval_prep<<<1, 1024, 0, st>>>(a, weights, n, target, target_b, adam_t);
    k_sg_infer<<<n_iblk, 256, 0, st>>>(a, n, 0, nullptr);
    if (train) {
        k_sg_train<<<n_tblk, 256, 0, st>>>(a, n, 0, nullptr);
        k_sg_adam<<<(SG_GP + 255) / 256, 256, 0, st>>>(a, n_tblk, 0, nullptr);
    }
    HIPCHK(hipGetLastError());
    return MDF_OK;
}

}  // extern "C"

// ---- batches --------------------------------------------------------------------------
// Many independent systems stepped together: every stage is one launch over all of them
// (block maps give each block its system), so small systems fill the GPU together and a
// step of the whole batch costs one graph launch.
enum { BR_VALU, BR_MMA32, BR_MMA64 };
enum { BK_GATHER, BK_SKIN, BK_COUNT, BK_FILL, BK_ORDER, BK_BUILD, BK_PAIR, BK_SF, BK_SFR, BK_BP, BK_EPI,
       BK_COLSUM, BK_N };

struct mdf_batch {
    std::vector<mdf_system *> sys;
    int device = 0;
    hipStream_t stream = nullptr;
    // with MDF_FORK_KSPACE_BATCH=1 the k-space kernels run on a second stream beside the skin
    // test, list build and pair kernel (forked after the gather, joined before the
    // epilogue); off by default: the 64-replica batch measured 4% slower forked
    hipStream_t stream2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    bool fork = false;
    std::vector<RepArgs> h_rep;
    DBuf<RepArgs> d_rep;
    DBuf<int2> maps[BK_N];
    int nmap[BK_N] = {};
    // reciprocal sum: the vector-ALU kernels, or the matrix-core kernels of the FP32 or FP64
    // form when every system runs those alone (then one launch per kernel template: systems
    // grouped by kn-tile count, or by K-steps and site groups of the FP64 back-projection)
    int rmode = 0;   // BR_*
    static constexpr int NGRP = 40;
    struct Group { int key = 0, n = 0, threads = 256; size_t lds = 0; DBuf<int2> map; };
    Group gsf[NGRP], gbp[NGRP];
    int n_gsf = 0, n_gbp = 0;
    std::vector<const double *> h_in;   // the last step's input pointers
    // a ring of input-pointer tables (pinned host staging, device copy) and of the step graphs
    // that read them: a step reuses a slot only after the step NB_RING back has finished, so
    // the host never waits for the step before it
    static constexpr int NB_RING = 4;
    const double **h_in_ring[NB_RING] = {};
    DBuf<const double *> d_in_ring[NB_RING];
    hipGraphExec_t graph_ring[NB_RING] = {};
    hipEvent_t done_ev[NB_RING] = {};
    bool slot_used[NB_RING] = {};
    int slot = 0, last_slot = 0;
    DBuf<DevFlags> d_flags;
    std::vector<DevFlags> h_flags;
    hipEvent_t in_event = nullptr;
    bool pending = false;
    int pair_g = 16;
    bool pair_strided = false;
    bool one_cls = false;   // every system has one moving lead type
    bool f32 = false, vir = false, acc = false;
};

static void fill_rep(mdf_system *s, RepArgs &r) {
    memset(&r, 0, sizeof r);
    const int nm = s->top.n_move, nr = s->top.n_rig, nk = (int)s->ak.size();
    r.S = s->S;
    r.dip_part = s->dip_part.p;
    r.dipz = s->dipz.p;
    r.skin_blocks = s->skin_blocks;
    r.force_rebuild = 0;
    r.lead_cell = s->lead_cell.p;
    r.lead_blocks = std::max(1, (s->n_lead + 255) / 256);
    r.build_blocks = s->build_blocks;
    r.order_blocks = s->order_blocks;
    r.items_tmp = s->cell_items_tmp.p;
    r.xyz_tmp = s->cell_xyz_tmp.p;
    r.ncells_xy = s->ncells_xy;
    r.gather_blocks = std::max(1, (std::max(nm, nr) + 255) / 256);
    const RecipArgs a = recip_args(s);
    r.rows = s->d_rows.p;
    r.nrows = (int)s->rows.size();
    r.cpb = s->cpb;
    r.n_sb = s->n_sb;
    r.n_cg = s->n_cg;
    r.nk = nk;
    r.n_kblk = s->n_kblk;
    r.n_rg = s->n_rg;
    r.n_bpsb = std::max(1, (nm + 256 * BP_NS - 1) / (256 * BP_NS));
    r.sf_partial = s->sf_partial.p;
    r.s_wall = s->d_swall.p;
    r.ak = s->d_ak.p;
    r.kvec = s->d_kvec.p;
    r.s_out = s->d_s.p;
    r.t_out = s->d_t.p;
    r.kpart = s->kpart.p;
    r.bp_row_start = s->d_bp_row_start.p;
    r.fpart = s->fpart.p;
    r.cux = a.cux; r.cuy = a.cuy; r.cuz = a.cuz;
    r.pipix = a.pipix; r.pipiy = a.pipiy; r.pipiz = a.pipiz;
    r.lamb = 1.0 / (4.0 * s->S.alfa * s->S.alfa);
    r.wl = s->d_wl.p;
    r.ew3dc = s->par.ew3dc;
    r.n_groups = (nk > 0 && nm > 0) ? s->n_rg : 0;
    r.n_epblk = s->n_epblk;
    r.colper = s->colper;
    r.n_colblk = s->n_colblk;
    r.crf0 = s->crf0;
    r.eppart = s->eppart.p;
    r.fcom = s->fcom.p;
    r.tq = s->torque.p;
    r.colpart = s->colpart.p;
    r.e_vdw = s->e_vdw.p;
    r.e_coul = s->e_coul.p;
    r.lead_slot = s->lead_slot.p;
    r.scal = s->d_scal.p;
    r.A = final_args(s);
    r.A.n_kblk = nk > 0 ? s->n_kblk : 0;   // the batch reduces k partials in blocks of SFR_K k-vectors
    // the matrix-core forms (used when the batch runs them: the k partials are per row then)
    r.knmax = s->knmax; r.n_nt = s->n_nt; r.sfm_sites = s->sfm_sites;
    r.sfm_nsb = s->n_sb_mma; r.sfm_nrg = s->n_rg_mma;
    r.rowkk = s->d_rowkk.p; r.tile_ks = s->d_tile_ks.p; r.tpack = s->d_tpack.p; r.tscale = s->d_tscale.p;
    r.klmax = s->klmax; r.kmmax = s->kmmax;
    r.s64_nsb = s->s64_nsb; r.s64_pcols = s->s64_pcols; r.s64_sites = s->s64_sites;
    r.b64_nsb = std::max(1, (nm + s->b64_sites - 1) / s->b64_sites);
    r.s64_rowkk = s->d_s64_rowkk.p; r.s64_waves = s->d_s64_waves.p; r.b64_tile = s->d_b64_tile.p;
    r.b64_rg = s->d_b64_rg.p; r.s64_pncol = s->d_s64_pncol.p; r.s64_pcoff = s->d_s64_pcoff.p;
    r.s64_sorted_of = s->d_s64_sorted_of.p; r.s64_part = s->d_s64_part.p; r.b64_rinfo = s->d_b64_rinfo.p;
    r.b64_tpk = s->d_b64_tpk.p;
}

// the batch's k-sum on the matrix cores: every system's partials, T and k partials in the
// layout its own kernels use
static void rep_mode(const mdf_batch *b, const mdf_system *s, RepArgs &r) {
    const int nk = (int)s->ak.size(), nm = s->top.n_move;
    if (b->rmode == BR_MMA32) {
        r.n_groups = (nk > 0 && nm > 0) ? 1 : 0;
        r.A.n_kblk = nk > 0 ? (int)s->rows.size() : 0;
    } else if (b->rmode == BR_MMA64) {
        r.n_groups = (nk > 0 && nm > 0) ? s->b64_nrg : 0;
        r.A.n_kblk = nk > 0 ? (int)s->rows.size() : 0;
    }
}

// block counts of every batched launch for one system
static void rep_blocks(const mdf_batch *b, const mdf_system *s, const RepArgs &r, int cnt[BK_N]) {
    const int nk = r.nk, nm = s->top.n_move;
    cnt[BK_GATHER] = r.gather_blocks;
    cnt[BK_SKIN] = r.skin_blocks;
    cnt[BK_COUNT] = r.lead_blocks;
    cnt[BK_FILL] = r.lead_blocks;
    cnt[BK_ORDER] = r.order_blocks;
    cnt[BK_BUILD] = r.build_blocks;
    cnt[BK_PAIR] = std::max(1, (s->n_lead + (256 / b->pair_g) - 1) / (256 / b->pair_g));
    cnt[BK_SF] = (nk > 0 && nm > 0) ? r.n_sb * r.n_cg : 0;
    cnt[BK_SFR] = nk > 0 ? r.n_kblk : 0;
    cnt[BK_BP] = (nk > 0 && nm > 0) ? r.n_bpsb * r.n_rg : 0;
    if (b->rmode != BR_VALU) {   // the matrix-core kernels go by template groups (batch_groups)
        cnt[BK_SF] = 0;
        cnt[BK_BP] = 0;
        cnt[BK_SFR] = nk > 0 ? r.nrows : 0;
    }
    cnt[BK_EPI] = r.n_epblk;
    cnt[BK_COLSUM] = NCOL * r.n_colblk;
}

// per-system arguments and block maps to the device; the captured graph reads both at run
// time, so it stays valid when only their contents change
// block maps of the matrix-core k-sum, one per kernel template: structure factor by kn tiles
// (FP32 n_nt, FP64 s64_nt), back-projection by kn tiles (FP32) or by K-steps and site groups
// (FP64); each with the largest dynamic LDS and the waves per block of its systems
static int batch_groups(mdf_batch *b) {
    for (int g = 0; g < mdf_batch::NGRP; g++) { b->gsf[g].map.release(); b->gbp[g].map.release(); b->gsf[g].n = b->gbp[g].n = 0; }
    b->n_gsf = b->n_gbp = 0;
    if (b->rmode == BR_VALU) return MDF_OK;
    std::vector<int2> msf[mdf_batch::NGRP], mbp[mdf_batch::NGRP];
    auto slot = [](mdf_batch::Group *gs, int &ng, int key) {
        for (int g = 0; g < ng; g++) if (gs[g].key == key) return g;
        gs[ng].key = key; gs[ng].lds = 0; gs[ng].threads = 256;
        return ng++;
    };
    for (int i = 0; i < (int)b->sys.size(); i++) {
        const mdf_system *s = b->sys[i];
        const int nk = (int)s->ak.size(), nm = s->top.n_move;
        if (nk == 0 || nm == 0) continue;
        const RepArgs &r = b->h_rep[i];
        if (b->rmode == BR_MMA32) {
            const int g = slot(b->gsf, b->n_gsf, s->n_nt);
            b->gsf[g].lds = std::max(b->gsf[g].lds, sfm_lds_bytes(s->sfm_sites, s->n_nt));
            for (int l = 0; l < r.sfm_nsb * r.sfm_nrg; l++) msf[g].push_back(make_int2(i, l));
            const int h = slot(b->gbp, b->n_gbp, s->n_nt);
            b->gbp[h].threads = 64 * s->bpm_waves;   // the same for every system of one n_nt
            const int ns = 16 * s->bpm_waves;
            for (int l = 0; l < (nm + ns - 1) / ns; l++) mbp[h].push_back(make_int2(i, l));
        } else {
            const int g = slot(b->gsf, b->n_gsf, s->s64_nt);
            b->gsf[g].lds = std::max(b->gsf[g].lds, sf64_lds_bytes(s->klmax, s->kmmax, s->s64_nt));
            for (int l = 0; l < r.s64_nsb * s->s64_npass; l++) msf[g].push_back(make_int2(i, l));
            const int h = slot(b->gbp, b->n_gbp, s->b64_nks * 4 + s->b64_ng);
            for (int l = 0; l < r.b64_nsb * s->b64_nrg; l++) mbp[h].push_back(make_int2(i, l));
        }
    }
    for (int g = 0; g < b->n_gsf; g++) {
        b->gsf[g].n = (int)msf[g].size();
        HIPCHK(upload(b->gsf[g].map, msf[g].data(), msf[g].size()));
    }
    for (int g = 0; g < b->n_gbp; g++) {
        b->gbp[g].n = (int)mbp[g].size();
        HIPCHK(upload(b->gbp[g].map, mbp[g].data(), mbp[g].size()));
    }
    return MDF_OK;
}

static int batch_upload(mdf_batch *b, bool maps) {
    const int n = (int)b->sys.size();
    b->h_rep.resize(n);
    std::vector<int2> mp[BK_N];
    for (int i = 0; i < n; i++) {
        fill_rep(b->sys[i], b->h_rep[i]);
        rep_mode(b, b->sys[i], b->h_rep[i]);
        int cnt[BK_N];
        rep_blocks(b, b->sys[i], b->h_rep[i], cnt);
        for (int k = 0; k < BK_N; k++)
            for (int l = 0; l < cnt[k]; l++) mp[k].push_back(make_int2(i, l));
    }
    HIPCHK(hipMemcpyAsync(b->d_rep.p, b->h_rep.data(), sizeof(RepArgs) * n, hipMemcpyHostToDevice, b->stream));
    if (maps) {
        for (int k = 0; k < BK_N; k++) {
            b->maps[k].release();
            b->nmap[k] = (int)mp[k].size();
            HIPCHK(upload(b->maps[k], mp[k].data(), mp[k].size()));
        }
        int rc = batch_groups(b);
        if (rc != MDF_OK) return rc;
    }
    HIPCHK(hipStreamSynchronize(b->stream));
    return MDF_OK;
}

static void batch_drop_graphs(mdf_batch *b) {
    for (int k = 0; k < mdf_batch::NB_RING; k++)
        if (b->graph_ring[k]) { (void)hipGraphExecDestroy(b->graph_ring[k]); b->graph_ring[k] = nullptr; }
}

// the batch's k-space kernels on stream st: structure factor, reduction, back-projection
static void batch_kspace(mdf_batch *b, hipStream_t st) {
    const RepArgs *R = b->d_rep.p;
    auto grid = [&](int k) { return b->nmap[k]; };
    if (b->rmode == BR_MMA32) {
        for (int g = 0; g < b->n_gsf; g++) {
            const mdf_batch::Group &G = b->gsf[g];
            switch (G.key) {
#define SFMB(NT) case NT: k_sf_mma_b<NT><<<G.n, 256, G.lds, st>>>(R, G.map.p); break;
                SFMB(1) SFMB(2) SFMB(3) SFMB(4) SFMB(5) SFMB(6) SFMB(7) SFMB(8)
#undef SFMB
            }
        }
        if (grid(BK_SFR)) k_sf_reduce_rows_b<<<grid(BK_SFR), SFR_THREADS, 0, st>>>(R, b->maps[BK_SFR].p);
        for (int g = 0; g < b->n_gbp; g++) {
            const mdf_batch::Group &G = b->gbp[g];
            switch (G.key) {
#define BPMB(NT) case NT: k_bp_mma_b<NT><<<G.n, G.threads, 0, st>>>(R, G.map.p); break;
                BPMB(1) BPMB(2) BPMB(3) BPMB(4) BPMB(5) BPMB(6) BPMB(7) BPMB(8)
#undef BPMB
            }
        }
    } else if (b->rmode == BR_MMA64) {
        for (int g = 0; g < b->n_gsf; g++) {
            const mdf_batch::Group &G = b->gsf[g];
            switch (G.key) {
#define SF64B(NT) case NT: k_sf64_b<NT><<<G.n, 64 * SF64_WAVES, G.lds, st>>>(R, G.map.p); break;
                SF64B(1) SF64B(2) SF64B(3) SF64B(4) SF64B(5) SF64B(6) SF64B(7) SF64B(8)
#undef SF64B
            }
        }
        if (grid(BK_SFR)) k_sf64_reduce_b<<<grid(BK_SFR), SFR_THREADS, 0, st>>>(R, b->maps[BK_SFR].p);
        for (int g = 0; g < b->n_gbp; g++) {
            const mdf_batch::Group &G = b->gbp[g];
            switch (G.key) {
#define BP64B(NKS, NG) case NKS * 4 + NG: k_bp64_b<NKS, NG><<<G.n, 256, 0, st>>>(R, G.map.p); break;
                BP64B(4, 1) BP64B(8, 1) BP64B(12, 1) BP64B(16, 1) BP64B(20, 1) BP64B(24, 1) BP64B(28, 1)
                BP64B(32, 1) BP64B(4, 2) BP64B(8, 2) BP64B(12, 2) BP64B(16, 2) BP64B(20, 2) BP64B(24, 2)
                BP64B(28, 2) BP64B(32, 2) BP64B(9, 3) BP64B(10, 3) BP64B(11, 3) BP64B(12, 3) BP64B(13, 3)
                BP64B(14, 3) BP64B(15, 3) BP64B(16, 3) BP64B(17, 3) BP64B(18, 3) BP64B(19, 3) BP64B(20, 3)
#undef BP64B
            }
        }
    } else {
        if (grid(BK_SF)) {
            if (b->f32) k_sf_b<float><<<grid(BK_SF), SF_BLOCK, 0, st>>>(R, b->maps[BK_SF].p);
            else k_sf_b<double><<<grid(BK_SF), SF_BLOCK, 0, st>>>(R, b->maps[BK_SF].p);
        }
        if (grid(BK_SFR)) k_sf_reduce_b<<<grid(BK_SFR), SFR_THREADS, 0, st>>>(R, b->maps[BK_SFR].p);
        if (grid(BK_BP)) {
            if (b->f32) k_bp_b<float><<<grid(BK_BP), 256, 0, st>>>(R, b->maps[BK_BP].p);
            else k_bp_b<double><<<grid(BK_BP), 256, 0, st>>>(R, b->maps[BK_BP].p);
        }
    }
}

static void batch_launch(mdf_batch *b, const double *const *d_in) {
    const int n = (int)b->sys.size();
    hipStream_t st = b->stream;
    const RepArgs *R = b->d_rep.p;
    auto grid = [&](int k) { return b->nmap[k]; };
    if (grid(BK_GATHER)) k_gather_b<<<grid(BK_GATHER), 256, 0, st>>>(R, b->maps[BK_GATHER].p, d_in);
    // the k-space kernels need only the gathered positions: forked here onto the second
    // stream, joined before the epilogue
    if (b->fork) {
        (void)hipEventRecord(b->ev_fork, st);
        (void)hipStreamWaitEvent(b->stream2, b->ev_fork, 0);
        batch_kspace(b, b->stream2);
        (void)hipEventRecord(b->ev_join, b->stream2);
    }
    k_skin_b<<<grid(BK_SKIN), 256, 0, st>>>(R, b->maps[BK_SKIN].p);
    k_decide_b<<<n, 256, 0, st>>>(R);
    k_cell_count_b<<<grid(BK_COUNT), 256, 0, st>>>(R, b->maps[BK_COUNT].p);
    k_cell_scan_build_b<<<n, 1024, 0, st>>>(R);
    k_cell_fill_b<<<grid(BK_FILL), 256, 0, st>>>(R, b->maps[BK_FILL].p);
    k_cell_order_b<<<grid(BK_ORDER), 256, 0, st>>>(R, b->maps[BK_ORDER].p);
    k_build_b<<<grid(BK_BUILD), BUILD_THREADS, 0, st>>>(R, b->maps[BK_BUILD].p);
    {
        const int2 *mp = b->maps[BK_PAIR].p;
        const int g = grid(BK_PAIR);
#define MDF_PB3(RT, G, ST, NM)                                                               \
    do {                                                                                     \
        if (b->acc) {                                                                        \
            if (b->vir) k_pair_b<RT, G, true, true, ST, NM><<<g, 256, 0, st>>>(R, mp);       \
            else k_pair_b<RT, G, false, true, ST, NM><<<g, 256, 0, st>>>(R, mp);             \
        } else {                                                                             \
            if (b->vir) k_pair_b<RT, G, true, false, ST, NM><<<g, 256, 0, st>>>(R, mp);      \
            else k_pair_b<RT, G, false, false, ST, NM><<<g, 256, 0, st>>>(R, mp);            \
        }                                                                                    \
    } while (0)
#define MDF_PB2(RT, G, ST)                                  \
    do {                                                    \
        if (b->one_cls) MDF_PB3(RT, G, ST, 1);              \
        else MDF_PB3(RT, G, ST, MAX_LEAD_TYPES);            \
    } while (0)
#define MDF_PB(RT, G)                                  \
    do {                                               \
        if (b->pair_strided) MDF_PB2(RT, G, true);     \
        else MDF_PB2(RT, G, false);                    \
    } while (0)
        if (b->f32) {
            if (b->pair_g == 8) MDF_PB(float, 8); else if (b->pair_g == 32) MDF_PB(float, 32); else MDF_PB(float, 16);
        } else {
            if (b->pair_g == 8) MDF_PB(double, 8); else if (b->pair_g == 32) MDF_PB(double, 32); else MDF_PB(double, 16);
        }
#undef MDF_PB
#undef MDF_PB2
#undef MDF_PB3
    }
    if (b->fork) (void)hipStreamWaitEvent(st, b->ev_join, 0);
    else batch_kspace(b, st);
    k_epilogue_b<<<grid(BK_EPI), EPI_BLOCK, 0, st>>>(R, b->maps[BK_EPI].p);
    k_colsum_b<<<grid(BK_COLSUM), 256, 0, st>>>(R, b->maps[BK_COLSUM].p);
    k_final_b<<<n, 256, 0, st>>>(R);
    k_flags_b<<<(n + 63) / 64, 64, 0, st>>>(R, n, b->d_flags.p);
}

static void batch_free(mdf_batch *b) {
    batch_drop_graphs(b);
    b->d_rep.release();
    for (int k = 0; k < BK_N; k++) b->maps[k].release();
    for (int k = 0; k < mdf_batch::NB_RING; k++) {
        b->d_in_ring[k].release();
        if (b->h_in_ring[k]) (void)hipHostFree(b->h_in_ring[k]);
        b->h_in_ring[k] = nullptr;
        if (b->done_ev[k]) (void)hipEventDestroy(b->done_ev[k]);
        b->done_ev[k] = nullptr;
    }
    b->d_flags.release();
    if (b->in_event) (void)hipEventDestroy(b->in_event);
    if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
    if (b->ev_join) (void)hipEventDestroy(b->ev_join);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    if (b->stream2) (void)hipStreamDestroy(b->stream2);
}

extern "C" {

int mdf_batch_create(mdf_system *const *systems, int32_t n, mdf_batch **out) {
    if (!systems || !out || n <= 0) return fail(MDF_ERR_ARG, "null argument or empty batch");
    *out = nullptr;
    for (int i = 0; i < n; i++)
        if (!systems[i]) return fail(MDF_ERR_ARG, "system %d is null", i);
    const mdf_system *s0 = systems[0];
    for (int i = 1; i < n; i++) {
        const mdf_system *s = systems[i];
        if (s->device != s0->device || s->par.precision != s0->par.precision || s->par.virial != s0->par.virial ||
            s->par.erfc_mode != s0->par.erfc_mode)
            return fail(MDF_ERR_ARG, "system %d differs from system 0 in device, precision, erfc form or virial", i);
        for (int j = 0; j < i; j++)
            if (systems[j] == s) return fail(MDF_ERR_ARG, "system %d appears twice", i);
    }
    for (int i = 0; i < n; i++) {
        if (systems[i]->mts_on) return fail(MDF_ERR_ARG, "system %d has multiple time steps on (not in batches)", i);
        if (systems[i]->obs_on) return fail(MDF_ERR_ARG, "system %d accumulates observables (not in batches)", i);
        if (systems[i]->ang_on) return fail(MDF_ERR_ARG, "system %d accumulates the angle histogram (not in batches)", i);
        if (systems[i]->chem_on) return fail(MDF_ERR_ARG, "system %d makes Widom insertions (not in batches)", i);
    }
    mdf_batch *b = new mdf_batch;
    b->sys.assign(systems, systems + n);
    b->device = s0->device;
    b->f32 = s0->par.precision == MDF_FP32_ACC64;
    b->vir = s0->par.virial != 0;
    b->acc = s0->par.erfc_mode == MDF_ERFC_ACCURATE;
    b->pair_g = s0->pair_g;
    b->pair_strided = s0->pair_strided;
    b->one_cls = true;
    for (int i = 0; i < n; i++) b->one_cls = b->one_cls && systems[i]->slot_type.size() <= 1;
    {   // the matrix-core k-sum when every system with a k-sum runs it alone
        bool m32 = true, m64 = true;
        for (int i = 0; i < n; i++) {
            const mdf_system *s = systems[i];
            if (s->ak.empty() || s->top.n_move == 0) continue;
            m32 = m32 && s->sf_mma && s->bp_mma;
            m64 = m64 && s->mma64;
        }
        b->rmode = b->f32 ? (m32 ? BR_MMA32 : BR_VALU) : (m64 ? BR_MMA64 : BR_VALU);
        if (const char *e = getenv("MDF_BATCH_VALU")) if (e[0] == '1') b->rmode = BR_VALU;
    }
    int rc = MDF_OK;
    hipError_t e = hipSetDevice(b->device);
    if (const char *fk = getenv("MDF_FORK_KSPACE_BATCH")) b->fork = fk[0] == '1';
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (b->fork) {
        if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream2, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev_join, hipEventDisableTiming);
    }
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->in_event, hipEventDisableTiming);
    for (int i = 0; i < n && e == hipSuccess; i++) e = hipStreamSynchronize(systems[i]->stream);
    if (e == hipSuccess) e = b->d_rep.alloc(n);
    for (int k = 0; k < mdf_batch::NB_RING && e == hipSuccess; k++) {
        e = b->d_in_ring[k].alloc(6 * (size_t)n);
        if (e == hipSuccess) e = hipHostMalloc((void **)&b->h_in_ring[k], sizeof(double *) * 6 * (size_t)n, 0);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&b->done_ev[k], hipEventDisableTiming);
    }
    if (e == hipSuccess) e = b->d_flags.alloc(n);
    if (e != hipSuccess) rc = fail(MDF_ERR_HIP, "batch set-up: %s", hipGetErrorString(e));
    if (rc == MDF_OK) rc = batch_upload(b, true);
    if (rc != MDF_OK) {
        batch_free(b);
        delete b;
        return rc;
    }
    b->h_flags.resize(n);
    b->h_in.assign(6 * (size_t)n, nullptr);
    *out = b;
    return MDF_OK;
}

int mdf_batch_compute_device(mdf_batch *b, const double *const *x, const double *const *y, const double *const *z,
                             const double *const *cx, const double *const *cy, const double *const *cz,
                             void *stream) {
    if (!b || !x || !y || !z) return fail(MDF_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(b->device));
    const int n = (int)b->sys.size();
    // this step's slot: free once the step that used it NB_RING steps back has finished
    const int sl = b->slot;
    if (b->slot_used[sl]) HIPCHK(hipEventSynchronize(b->done_ev[sl]));
    for (int i = 0; i < n; i++) {
        if (!x[i] || !y[i] || !z[i]) return fail(MDF_ERR_ARG, "null positions for system %d", i);
        if (b->sys[i]->top.n_rig > 0 && (!cx || !cy || !cz || !cx[i] || !cy[i] || !cz[i]))
            return fail(MDF_ERR_ARG, "COM arrays required for system %d", i);
        const mdf_system *si = b->sys[i];
        if (si->h_mirror && *(volatile int *)si->h_mirror)
            return fail(MDF_ERR_LIST_OVERFLOW, "system %d: an earlier step's pair list overflowed its rows; the step "
                        "was not enqueued: call mdf_batch_read_outputs", i);
        const double *p[6] = {x[i], y[i], z[i], cx ? cx[i] : nullptr, cy ? cy[i] : nullptr, cz ? cz[i] : nullptr};
        for (int c = 0; c < 6; c++) b->h_in[6 * i + c] = p[c];
    }
    memcpy(b->h_in_ring[sl], b->h_in.data(), sizeof(double *) * 6 * n);
    HIPCHK(hipMemcpyAsync(b->d_in_ring[sl].p, b->h_in_ring[sl], sizeof(double *) * 6 * n, hipMemcpyHostToDevice,
                          b->stream));
    HIPCHK(hipEventRecord(b->in_event, (hipStream_t)stream));
    HIPCHK(hipStreamWaitEvent(b->stream, b->in_event, 0));
    if (!b->graph_ring[sl]) {
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeRelaxed));
        batch_launch(b, b->d_in_ring[sl].p);
        hipError_t ec = hipStreamEndCapture(b->stream, &g);
        if (ec != hipSuccess) return fail(MDF_ERR_HIP, "batch graph capture: %s", hipGetErrorString(ec));
        ec = hipGraphInstantiate(&b->graph_ring[sl], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ec != hipSuccess) return fail(MDF_ERR_HIP, "batch graph instantiate: %s", hipGetErrorString(ec));
    }
    HIPCHK(hipGraphLaunch(b->graph_ring[sl], b->stream));
    HIPCHK(hipEventRecord(b->done_ev[sl], b->stream));
    HIPCHK(hipGetLastError());
    b->slot_used[sl] = true;
    b->last_slot = sl;
    b->slot = (sl + 1) % mdf_batch::NB_RING;
    b->pending = true;
    return MDF_OK;
}

int mdf_batch_read_outputs(mdf_batch *b, mdf_outputs *outs) {
    if (!b || !outs) return fail(MDF_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(b->device));
    HIPCHK(hipStreamSynchronize(b->stream));
    b->pending = false;
    const int n = (int)b->sys.size();
    HIPCHK(hipMemcpy(b->h_flags.data(), b->d_flags.p, sizeof(DevFlags) * n, hipMemcpyDeviceToHost));
    bool regrown = false;
    int late = -1;   // a system whose overflow came from an earlier step (-1: none)
    for (int i = 0; i < n; i++) {
        mdf_system *s = b->sys[i];
        s->last_flags = b->h_flags[i];
        s->last_rebuilt = s->last_flags.rebuild;
        const DevFlags f = s->last_flags;
        if (f.overflow_mm || f.overflow_mw) {
            // this system's list outgrew its rows: grow them and redo its step on its own
            // (the batch kernels already gathered its positions)
            // with a fixed capacity the flags stay set: the batch keeps reporting the overflow
            // (the reference stops, nnlist.f90:136-139) since its graph cannot rebuild one system
            if (s->fixed_cap)
                return fail(MDF_ERR_LIST_OVERFLOW, "system %d: nnlist too big (need %d / %d partners per lead)%s", i,
                            f.overflow_mm, f.overflow_mw,
                            f.steps != f.overflow_step ? "; later steps ran on the truncated rows" : "");
            int rc = clear_overflow(s);
            if (rc) return rc;
            if (f.steps != f.overflow_step && late < 0) late = i;
            rc = alloc_lists(s, std::max(s->S.cap_mm, (int)(f.overflow_mm * 1.25) + 16),
                             std::max(s->S.cap_mw, (int)(f.overflow_mw * 1.25) + 16));
            if (rc) return rc;
            rc = run_step(s, 1);
            if (rc) return rc;
            int redo = 0;
            rc = finish_step(s, &redo);
            if (rc) return rc;
            if (redo) return fail(MDF_ERR_LIST_OVERFLOW, "system %d: pair list kept overflowing", i);
            regrown = true;
        }
    }
    if (regrown) {
        int rc = batch_upload(b, false);   // new list rows; block counts are unchanged
        if (rc) return rc;
    }
    for (int i = 0; i < n; i++) {
        int rc = copy_outputs(b->sys[i], &outs[i]);
        if (rc) return rc;
    }
    if (late >= 0)
        return fail(MDF_ERR_LIST_OVERFLOW, "system %d: a pair-list row overflowed before the last batch step; the "
                    "steps since ran on truncated rows (lists grown, the last step redone: the outputs are valid)", late);
    return MDF_OK;
}

int mdf_batch_autotune(mdf_batch *b, int32_t reps, int32_t *chosen) {
    if (!b) return fail(MDF_ERR_ARG, "null batch");
    if (reps <= 0) reps = 5;
    const int n = (int)b->sys.size();
    for (int i = 0; i < n; i++)
        if (!b->h_in[6 * i])
            return fail(MDF_ERR_ARG, "tuning times steps on the last inputs: call mdf_batch_compute_device first");
    HIPCHK(hipSetDevice(b->device));
    HIPCHK(hipStreamSynchronize(b->stream));
    b->pending = false;
    // reciprocal work share of each system: k-vectors x moving sites
    std::vector<double> w(n);
    double wt = 0.0;
    for (int i = 0; i < n; i++) {
        HIPCHK(hipStreamSynchronize(b->sys[i]->stream));
        w[i] = (double)b->sys[i]->ak.size() * b->sys[i]->top.n_move;
        wt += w[i];
    }
    wt = std::max(wt, 1.0);
    // new block maps and per-system tables; the captured step holds the old maps
    auto reshape = [&](int sf_total, int bp_total) -> int {
        batch_drop_graphs(b);
        for (int i = 0; i < n; i++) {
            mdf_system *s = b->sys[i];
            s->sf_target_blocks = std::max(1, (int)llround(sf_total * w[i] / wt));
            s->bp_target_blocks = std::max(1, (int)llround(bp_total * w[i] / wt));
            int rc = apply_geometry(s);
            if (rc) return rc;
        }
        return batch_upload(b, true);
    };
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    int rc = MDF_OK;
    // mean time of one whole batch step (direct launches), after one untimed step
    auto timed = [&](float *ms) -> int {
        batch_launch(b, b->d_in_ring[b->last_slot].p);
        HIPCHK(hipEventRecord(e0, b->stream));
        for (int r = 0; r < reps; r++) batch_launch(b, b->d_in_ring[b->last_slot].p);
        HIPCHK(hipEventRecord(e1, b->stream));
        HIPCHK(hipEventSynchronize(e1));
        HIPCHK(hipGetLastError());
        float t = 0.0f;
        HIPCHK(hipEventElapsedTime(&t, e0, e1));
        *ms = t / reps;
        return MDF_OK;
    };
    float best = 1e30f, t = 0.0f;
    int sf_best = 8192, bp_best = 2048, g_best = b->pair_g;
    bool st_best = b->pair_strided;
    if ((rc = reshape(sf_best, bp_best)) != MDF_OK) goto done;
    for (int g : {8, 16, 32}) {
        b->pair_g = g;
        if ((rc = batch_upload(b, true)) != MDF_OK) goto done;
        for (int ro = 0; ro < 2; ro++) {
            b->pair_strided = ro != 0;
            if ((rc = timed(&t)) != MDF_OK) goto done;
            if (t < best) { best = t; g_best = g; st_best = ro != 0; }
        }
    }
    b->pair_g = g_best;
    b->pair_strided = st_best;
    best = 1e30f;
    for (int tb : {4096, 8192, 16384, 32768}) {
        if ((rc = reshape(tb, bp_best)) != MDF_OK) goto done;
        if ((rc = timed(&t)) != MDF_OK) goto done;
        if (t < best) { best = t; sf_best = tb; }
    }
    best = 1e30f;
    for (int tb : {1024, 2048, 4096, 8192}) {
        if ((rc = reshape(sf_best, tb)) != MDF_OK) goto done;
        if ((rc = timed(&t)) != MDF_OK) goto done;
        if (t < best) { best = t; bp_best = tb; }
    }
    rc = reshape(sf_best, bp_best);
    if (rc == MDF_OK && chosen) {
        chosen[0] = b->pair_g;
        chosen[1] = 0;
        chosen[2] = sf_best;
        chosen[3] = bp_best;
    }
done:
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

void mdf_batch_destroy(mdf_batch *b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    batch_free(b);
    delete b;
}

}  // extern "C"
