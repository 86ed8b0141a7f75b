// Internal definitions shared by the force engine's HIP translation units.
// Device data layout (all in HBM, allocated once per handle):
//   positions x,y,z [n_total] SoA FP64 (frozen part uploaded once), charges, site types;
//   leads [n_lead] = moving sites that carry a molecule (nsi > 0), ascending site index;
//   neighbour lists in ELL form, one row of `cap` int32 per lead (full: both directions
//   for moving pairs, lead -> frozen sites for wall pairs), counts per lead;
//   per-lead partial sums in column-major [NCOL][n_lead] for deterministic reduction;
//   k-set as rows (kl, km, kn0..kn1) plus flat per-k arrays (Ak, S_wall, S, T = Ak conj S).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mdf {

constexpr int NCLS = 6;          // energy classes of a partner type (moving lead types + 2 crystal)
constexpr int MAX_TYPES = 14;
constexpr int MAX_LEAD_TYPES = 4;
// per-lead partial columns: vdw energy per class, coulomb energy per class, virial(9),
// wall force x/y/z for sgn=-1 and sgn=+1 (6)
constexpr int COL_EV = 0;
constexpr int COL_EC = NCLS;
constexpr int COL_VIR = 2 * NCLS;
constexpr int COL_FWS = 2 * NCLS + 9;
constexpr int NCOL = 2 * NCLS + 9 + 6;
constexpr int KC = 16;           // k-vectors per structure-factor chunk held in registers
constexpr int SF_SITES = 4;      // sites per thread in the structure-factor kernel
constexpr int SF_BLOCK = 256;
constexpr int SFR_K = 64;        // k-vectors per structure-factor reduction block (one wave wide)
constexpr int SFR_THREADS = 1024; // threads of a reduction block: 16 slices over the site blocks
constexpr int BP_ROWGROUPS_MAX = 64;
// dynamic LDS of k_sf_mma for a block of `sites` sites and n_nt kn tiles: cos/sin steps
// (float4), site phases (3 doubles) and charges per site, then one 32-site stage's split E
// tables (kn stride 16 n_nt + 4 words) and P tables (row stride 68 words), hi and lo each
__host__ __device__ inline size_t sfm_lds_bytes(int sites, int n_nt) {
    return (size_t)sites * 44 + (size_t)2 * 32 * (16 * n_nt + 4) * 4 + (size_t)2 * 32 * 68 * 4;
}
constexpr int BP_GROUP_K = 1024;     // k-vectors per back-projection row group (its LDS stage)
// FP64 reciprocal sum on the matrix cores: sites per LDS stage of the structure factor, waves
// of its block, column tiles per wave, and the powers e, e^2, e^4, e^8 of a site's base
// phases kept for the tables (a table entry n = h + 8 i starts from e^h, h < 8, and steps by
// e^8; q e^{i pi kl u} and e^{i pi km v} by e^4)
constexpr int SF64_SUB = 32;
constexpr int SF64_WAVES = 16;
constexpr int SF64_SLOTS = 3;
constexpr int SF64_POWS = 4;
// dynamic LDS of k_sf64: base-phase powers, q e^{i pi kl u}, e^{i pi km v} for km >= 0
// (complex) and the cos/sin columns (16 n_nt + 2 per site) of one stage
// cos/sin columns per site in LDS: at least the 16 n_nt columns the products read, and an odd
// multiple of 16 doubles, so that the 16-column windows of two consecutive sites of a K-step
// fall on disjoint halves of the 64 four-byte banks
__host__ __device__ constexpr int sf64_ep(int n_nt) { return 16 * n_nt + ((n_nt & 1) ? 0 : 16); }
__host__ __device__ inline size_t sf64_lds_bytes(int klmax, int kmmax, int n_nt) {
    return (size_t)SF64_SUB * (16 * (3 * SF64_POWS + (klmax + 1) + (kmmax + 1)) + 8 * sf64_ep(n_nt));
}
constexpr size_t SF64_LDS_MAX = 159 * 1024;   // dynamic LDS of a k_sf64 block, at most (gfx950: 160 KiB per unit)

// k-set rows (contiguous kn) and structure-factor chunks (<= KC consecutive kn of a row)
struct KChunk { int kl, km, kn0, len, koff; };
struct KRow { int kl, km, kn0, kn1, koff; };

// wall long-range tail tables per site type (src/forcem.f90:157-231)
struct WallLR {
    double flr_l[MAX_TYPES + 1], flr_w[MAX_TYPES + 1], flr_lja[MAX_TYPES + 1], flr_ljc[MAX_TYPES + 1];
    double flr_bua[MAX_TYPES + 1], flr_buc[MAX_TYPES + 1], flr_bur[MAX_TYPES + 1];
    double ulr_l[MAX_TYPES + 1], ulr_w[MAX_TYPES + 1];
    double rcut;
    int on;
};

constexpr int EPI_BLOCK = 192;  // epilogue block: 64 rigid molecules x 3 sites
constexpr int EP_N = 6;  // epilogue partials: ulr, fwsz(-1), fwsz(+1), sum fx, sum fy, sum fz
constexpr int RE_N = 8;  // k partials: A|S|^2, A|S_moving|^2, virial xx yy zz xy xz yz

// scalar outputs of one step, laid out for one device->host copy
struct Scalars {
    double e_recip, e_recip_ll, e_recip_wl, e_recip_ww, e_self, e_slab, e_wall_lr, e_conf, e_avex, press;
    double f_wall[9];
    double virial[9];
    int nonfinite;
    int pad;
};

struct FinalArgs {
    int n_colblk;                 // blocks per column in colpart
    int n_slots;                  // lead type slots in use
    int slot_type[MAX_LEAD_TYPES];
    int cls_type[NCLS];           // type of each energy class (0: unused)
    int n_kblk;                   // blocks of k partials
    int n_epblk;                  // blocks of epilogue partials
    double e_self, e_recip_ww;
    double crf0;                  // -2 pi / (lx ly lz) (src/forcem.f90:925)
    int ew3dc;
    double lx, ly, na_scale, bar; // pressure formula pieces (src/forcem.f90:1081)
};

// flags written by device kernels, read back by the host after a step
// The overflow fields are sticky: a build that outgrows a row sets them, and only the host
// clears them after reading them, so an overflow is never lost behind later steps that did
// not rebuild (they ran on the same truncated rows).
struct DevFlags {
    int rebuild;                         // 1: the list is rebuilt this step
    int overflow_mm, overflow_mw;        // max count seen when a list row overflowed
    int overflow_step;                   // value of `steps` at the step whose build overflowed
    int steps;                           // steps since creation (counted by the rebuild decision)
    int ncz;                             // cells in z for this build
    double z0, csz;                      // z origin and cell size for this build
    int n_pairs_mm, n_pairs_mw;          // listed pairs (i<j count and wall count)
    int n_builds;                        // builds since creation (not reset per step)
};

struct Grid {
    int ncx, ncy;        // periodic x,y cells
    int ncz_max;         // capacity of z cells
    double csx, csy;     // cell sizes
    int zper;            // z periodic in the list
    int ncz_fixed;       // with periodic z: fixed cell count
    double csz_fixed;
};

struct DevSys {
    // sizes
    int n_total, n_move, n_rig, n_rigid_sites, n_liquid_sites, n_lead, n_wall, n_types, n_ltype, nijt;
    // topology
    const int *itype;      // [n_total]
    const double *q;       // [n_total]
    const int *nsi;        // [n_types+1] indexed by 1-based type
    const int *ijtype;     // [(n_types+1)^2] 1-based types, 1-based pair numbers
    const int *vkind;      // [nijt+1] 1-based pair number
    const double *eps4, *sig6, *aa, *rhoi, *cc;  // [nijt+1]
    const int *cls;        // [n_types+1] energy class of a type
    const int *lead;       // [n_lead] site index
    const int *con_ptr;    // [n_move+1] CSR of constraint partners (surface sites)
    const int *con_idx;
    // angle bending (force_intra): per moving site the angles whose force moves it
    int n_angles;
    const int *ang_ptr;    // [n_move+1] CSR over moving sites
    const int *ang_inc;    // angle * 2 + role (0: vertex site j, 1: end site k)
    const int *ang_ijk;    // [3 * n_angles] sites i, j, k
    const double *ang_kt, *ang_th;  // [n_angles] force constant, equilibrium angle (rad)
    // state
    double *x, *y, *z;     // [n_total]
    const double *cx, *cy, *cz;  // [n_rig]
    double *ocx, *ocy, *ocz;     // [n_rig] COM snapshot at the last build
    double *ox, *oy, *oz;        // [n_move] site snapshot (non-rigid sites)
    // box and cut-offs
    double lx, ly, lz, lxi, lyi, lzi, lwh;
    int zper_list, zper_force;
    double rcutsq, rskinsq, skin2;
    double alfa, spii, facq;
    double a1, a2, a3, a4, a5, p;
    int erfc_mode;
    // lists
    int *nbr_mm, *cnt_mm; int cap_mm;
    int *cnt_lo;                 // [n_lead] partners j > i per row (the reference's i<j pairs)
    int *nbr_mw, *cnt_mw; int cap_mw;
    // cells
    Grid g;
    int *cell_cnt, *cell_start, *cell_items;   // moving leads
    int pair_cell_order;                       // the pair kernel takes leads in cell_items order
    double4 *cell_xyz;                         // the cell-ordered leads' x, y, z and site index
    int *cell_cursor;
    int n_cells_cap;
    const int *wcell_start, *wcell_items;      // frozen sites (built once)
    const double4 *wcell_xyz;                  // and their x, y, z and site index
    int wncz; double wz0, wcsz;
    // outputs
    double *fx, *fy, *fz;        // [n_move] site forces
    double *part;                // [NCOL][n_lead]
    DevFlags *flags;
    int *flag_mirror;            // host-mapped: [0] = 1 once a list overflow is pending (or null)
    // packed inputs of the pair kernel: per lead site i the molecule record (PREC doubles:
    // x, y, z, q of the lead, then offset x, y, z from the lead and q of sites 1 and 2),
    // rewritten by the skin test every step; per moving site its lead-type slot | nsi << 4
    // (static); per frozen site (x, y, z, q) and its crystal class 0/1 (static); the vdW
    // parameters per (lead slot, partner class): kind, 4 eps, sigma^6, A, 1/rho, C
    double *prec;
    const int *pmeta;
    const double4 *wrec;
    const int *wcls;
    const double *ptab;          // [MAX_LEAD_TYPES][NCLS][PTAB_W]
};
constexpr int PREC = 12;         // doubles per molecule record
constexpr int PTAB_W = 8;        // doubles per vdW table entry


// ---- helpers ---------------------------------------------------------------------
// Fortran anint: round half away from zero
__device__ __forceinline__ double anint_d(double v) { return round(v); }

// squared minimum-image distance of the reference's list predicate, evaluated with the
// reference's operation order and no FMA contraction (src/nnlist.f90:124-130)
__device__ __forceinline__ double list_d2(const DevSys &S, int i, int j) {
    double dx = __dsub_rn(S.x[i], S.x[j]);
    dx = __dsub_rn(dx, __dmul_rn(S.lx, anint_d(__dmul_rn(dx, S.lxi))));
    double dy = __dsub_rn(S.y[i], S.y[j]);
    dy = __dsub_rn(dy, __dmul_rn(S.ly, anint_d(__dmul_rn(dy, S.lyi))));
    double dz = __dsub_rn(S.z[i], S.z[j]);
    if (S.zper_list) dz = __dsub_rn(dz, __dmul_rn(S.lz, anint_d(__dmul_rn(dz, S.lzi))));
    return __dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)), __dmul_rn(dz, dz));
}

// ---- reciprocal-force surrogate and multiple time steps --------------------------------
// Per moving site: 48 invariant inputs and 16 vector features from the sites within
// SG_RC; a 48-64-64-16 SiLU network gives the coefficients of the vector features.
constexpr int SG_D = 48, SG_H = 64, SG_M = 16;
constexpr int SG_MV = 8;   // vector features computed (the rest of the 16 outputs' partners are zero)
constexpr float SG_RC = 6.0f;
constexpr int SG_OFF_W1 = 0, SG_OFF_B1 = 3072, SG_OFF_W2 = 3136, SG_OFF_B2 = 7232, SG_OFF_W3 = 7296,
              SG_OFF_B3 = 8320, SG_NP = 8336;
// bfloat16 copies of the weights: W1 [64][48], W2 [64][64], W3 [16][64], and the transposes
// W2T [64][64], W3T [64][16] that the backward pass multiplies by
constexpr int SG_WB_W1 = 0, SG_WB_W2 = 3072, SG_WB_W3 = 7168, SG_WB_W2T = 8192, SG_WB_W3T = 12288, SG_WB_N = 13312;
constexpr int SG_NORM_N = 2 * SG_D + SG_M;   // mu[48], sigma[48], su[16]
constexpr int SG_TRAIN_BLOCKS = 128;         // most blocks of one training launch
constexpr int SG_GP = SG_NP + 1;             // gradient partial row: parameters, then the loss

// device-resident state of the scheme, one per handle
struct MtsState {
    int state;        // 0: collect (full k-sum every step, training), 1: multiple time steps
    int exact;        // this step runs the full k-sum
    int phase;        // steps since the last exact step
    int frames;       // exact frames in the current collect run
    int train_now;    // this exact step trains
    int have_delta;   // a correction from an exact step with the current weights exists
    int norm_ready;   // input normalisation fixed (first frame, or given weights)
    int prior;        // set the physics prior (c_0 = su_0) when the normalisation is fixed
    int adam_t;       // Adam steps taken
    int have_prev;    // a previous full-sum frame is stored (training pairs)
    int have_slope;   // the last two corrections came from the same weights: extrapolate
    int adam_base;    // Adam steps taken before this step's training iterations
    long long step;                      // steps since the scheme was switched on
    long long delta_step, delta_prev_step;  // steps of the last two corrections
    long long n_exact, n_surrogate, n_fallback, n_judged;
    double err_last, err_max, err_sum;   // judged relative force RMS errors
    double t2;        // mean |F_k - F_k(previous full sum)|^2 of the last full sum (loss scale)
    double loss_last;
    double de_acc;    // reciprocal energy change since the last exact step (path integral)
};

struct MtsArgs {
    MtsState *st;
    int interval, train_frames, model, n_sblk, n_tblk, train_iters;   // model 1: network, 0: held force
    int train_stride, pad2;                    // an Adam step trains on one 16-site tile in train_stride
    double tol;
    float lr;
    float *feat, *vec, *sg;                    // [n][48], [n][16][3], [3][n]
    float *sg_prev;                            // [3][n] the previous full-sum frame, new weights
    float *feat_prev, *vec_prev;               // the previous full-sum frame's
    float4 *pos4;                              // [n_total] x, y, z, q as float
    const int *msize;                          // [n_total] sites of the molecule a lead carries (0: not a lead)
    double *delta, *frec, *fmts, *rprev, *fprev; // [3][n]
    double *frec_prev;                         // [3][n]
    double *delta_prev;                        // [3][n] the correction before the last
    float *params, *adam_m, *adam_v, *gsum;    // [SG_NP] (gsum: [SG_GP])
    unsigned short *wb;                        // [SG_WB_N]
    float *norm;                               // [SG_NORM_N]
    float *gpart;                              // [SG_TRAIN_BLOCKS][SG_GP]
    double *jpart;                             // [n_sblk][4] judge partials
    double *epart;                             // [n_sblk] energy path partials
};

// ---- observables of the reference's hot loops (measurement only) -----------------------
// bins of force_intra's angle histogram: ngang(:, 0:langmax), langmax = 201 (src/measure.f90:96-99)
constexpr int MDF_NGANG_BINS = 202;

struct ObsArgs {
    unsigned long long *ngc;   // [nbins][nijt] site-pair distance histogram of the pair loops
    double *efi;               // [n_move] this step's z-force of each mobile ion of types 3, 4
    double *efipos;            // [nbins][2] accumulated z-force profile of those ions
    double rdeli, lwh;
    int nbins;
    int count_mm;              // 1: moving-moving site pairs count too (3-D pair loops)
};

// ---- Widom test-particle insertions (measurement only) ---------------------------------
constexpr int CHEM_ROT_MAX = 20;        // orientations per insertion in the far bins
constexpr int CHEM_TAB_MAX = 1024;      // e^{i k r} table entries per site (LDS)
struct ChemArgs {
    // this call's draws, in insertion order
    const double *ins;          // [n_ins][3] oxygen positions
    const int *ins_il;          // [n_ins] 1-based bin
    const int *ins_rot;         // [n_ins + 1] prefix of the orientation counts
    const int *rot_ins;         // [n_rot] insertion of each orientation
    const double *rot;          // [n_rot][6] the two hydrogens (x, y, z each)
    // oxygen scan
    int *overlap;               // [n_ins]
    int *n_nbr;                 // [n_ins] partners within rcut (may exceed cap)
    double *u_ox;               // [n_ins]
    int *nbr;                   // [n_ins][cap]
    double *nbr_d;              // [n_ins][cap][3] oxygen - partner, minimum image in x, y
    double *boltz;              // [n_rot] exp(-beta u)
    double *bsum;               // [n_ins] sum over the insertion's orientations
    int *status;                // [4]: first overlap (n_ins: none), partner overflow, skipped, max partners
    double *pave;               // [n_bins]
    unsigned long long *inave;  // [n_bins]
    const double *dipz;         // the step's sum q z over moving sites
    const int *kvec;            // [nk][3]
    const double *ak;           // [nk]
    const double2 *sk;          // [nk] this step's S(k)
    int n_ins, n_rot, cap, n_bins, nk, klmax, kmmax, knmax;
    double rcut, beta, rinsmin, rwinsmin, qsw[3];
    double crf_scale;           // crf = crf_scale * dipz (0: no slab term)
    double lwh, pipix, pipiy, pipiz;
    double ulr_w, ulr_l, lja, ljc, bua, buc, bur;  // wall tail tables of type 1
};

// list build: blocks of BUILD_THREADS over (cell, split) work items, candidates staged
// BUILD_CH at a time, leads in rounds of BUILD_LEADS; BUILD_ITEMS_PER_CELL items per cell
constexpr int BUILD_THREADS = 256, BUILD_WAVES = BUILD_THREADS / 64, BUILD_CH = 512, BUILD_LEADS = 16;
constexpr int BUILD_ITEMS_PER_CELL = 4;

// ---- batches of independent systems ----------------------------------------------------
// One entry per system: everything its launches take.  A batched kernel covers the blocks of
// every system in one launch; a block map entry (system, local block) tells each block whose
// work it does.  Kernels with one block per system index the table by blockIdx.x.
struct RepArgs {
    DevSys S;
    int n_rig_sites_pad;                 // unused; keeps the table entries 8-byte aligned
    // list stage
    double *dip_part, *dipz;
    int skin_blocks, force_rebuild;
    int *lead_cell, *items_tmp;
    double4 *xyz_tmp;
    int lead_blocks, order_blocks, build_blocks, ncells_xy, gather_blocks;
    // reciprocal sum
    const KRow *rows;
    int nrows, cpb, n_sb, n_cg, nk, n_kblk, n_rg, n_bpsb;
    double2 *sf_partial;
    const double2 *s_wall;
    const double *ak;
    const int *kvec;
    double2 *s_out;
    double4 *t_out;
    double *kpart;
    const int *bp_row_start;
    double *fpart;
    double cux, cuy, cuz, pipix, pipiy, pipiz, lamb;
    // epilogue, column sums, final sums
    // the reciprocal sum on the matrix cores: FP32 form (k_sf_mma, k_sf_reduce_rows, k_bp_mma;
    // sf_partial holds FP32 partial rows) and FP64 form (k_sf64, k_sf64_reduce, k_bp64)
    int knmax, n_nt, sfm_sites, sfm_nsb, sfm_nrg;
    const int2 *rowkk, *tile_ks;
    unsigned *tpack;
    int *tscale;
    int klmax, kmmax, s64_nsb, s64_pcols, s64_sites, b64_nsb;
    const int2 *s64_rowkk, *b64_tile, *b64_rg;
    const int4 *s64_waves;
    const int *s64_pncol, *s64_pcoff, *s64_sorted_of;
    double2 *s64_part;
    const int4 *b64_rinfo;
    double *b64_tpk;
    const WallLR *wl;
    int ew3dc, n_groups, n_epblk, colper, n_colblk, pad2;
    double crf0;
    double *eppart, *fcom, *tq, *colpart, *e_vdw, *e_coul;
    const int *lead_slot;
    Scalars *scal;
    FinalArgs A;
};

}  // namespace mdf
