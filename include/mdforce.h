/*
 * mdforce.h - C ABI of the MI355X force engine for md_sim's force evaluation.
 *
 * Replaces the reference's force module entry points, which take no arguments and
 * communicate through module variables:
 *   force_init  src/forcem.f90:54-81   (constants, k-set, static wall S(k), wall tail
 *                                       tables, first neighbour list)   -> mdf_create
 *   force_get   src/forcem.f90:93-129  (list skin test + rebuild, real-space loops,
 *                                       reciprocal sum, slab correction, angle terms,
 *                                       COM force/torque/energy reduction) -> mdf_compute
 *   force_zero  src/forcem.f90:84-90   (running-average reset)         -> mdf_zero
 *   stop '...' at the reference's check sites                          -> status codes
 * and the neighbour-list module nnlist (src/nnlist.f90:14-164), which the library owns.
 *
 * Conventions
 *  - Site order is the reference's (src/sysdef.f90:381-419): rigid 3-site molecules
 *    first (sites 3m, 3m+1, 3m+2, lead = 3m), then other moving sites, then frozen
 *    crystal ("wall") sites.  Indices in this ABI are 0-based; site types are the
 *    reference's 1-based type numbers.
 *  - Positions are unwrapped, SoA, FP64, Angstrom; energies kJ/mol; forces kJ/mol/A.
 *  - Frozen-site positions are taken once, at mdf_create (the reference computes the
 *    static wall structure factor once, src/forcem.f90:766-767).
 *  - One handle = one system on one GPU.  Calls on a handle are synchronous (outputs
 *    valid on return) except mdf_compute_device, and are not re-entrant per handle.
 */
#ifndef MDFORCE_H
#define MDFORCE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes; each stop site of the reference maps to one */
enum {
    MDF_OK = 0,
    MDF_ERR_ARG = 1,            /* invalid argument or unsupported geometry            */
    MDF_ERR_HIP = 2,            /* HIP runtime error                                   */
    MDF_ERR_LIST_OVERFLOW = 3,  /* 'nnlist_make(): nnlist too big!'  nnlist.f90:136-139,157-160 */
    MDF_ERR_VDW_TYPE = 4,       /* 'unknown interaction type'  forcem.f90:208-211,326-328,455-456 */
    MDF_ERR_KSET = 5,           /* 'too large kcount'  forcem.f90:751-754,788-791      */
    MDF_ERR_NONFINITE = 6,      /* non-finite energy (the reference only prints, forcem.f90:544-547) */
    MDF_ERR_NO_DEVICE = 7       /* no HIP device visible                               */
};

enum { MDF_FP64 = 0, MDF_FP32_ACC64 = 1 };               /* arithmetic of pair and k-sum kernels */
enum { MDF_CONST_REFERENCE = 0, MDF_CONST_EXACT = 1 };   /* widened single literals / exact     */
enum { MDF_ERFC_AS7126 = 0, MDF_ERFC_ACCURATE = 1 };     /* A&S 7.1.26 as forcem.f90:11-13,335-337 */
enum { MDF_ALPHA_REFERENCE = 0, MDF_ALPHA_GIVEN = 1 };   /* alfa = 3.09/rcut (forcem.f90:59) or given */
enum { MDF_Z_OPEN = 0,            /* forcem.f90: min image in x,y only                       */
       MDF_Z_PERIODIC = 1,        /* z min image in pair loops and in the list               */
       MDF_Z_PERIODIC_OPEN_LIST = 2 /* forcem_back.f90 pair loops over an open-z list (nnlist.f90:129,150) */ };

typedef struct mdf_topology {
    int32_t n_total;         /* ntota                                                 */
    int32_t n_move;          /* nmove: moving sites                                   */
    int32_t n_rigid_sites;   /* nriga = 3 * n_rig                                     */
    int32_t n_liquid_sites;  /* nliqa (rigid + mobile atoms; surface sites follow)   */
    int32_t n_rig;           /* rigid 3-site molecules                                */
    int32_t n_types;         /* nit                                                   */
    int32_t n_ltype;         /* nltype: compared with a site type in the wall-force
                                bookkeeping, as forcem.f90:481,540 do                 */
    const int32_t *site_type;    /* [n_total] 1-based                                 */
    const double  *charge;       /* [n_total] e                                       */
    const int32_t *nsi;          /* [n_types] sites carried by a lead of the type, 0 = not a lead */
    const int32_t *ijtype;       /* [n_types*n_types] row-major, 1-based pair number  */
    const int32_t *vdw_kind;     /* [n_pairs] 0 none, 1 Lennard-Jones, 2 Buckingham   */
    const double  *vdw_eps4;     /* [n_pairs] 4 eps          (sysdef.f90:372)         */
    const double  *vdw_sig6;     /* [n_pairs] sigma^6        (sysdef.f90:373)         */
    const double  *vdw_a;        /* [n_pairs] Buckingham A                            */
    const double  *vdw_rhoi;     /* [n_pairs] 1/rho          (sysdef.f90:374)         */
    const double  *vdw_c;        /* [n_pairs] Buckingham C                            */
    /* distance constraints of surface groups: excluded from the pair list
       (nnlist.f90:41-43,69-71) and their self term (forcem.f90:808-819)              */
    int32_t        n_constraints;
    const int32_t *con_a, *con_b;   /* [n_constraints] 0-based site indices           */
    const double  *con_len;         /* [n_constraints] constrained length             */
    /* harmonic angles of surface groups (forcem.f90:565-652)                         */
    int32_t        n_angles;
    const int32_t *ang_i, *ang_j, *ang_k;  /* [n_angles] 0-based                      */
    const double  *ang_kt, *ang_theta;     /* [n_angles] force constant, radians      */
    /* wall tail-correction inputs (force_real_init, forcem.f90:167-228)             */
    double  wall_lw;             /* lw: walls at z = +-lw/2                            */
    double  wall_cell[3];        /* cx, cy, cz                                         */
    int32_t wall_nss[2];         /* nss of the two crystal site types                 */
} mdf_topology;

typedef struct mdf_params {
    double  box[3];          /* lx, ly, lz; the reference needs lx = ly = min (forcem.f90:674-685) */
    double  rcut;            /* group cutoff on the lead-lead distance (forcem.f90:299,429) */
    double  skin;            /* Verlet skin (nnlist.f90:23-26, forcem.f90:245-250)  */
    double  alpha;           /* Ewald alpha, used with MDF_ALPHA_GIVEN               */
    int32_t alpha_mode;
    int32_t kmax;            /* k index bound; klmax/kmmax/knmax from the box (forcem.f90:673-678) */
    int32_t z_mode;          /* MDF_Z_*                                              */
    int32_t ew3dc;           /* 1: Yeh-Berkowitz slab correction (forcem.f90:923-928) */
    int32_t wall_lr;         /* 1: wall long-range tail correction (forcem.f90:509-543) */
    int32_t precision;       /* MDF_FP64 / MDF_FP32_ACC64                            */
    int32_t const_mode;      /* MDF_CONST_*                                          */
    int32_t erfc_mode;       /* MDF_ERFC_*                                           */
    int32_t max_neighbors;   /* per-lead list capacity, 0 = default                  */
    int32_t device;          /* HIP device ordinal                                   */
    int32_t virial;          /* 1: accumulate the real-space pair virial (the reference
                                has none; 0 leaves that part of mdf_outputs.virial zero) */
} mdf_params;

typedef struct mdf_outputs {
    /* arrays: caller-owned, may be NULL to skip                                       */
    double *f_site;          /* [3*n_move]  x block, y block, z block: site forces
                                before the COM reduction (fx, fy, fz of force_sum)    */
    double *f_com;           /* [3*n_rig]   COM forces fcx, fcy, fcz                  */
    double *torque;          /* [3*n_rig]   tx, ty, tz                                */
    double *e_vdw;           /* [n_pairs]   elj                                       */
    double *e_coul_real;     /* [n_pairs]   eqq                                       */
    /* scalars                                                                         */
    double e_recip;          /* uewa   */
    double e_recip_ll, e_recip_wl, e_recip_ww;   /* ewqLL, ewqWL, ewqWW */
    double e_self;           /* uself  */
    double e_slab;           /* urf    */
    double e_wall_lr;        /* ulr    */
    double e_conf;           /* uconf  */
    double e_avex;           /* uavex after force_sum */
    double press_wall;       /* press  */
    double f_wall[9];        /* fwsx(-1:1), fwsy(-1:1), fwsz(-1:1) */
    double virial[9];        /* site-pair virial tensor W_ab (row-major): real-space pairs,
                                reciprocal and slab terms; the reference has none      */
    int32_t list_rebuilt;    /* 1 if this call rebuilt the neighbour list              */
    int32_t n_pairs_listed;  /* ipack  */
    int32_t n_wall_pairs_listed; /* wpack */
    int32_t n_kvec;          /* k-vectors in the set */
    int32_t n_list_builds;   /* list builds since mdf_create, the first one included */
} mdf_outputs;

typedef struct mdf_system mdf_system;

/* force_init: validate, derive constants and tables, upload, build the first list.  */
int mdf_create(const mdf_topology *top, const mdf_params *par,
               const double *x, const double *y, const double *z,
               const double *cx, const double *cy, const double *cz,
               mdf_system **out);

/* force_get on host arrays: x,y,z have n_total entries (only moving sites are read),
   cx,cy,cz have n_rig entries.  Outputs are written into *out.                       */
int mdf_compute(mdf_system *sys,
                const double *x, const double *y, const double *z,
                const double *cx, const double *cy, const double *cz,
                mdf_outputs *out);

/* force_get on device-resident inputs (n_move and n_rig entries).  Enqueued on the
   handle's stream and not waited for; results stay on the device until
   mdf_read_outputs.  stream is the stream that produced the inputs (a hipStream_t): the
   step starts after the work enqueued on it so far; NULL means the null stream.      */
int mdf_compute_device(mdf_system *sys,
                       const double *d_x, const double *d_y, const double *d_z,
                       const double *d_cx, const double *d_cy, const double *d_cz,
                       void *stream);

/* Wait for the last mdf_compute_device and copy its results into *out.              */
int mdf_read_outputs(mdf_system *sys, mdf_outputs *out);

/* Device pointers of the last results: site forces [3*n_move], COM forces and torques
   [3*n_rig] each (valid until the next compute).                                      */
int mdf_device_results(mdf_system *sys, const double **d_f_site, const double **d_f_com,
                       const double **d_torque);

/* Pair list as it stands (the reference's list1/list2 and listw1/listw2 sets, 0-based,
   i < j, ordered by i then j).  Capacities in, counts out.                            */
int mdf_get_pairs(mdf_system *sys, int32_t *i_mm, int32_t *j_mm, int32_t cap_mm, int32_t *n_mm,
                  int32_t *i_mw, int32_t *j_mw, int32_t cap_mw, int32_t *n_mw);

/* force_zero */
int mdf_zero(mdf_system *sys);

/* k-vector set of the handle: (kl,km,kn) triples and Ak, in the reference's order.   */
int mdf_get_kvectors(mdf_system *sys, int32_t *kvec, double *ak, int32_t cap, int32_t *n);

/* Last structure factor S(k) = S_moving + S_wall (re, im interleaved) per k-vector.  */
int mdf_get_structure_factor(mdf_system *sys, double *s, int32_t cap);

void mdf_destroy(mdf_system *sys);

/* Message for the last failing call (per thread).                                    */
const char *mdf_last_error(void);

/* ---- batches of independent systems (no counterpart in the reference) ----------------
   Many systems - replicas of uneven size, each with its own box, lists and k-set - stepped
   together: every stage of the step is one launch over all of them, and the whole step is
   one replayed graph.  The systems stay owned by the caller and must not be stepped on their
   own while the batch is in flight.  All must share device, precision, erfc form and the virial switch;
   the batch uses the first system's pair lanes per lead until mdf_batch_autotune picks its own. */
typedef struct mdf_batch mdf_batch;

int mdf_batch_create(mdf_system *const *systems, int32_t n, mdf_batch **out);

/* Enqueue one step of every system on device-resident positions: x[i] .. cz[i] are system
   i's arrays, as for mdf_compute_device (cx/cy/cz may be NULL when no system has rigid
   molecules); stream as for mdf_compute_device.  Returns without waiting. */
int mdf_batch_compute_device(mdf_batch *b, const double *const *x, const double *const *y,
                             const double *const *z, const double *const *cx, const double *const *cy,
                             const double *const *cz, void *stream);

/* Wait for the step and copy every system's results into outs[0 .. n-1] (as
   mdf_read_outputs).  A system whose pair list overflowed is grown and redone alone. */
int mdf_batch_read_outputs(mdf_batch *b, mdf_outputs *outs);

/* Pick the batch's launch shapes by timing whole batch steps on the inputs of the last
   mdf_batch_compute_device (those arrays must still hold valid positions): lanes per lead of
   the pair kernel, and the structure-factor and back-projection blocks per launch of the
   whole batch, shared out to the systems in proportion to their k-vectors x moving sites.
   Each candidate runs reps + 1 times.  chosen (optional, 4 entries) receives lanes per lead,
   0, structure-factor and back-projection blocks of the batch.  Results of the last step are
   overwritten: call it between steps.  No counterpart in the reference. */
int mdf_batch_autotune(mdf_batch *b, int32_t reps, int32_t *chosen);

void mdf_batch_destroy(mdf_batch *b);

/* Pick the launch shapes of this handle's kernels by timing candidates on its current
   state: lanes per lead (8/16/32) and loop form of the pair kernel, blocks per launch of the
   structure-factor and back-projection kernels.  Each candidate runs reps + 1 times; the
   fastest is kept for later steps.  chosen (optional, 4 entries) receives lanes per lead,
   pair loop form (0: each lead's partners inside the cut-off compacted before the pair
   arithmetic, 1: row-strided), structure-factor blocks and back-projection blocks.  Device results of the
   last step are overwritten: call between steps, before reading them again.  No
   counterpart in the reference (its loops have no launch shape). */
int mdf_autotune(mdf_system *sys, int32_t reps, int32_t *chosen);

/* Set the launch shapes mdf_autotune chooses, by hand: pair lanes per lead (8, 16, 32),
   pair loop form (0 compacted, 1 row-strided), structure-factor and back-projection blocks
   per launch.  No counterpart in the reference. */
int mdf_set_launch_shapes(mdf_system *sys, int32_t pair_lanes, int32_t pair_strided, int32_t sf_blocks,
                          int32_t bp_blocks);

/* Which reciprocal-sum kernels the handle runs: 1 = on the matrix cores (the FP64 form's
   structure factor and back-projection on v_mfma_f64_16x16x4_f64, environment MDF_MMA64=0 at
   creation selecting the vector-ALU kernels; the FP32 form's as FP16 hi/lo split products on
   v_mfma_f32_16x16x32_f16, MDF_SF_MMA=0 / MDF_BP_MMA=0 selecting the vector ALU), 0 = the
   vector ALU.  No counterpart in the reference (its loops have no kernel choice). */
int mdf_reciprocal_kernels(const mdf_system *sys, int32_t *sf_matrix, int32_t *bp_matrix);

/* ---- observables of the reference's hot loops (src/forcem.f90:347-354,388-395,486-495,
   930-934; arrays of src/measure.f90:18-22,80-86) -------------------------------------
   Accumulated on the device over the steps since they were switched on (or last reset):
   ngc, the site-pair distance histogram per type pair, bin int(r / rdel + 0.5), as int64
   [nbins][n_pairs] (the memory of the reference's ngc(nijt, 0:lmax)), over every site of a
   lead's molecule against each frozen site inside the cut-off and, with a periodic z mode,
   also every site pair of the moving molecule pairs inside the cut-off
   (src/forcem_back.f90:348-349,380-382, counts the slab form leaves commented out at
   src/forcem.f90:357,399); efipos, the reference's z-force profile of
   the mobile ions of types 3 and 4, bin int((z + lw/2) / rdel + 0.5), as double [nbins][2]
   (efipos(1:2, 0:lmax)).  nbins = int(lw / rdel + 0.5) + 2.  rdel <= 0 switches them off. */
int mdf_observables_enable(mdf_system *sys, double rdel);
/* Copy the accumulated observables (either pointer may be NULL); reset non-zero zeroes them. */
int mdf_get_observables(mdf_system *sys, int64_t *ngc, double *efipos, int32_t *nbins, int32_t reset);

/* ---- ngang, the angle histogram of force_intra (src/forcem.f90:617-625; the array of
   src/measure.f90:20,96-99, reset at :198, written out at :612-618) ----------------------
   Accumulated on the device at every step from the call on: per angle term, bin
   l = int((cos theta + 1) * 100) + 1 (0..201) of the term's group type, with the reference's
   cosine of a = r_k - r_j, b = r_j - r_i (x, y minimum image; z too with a periodic z mode).
   ang_type[n_angles] is the group type of each angle term of the topology (the reference's
   iatype, 1..n_atypes); the counts are int64 [202][n_atypes], the memory of the reference's
   ngang(n_atypes, 0:201).  ang_type NULL switches it off. */
int mdf_angle_histogram_enable(mdf_system *sys, const int32_t *ang_type, int32_t n_atypes);
/* Copy the accumulated counts (ngang may be NULL); reset non-zero zeroes them. */
int mdf_get_angle_histogram(mdf_system *sys, int64_t *ngang, int32_t reset);

/* ---- Widom test-particle insertions: the reference's chempot (src/forcem.f90:1226-1507,
   called at the end of force_get when lchempot is set, :126; inputs src/main.f90:148-152,
   set-up src/measure.f90:149-166, output src/measure.f90:700-705) --------------------------
   After every step, up to n_insert SPC/E waters are inserted with the reference's generator
   (src/rndmod.f90:28-35): oxygen uniform in x, y and at a distance from the pore center in
   (lwh - wall_max, lwh - wall_min); 20 orientations in bins above bin_mid, 1 in the others.
   An insertion whose oxygen comes within 2.35 A of an oxygen or 2 A of another non-hydrogen
   site ends the step's insertions, as in the reference.  pave(bin) accumulates
   exp(-u / RT) over the orientations and inave(bin) their number; the excess chemical
   potential is -RT ln(sum pave / sum inave) (src/measure.f90:701).  The insertion energy
   is the reference's: the oxygen's vdW and real-space Coulomb with every non-hydrogen site
   within rcut and whole partner molecules for the hydrogens, 0.5 sum A_k Re(rho_ins conj S)
   with this step's S(k), the wall tail terms of type 1, and the slab term below.
   ew3dc: the reference's chempot multiplies a local `crf` that nothing assigns
   (src/forcem.f90:1236,1477); 1 uses the step's slab-correction coefficient
   2 crf0 facq sum(q z) (the module variable it shadows, :925-927), 0 leaves the term out.
   Runs in FP64 on the handle's stream after each step (not in batches, not with the
   k-space surrogate).  The host draws each step's geometry, so a step with insertions
   waits for the previous step's insertions to finish. */
typedef struct {
    int32_t n_insert;        /* ichmax                                                   */
    int32_t n_bins;          /* lchmax                                                   */
    int32_t bin_mid;         /* ilmid                                                    */
    int32_t ew3dc;           /* 1: slab term with the step's coefficient; 0: none        */
    double wall_min;         /* lwmin as tot.inp gives it: distance from the wall        */
    double wall_max;         /* lwmax                                                    */
    double temperature;      /* chtemper (K); <= 0: 300 K as the reference without temper */
    double seed;             /* generator state; <= 0: the reference's rank-0 start      */
    int32_t max_partners;    /* partner list per insertion; 0: min(non-H sites, 4096)    */
    int32_t pad;
} mdf_chempot_params;
/* NULL params switch the insertions off. */
int mdf_chempot_enable(mdf_system *sys, const mdf_chempot_params *params);
/* Copy pave [n_bins], inave [n_bins] and the generator state (any may be NULL); reset
   non-zero zeroes pave and inave (not the generator). */
int mdf_get_chempot(mdf_system *sys, double *pave, int64_t *inave, double *seed, int32_t reset);

/* ---- reciprocal-force surrogate with multiple time steps (no counterpart in the reference,
   whose force_get evaluates the full k-sum every step, src/forcem.f90:827-918) ----------
   Between full k-sums the k-space site force is predicted by a small per-site network of
   the surroundings within 6 A plus the correction [full - predicted] from the last full
   sum, which runs every `interval`-th step and also judges the prediction: when the
   relative force RMS error it finds exceeds `tol`, the scheme falls back to full sums
   every step (training again) for `train_frames` steps before it resumes.  The network is
   trained on the GPU, with Adam, on the full-sum frames of the run itself: the first
   `train_frames` steps run the full sum and train (`train_iters` Adam steps per frame).
   model MDF_MTS_HELD uses no network: the last full sum's force is held in between.
   On steps without the full sum, e_recip (and e_conf, e_avex) is the last full sum's value
   plus the work of the k-space force along the path moved since (trapezoid rule), and the
   k-space virial is the last full sum's.  Needs every moving site in a lead's molecule and
   6 A + twice the largest site-to-lead distance below rc; not available in batches. */
enum { MDF_MTS_HELD = 0, MDF_MTS_NETWORK = 1 };

typedef struct {
    int32_t interval;       /* full k-sum every interval-th step (>= 1)                   */
    int32_t model;          /* MDF_MTS_NETWORK or MDF_MTS_HELD                             */
    int32_t train_frames;   /* full-sum steps before multiple time steps start, and after a fallback */
    int32_t train_iters;    /* Adam steps per full-sum frame while collecting (0..16)      */
    double tol;             /* judged relative force RMS error that triggers the fallback  */
    float lr;               /* Adam learning rate                                          */
    uint32_t seed;          /* initial weights                                             */
} mdf_mts_params;

typedef struct {
    int32_t state;          /* 0 collecting (full sum every step), 1 multiple time steps   */
    int32_t last_exact;     /* the last step ran the full sum                              */
    int64_t n_exact, n_surrogate, n_fallback, n_judged, adam_steps;
    double err_last, err_max, err_mean;   /* judged relative force RMS errors              */
    double loss_last;       /* training loss of the last Adam step                         */
} mdf_mts_stats;

/* Switch the scheme on (p non-NULL: fresh network unless weights are set afterwards) or off
   (p NULL).  Takes effect from the next step. */
int mdf_mts_enable(mdf_system *sys, const mdf_mts_params *p);
int mdf_mts_get_stats(mdf_system *sys, mdf_mts_stats *out);

/* Network state as one float array: parameters (W1 [64][48], b1 [64], W2 [64][64], b2 [64],
   W3 [16][64], b3 [16]), then input means [48], input deviations [48] and vector scales [16]
   (MDF_MTS_WEIGHTS floats).  Setting weights restarts the collect phase with them. */
#define MDF_MTS_WEIGHTS 8448
int mdf_mts_get_weights(mdf_system *sys, float *w, int32_t cap);
int mdf_mts_set_weights(mdf_system *sys, const float *w, int32_t n);

/* The surrogate's inputs [n_move][48] and vector features [n_move][16][3] at the positions of
   the last step (scheme on, network model). */
int mdf_mts_features(mdf_system *sys, float *feat, float *vec);

/* One evaluation of the network on given inputs, outside any handle: predicted forces
   f_out [3][n] of frame a (feat [n][48], vec [n][16][3]); with targets of frame a and of a
   second frame b (feat_b, vec_b, target_b), the gradient of the loss on their change,
   sum |(S_a - S_b) - (T_a - T_b)|^2 / (n mean |T_a - T_b|^2), in grad_out
   [MDF_MTS_WEIGHTS - 112 + 1] (the last entry is the loss) and, for lr > 0, the parameters
   after one Adam step from zero moments (step number adam_t + 1).  weights: as
   mdf_mts_get_weights.  Outputs may be NULL; target NULL: forward only.  Runs the scheme's
   own kernels, on the current device. */
int mdf_mts_eval(const float *weights, int32_t n, const float *feat, const float *vec, const double *target,
                 const float *feat_b, const float *vec_b, const double *target_b, float lr, int32_t adam_t,
                 float *f_out, float *grad_out, float *params_out);

/* mdf_mts_eval on device-resident arrays, allocating nothing.  No counterpart in the
   reference (nor an allocating one there): for callers that cannot allocate device memory -
   code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.  Every
   pointer is a device pointer: weights [MDF_MTS_WEIGHTS], feat [n][48], vec [n][16][3], targets
   [3][n] (FP64), f_out [3][n] (required), grad_out [MDF_MTS_WEIGHTS - 112 + 1] and params_out
   [MDF_MTS_WEIGHTS - 112] as in mdf_mts_eval (written only with targets).  The weights are not
   checked (mdf_mts_eval checks its host copy); equal targets make the gradient non-finite.
   Enqueued on `stream` (a hipStream_t; NULL: the null stream) and not waited for.  The scratch
   (the scheme state, bfloat16 weight copies, Adam moments, per-block gradient partials:
   MDF_MTS_EVAL_SCRATCH_BYTES) is one static array per device, which every call on that device
   shares and which keeps the last call's contents (each call sets it up again): calls that
   overlap on one device race.  n <= MDF_MTS_EVAL_MAX_SITES (the kernels index sites in 32-bit
   arithmetic); larger n is refused with MDF_ERR_ARG. */
#define MDF_MTS_EVAL_SCRATCH_BYTES 4430080   /* 4.2 MiB */
#define MDF_MTS_EVAL_MAX_SITES (1 << 24)
int mdf_mts_eval_device(const float *weights, int32_t n, const float *feat, const float *vec, const double *target,
                        const float *feat_b, const float *vec_b, const double *target_b, float lr, int32_t adam_t,
                        float *f_out, float *grad_out, float *params_out, void *stream);

/* Timing.  Steps normally replay one captured HIP graph; with stage timing on they use
   direct launches with HIP events between the stages instead (events recorded during
   graph capture cannot be timed).  mdf_stage_times gives names and mean milliseconds per
   step over the steps since mdf_reset_stage_times: the stages (from steps run with stage
   timing on) followed by "step", the whole step measured in both modes, then "k_pair",
   "k_sf", "k_sf_reduce" and "k_bp": mean durations per launch of the pair kernel and
   the three reciprocal-sum kernels, from start and stop events their own dispatches
   record in the steps with stage timing on.  cap: capacity of names/ms; n out: number
   of entries.  on = 2 also rebuilds the pair list on every
   step (the "build" stage then times full list builds; the forces do not change: the
   list holds every pair within the cutoff either way).                                */
int mdf_set_stage_timing(mdf_system *sys, int32_t on);
int mdf_stage_times(mdf_system *sys, const char **names, float *ms, int32_t cap, int32_t *n);
int mdf_reset_stage_times(mdf_system *sys);

#ifdef __cplusplus
}
#endif
#endif
