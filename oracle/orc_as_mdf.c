/*
 * The CPU oracle behind the engine's C entry points - TEST INFRASTRUCTURE ONLY.
 *
 * Exports mdf_create / mdf_compute / mdf_zero / mdf_destroy / mdf_last_error forwarding to
 * the oracle (oracle.c), so a program written against include/mdforce.h - here the NVE
 * driver with the engine's drop-in Fortran force module - can run on the oracle instead of
 * the GPU.  That gives the defect-free FP64 reference trajectory for the drift test.
 * Built by oracle/build_ref.sh into oracle/_ref; never linked into the product.
 */
#include "../include/mdforce.h"

typedef struct orc_system orc_system;
int orc_create(const mdf_topology *top, const mdf_params *par, const double *x, const double *y, const double *z,
               const double *cx, const double *cy, const double *cz, orc_system **out);
int orc_compute(orc_system *s, const double *x, const double *y, const double *z, const double *cx,
                const double *cy, const double *cz, mdf_outputs *o);
void orc_destroy(orc_system *s);
const char *orc_last_error(void);

int mdf_create(const mdf_topology *top, const mdf_params *par, const double *x, const double *y, const double *z,
               const double *cx, const double *cy, const double *cz, mdf_system **sys) {
    return orc_create(top, par, x, y, z, cx, cy, cz, (orc_system **)sys);
}

int mdf_compute(mdf_system *sys, const double *x, const double *y, const double *z, const double *cx,
                const double *cy, const double *cz, mdf_outputs *out) {
    return orc_compute((orc_system *)sys, x, y, z, cx, cy, cz, out);
}

int mdf_zero(mdf_system *sys) { (void)sys; return MDF_OK; }

void mdf_destroy(mdf_system *sys) { orc_destroy((orc_system *)sys); }

const char *mdf_last_error(void) { return orc_last_error(); }

/* The oracle computes every step with the full k-sum: multiple time steps are refused
   (switching them off is accepted), and there are no counters to report. */
int mdf_mts_enable(mdf_system *sys, const mdf_mts_params *p) {
    (void)sys;
    return p ? MDF_ERR_ARG : MDF_OK;
}

int mdf_mts_get_stats(mdf_system *sys, mdf_mts_stats *out) {
    (void)sys;
    (void)out;
    return MDF_ERR_ARG;
}

/* the CPU statement has no observables or insertions: it reports empty histograms and
   refuses insertions */
int mdf_observables_enable(mdf_system *sys, double rdel) {
    (void)sys;
    (void)rdel;
    return MDF_OK;
}

int mdf_get_observables(mdf_system *sys, int64_t *ngc, double *efipos, int32_t *nbins, int32_t reset) {
    (void)sys;
    (void)ngc;
    (void)efipos;
    (void)reset;
    if (nbins) *nbins = 0;
    return MDF_OK;
}

/* the angle histogram is not restated here: counts stay zero */
static int32_t shim_n_atypes = 0;
int mdf_angle_histogram_enable(mdf_system *sys, const int32_t *ang_type, int32_t n_atypes) {
    (void)sys;
    shim_n_atypes = ang_type ? n_atypes : 0;
    return MDF_OK;
}

int mdf_get_angle_histogram(mdf_system *sys, int64_t *ngang, int32_t reset) {
    (void)sys;
    (void)reset;
    if (ngang)
        for (int32_t k = 0; k < 202 * shim_n_atypes; k++) ngang[k] = 0;
    return MDF_OK;
}

int mdf_chempot_enable(mdf_system *sys, const mdf_chempot_params *p) {
    (void)sys;
    return p ? MDF_ERR_ARG : MDF_OK;
}

int mdf_get_chempot(mdf_system *sys, double *pave, int64_t *inave, double *seed, int32_t reset) {
    (void)sys;
    (void)pave;
    (void)inave;
    (void)seed;
    (void)reset;
    return MDF_ERR_ARG;
}
