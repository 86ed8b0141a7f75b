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
