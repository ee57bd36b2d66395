/*
 * umwm_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * FP64 CPU restatement of the UMWM timestep hot path, used by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg as the checker.
 * It is never linked into, loaded by or called from the product library.
 *
 * Arrays keep the reference's layout and 1-based index values:
 *   spectra   e(o,p,i), i = 0..im (0 = ghost)   -> [(i*pm + p-1)*om + o-1]
 *   tables    t(o,i),   i = 0..im               -> [i*om + o-1]
 *   unrolled  a(i),     i = 1..imm              -> [i]
 *   2-D       a(m,n)                            -> [(n-1)*mm + m-1]
 *   Stokes    us(i,l)                           -> [(l-1)*im + i-1]
 */
#ifndef UMWM_ORACLE_H
#define UMWM_ORACLE_H

#include "../include/umwm_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_model oracle_model;

/* umwm_initialize without I/O: grid, masks, remap, input_nc, init,
 * dispersion.  Returns NULL and fills err (if given) on a bad namelist. */
oracle_model* oracle_create(const umwm_params* p, const umwm_grid_input* g,
                            const umwm_forcing_input* f0, char* err,
                            int errlen);
void oracle_free(oracle_model* m);
void oracle_set_threads(int n);

/* dims[0..7] = im, imm, om, pm, mm, nm, lm, 8 */
void oracle_dims(const oracle_model* m, int dims[8]);

/* Pointer to internal storage of a named field; *n = element count,
 * *is_int = 1 for integer fields.  NULL if unknown. */
void* oracle_field(oracle_model* m, const char* name, long* n, int* is_int);
double oracle_get_scalar(oracle_model* m, const char* name);
int oracle_set_scalar(oracle_model* m, const char* name, double v);

/* Reference operators (same names as the Fortran subroutines). */
void oracle_forcinginput(oracle_model* m, const umwm_forcing_input* next);
void oracle_forcinginterpolate(oracle_model* m);
void oracle_sin_d12(oracle_model* m);
void oracle_sds_d12(oracle_model* m);
void oracle_snl_d12(oracle_model* m);
void oracle_s_ice(oracle_model* m);
void oracle_source(oracle_model* m);
void oracle_propagation(oracle_model* m);
void oracle_refraction(oracle_model* m);
void oracle_update(oracle_model* m);
void oracle_stress_atm(oracle_model* m);
void oracle_stress_ocn(oracle_model* m);
void oracle_diag(oracle_model* m);
void oracle_stokes_drift(oracle_model* m, int init);
void oracle_init_jonswap(oracle_model* m, const umwm_jonswap* js);

/* One inner-loop iteration (src/umwm_top.F90:91-141 without output). */
void oracle_substep(oracle_model* m);
/* Inner loop to sumt >= dtg, then stokes (if enabled) and diag; returns
 * the number of substeps. */
int oracle_run_dtg(oracle_model* m);

/* Closed-form building blocks exposed for pinning tests. */
double oracle_sheltering_coare35(double wspd);
double oracle_stress_tail(double wspd, double kmax);
double oracle_friction_velocity_skin(double wspd, double z, double nu,
                                     double kappa);
double oracle_drag_coefficient_skin(double cd_form, double wspd,
                                    double wspdrel, double z, double nu,
                                    double kappa);
double oracle_distance_haversine(double lon1, double lon2, double lat1,
                                 double lat2);

#ifdef __cplusplus
}
#endif

#endif
