/*
 * umwm_hip.h -- C-ABI of the MI355X (gfx950) implementation of the UMWM
 * timestep hot path.
 *
 * The reference operators are argument-less Fortran module subroutines that
 * read and write module globals (src/umwm_module.F90:103-207).  Here every
 * operator takes an opaque domain handle that owns the device-resident state
 * of one model domain; a Fortran host binds each entry point with
 * iso_c_binding (bind(C, name='umwm_hip_...')).  All entry points return 0 on
 * success and a negative umwm_status code on failure; umwm_hip_error_string()
 * and umwm_hip_last_error() describe the failure.  No entry point aborts the
 * process: the caller keeps the reference's fatal convention
 * (src/umwm_util.F90:129-153, raiseexception + stop).
 *
 * Array conventions at the boundary (what the get/set entry points exchange):
 *   - host arrays are double precision and use the reference's column-major
 *     shapes: spectra (om,pm,im), per-frequency tables (om,im), 2-D fields
 *     (mm,nm), per-point fields (im) or (imm);
 *   - index maps use the reference's 1-based values (ghost point = 0).
 * On the device the spectrum is stored E[o][p][i] with the sea-point index
 * contiguous; the transposition happens only inside get/set calls.
 */
#ifndef UMWM_HIP_H
#define UMWM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes returned by every entry point. */
typedef enum umwm_status {
  UMWM_OK = 0,
  UMWM_ERR_ARGUMENT = -1,   /* bad argument or namelist value             */
  UMWM_ERR_HIP = -2,        /* HIP runtime error (message has details)    */
  UMWM_ERR_NOMEM = -3,      /* device or host allocation failed           */
  UMWM_ERR_STATE = -4,      /* call made in the wrong order               */
  UMWM_ERR_UNKNOWN_FIELD = -5,
  UMWM_ERR_SIZE = -6,       /* buffer size does not match the field       */
  UMWM_ERR_IO = -7          /* file cannot be opened, written or parsed   */
} umwm_status;

/* Floating-point type of the spectrum and of the source/advection math. */
enum { UMWM_FP32 = 4, UMWM_FP64 = 8 };

/*
 * Namelist contents (namelists/main.nml groups DOMAIN, PHYSICS, GRID,
 * FORCING, FORCING_CONSTANT, OUTPUT, STOKES), read by the reference in
 * src/umwm_init.F90:67-92 and src/umwm_stokes.F90:231-240.  Logical switches
 * are int32 0/1.  Layout is plain C so a Fortran bind(C) derived type with
 * the same member order maps onto it.
 */
typedef struct umwm_params {
  /* &DOMAIN */
  int32_t mm, nm, om, pm;
  int32_t isglobal, restart;
  /* &GRID switches */
  int32_t gridfromfile, topofromfile, fillestuaries, filllakes;
  /* &FORCING switches: field read from forcing snapshots when 1 */
  int32_t winds, currents, air_density, water_density, seaice;
  /* &OUTPUT */
  int32_t stokes;
  int32_t lm;          /* ignored on input: the library counts the levels
                          as count(depths >= 0), src/umwm_stokes.F90:243 */
  int32_t reserved0;
  double fmin, fmax, fprog, dtg;
  /* &PHYSICS */
  double g, nu_air, nu_water, sfct, kappa, z, gustiness, dmin, explim;
  double sin_fac, sin_diss1, sin_diss2, sds_fac, sds_power, mss_fac;
  double snl_fac, sdt_fac, sbf_fac, sbp_fac;
  /* &GRID values */
  double delx, dely, dpt;
  /* &FORCING_CONSTANT */
  double wspd0, wdir0, uc0, vc0, rhoa0, rhow0, fice0, fice_lth, fice_uth;
  /* &STOKES: depths [m], positive down, first lm entries used */
  double depths[100];
} umwm_params;

/*
 * Grid inputs that the reference reads from input/umwm.gridtopo
 * (src/umwm_init.F90:397-552).  Each pointer may be NULL when the matching
 * switch in umwm_params is 0.  2-D arrays are (mm,nm), m fastest.
 */
typedef struct umwm_grid_input {
  const double* lon;        /* degrees east, used when gridfromfile        */
  const double* lat;        /* degrees north, used when gridfromfile       */
  const double* z;          /* bathymetry, negative below sea level; cells
                               with z >= 0 are land (topofromfile)         */
  const int32_t* exclude;   /* 2*nexclude 1-based (m,n) lake seeds         */
  int32_t nexclude;         /* (namelists/exclude.nml, filllakes)          */
  int32_t reserved0;
} umwm_grid_input;

/*
 * One forcing snapshot, the content of input/umwmin_<time>.nc
 * (src/umwm_io.F90:13-93).  A field is used only when its switch in
 * umwm_params is set; NULL otherwise.  2-D arrays are (mm,nm).
 */
typedef struct umwm_forcing_input {
  const double* uw;   const double* vw;     /* winds    */
  const double* uc;   const double* vc;     /* currents */
  const double* rhoa; const double* rhow;   /* densities */
  const double* fice;                       /* sea ice fraction */
} umwm_forcing_input;

/* JONSWAP initial spectrum (no counterpart in the reference, which
 * cold-starts from e = tiny at src/umwm_init.F90:1182).  See
 * umwm_hip_init_jonswap. */
typedef struct umwm_jonswap {
  double alpha, gamma, sigma_a, sigma_b;
  double fp;          /* peak frequency [Hz] when fp_from_wind == 0       */
  int32_t fp_from_wind;   /* 1: fp = clamp(0.13 g / max(U,3), 0.05, 0.5)  */
  int32_t dir_from_wind;  /* 1: mean direction = local wdir, else thw     */
  double thw;         /* mean direction [rad, mathematical convention]   */
} umwm_jonswap;

typedef struct umwm_domain umwm_domain;   /* opaque, one model domain      */
typedef struct umwm_batch umwm_batch;     /* opaque, many domains at once  */

/* ---------------------------------------------------------------------- */
/* Library                                                                 */
/* ---------------------------------------------------------------------- */
const char* umwm_hip_error_string(int status);
/* Message of the most recent failure on the calling thread. */
const char* umwm_hip_last_error(void);
/* Number of visible HIP devices (0 when none). */
int umwm_hip_device_count(int* count);

/* ---------------------------------------------------------------------- */
/* Setup: replaces umwm_initialize (src/umwm_top.F90:7-29)                 */
/* ---------------------------------------------------------------------- */
/*
 * Builds grid, masks, the sea-first unrolled index maps and every
 * coefficient table (src/umwm_init.F90 grid :381-568, masks :571-694,
 * remap :875-1237, init :1240-1419, dispersion :1422-1658) on the host in
 * double precision, then allocates the device state and uploads it.
 * `forcing0` holds the snapshot at the start time (input_nc(starttime),
 * src/umwm_io.F90:13-93) and may be NULL when no forcing switch is set.
 * `precision` is UMWM_FP32 or UMWM_FP64.  The spectrum starts at
 * e = tiny(e) as in the reference.
 */
int umwm_hip_create(const umwm_params* params, const umwm_grid_input* grid,
                    const umwm_forcing_input* forcing0, int precision,
                    umwm_domain** out);
int umwm_hip_destroy(umwm_domain* dom);
/* Dimensions: out[0]=im, out[1]=imm, out[2]=om, out[3]=pm, out[4]=mm,
 * out[5]=nm, out[6]=lm, out[7]=precision. */
int umwm_hip_dims(const umwm_domain* dom, int32_t out[8]);

/* Replaces the cold start with a JONSWAP x cos^2 spectrum. */
int umwm_hip_init_jonswap(umwm_domain* dom, const umwm_jonswap* js);

/* stokes_drift('init') and stokes_drift() (src/umwm_stokes.F90:13-132). */
int umwm_hip_stokes_drift(umwm_domain* dom, int init);

/* ---------------------------------------------------------------------- */
/* Reference operators, one entry point each.  Together they reproduce one */
/* iteration of the inner loop of umwm_run (src/umwm_top.F90:89-150) with  */
/* the reference's module arrays (ssin, sds, snl, ...) materialised.       */
/* ---------------------------------------------------------------------- */
/* forcinginterpolate (src/umwm_forcing.F90:40-99). */
int umwm_hip_forcinginterpolate(umwm_domain* dom);
/* forcinginput (src/umwm_forcing.F90:11-37): level n <- n+1, load n+1. */
int umwm_hip_forcinginput(umwm_domain* dom, const umwm_forcing_input* next);

/* A translating tropical cyclone as wind forcing, generated on the device:
 * no counterpart in the reference, which reads winds from files
 * (src/umwm_io.F90:13-93); it stands for the snapshot a driver would pass to
 * umwm_hip_forcinginput for a storm.  Holland (1980) gradient wind
 * v(r) = sqrt(B dp / rho (Rmax/r)^B exp(-(Rmax/r)^B) + (r f / 2)^2) - r f / 2,
 * r = distance from (xc, yc) + 1 m, f = 2 Omega sin(lat), turned inward by
 * `inflow_deg` from the cyclonic tangent, plus half the translation velocity
 * (trans_u, trans_v).  dp <= 0 takes dp = rho e vmax^2 / B.  Planar grids
 * only (gridfromfile = 0, isglobal = 0): the point (m, n) lies at
 * ((m-1) delx, (n-1) dely). */
typedef struct umwm_holland {
  double xc, yc;            /* storm centre (m) */
  double vmax, rmax, b;     /* m/s, m, shape parameter */
  double rho, dp;           /* air density (kg/m3), pressure deficit (Pa) */
  double trans_u, trans_v;  /* translation velocity (m/s) */
  double lat_deg, inflow_deg;
} umwm_holland;
/* umwm_hip_forcinginput with the next level's winds computed on the device
 * from `storm` (the domain's only switched forcing must be the wind). */
int umwm_hip_holland_forcinginput(umwm_domain* dom, const umwm_holland* storm);
/* sin_d12, sds_d12, snl_d12, s_ice (src/umwm_source_functions.F90:12). */
int umwm_hip_sin_d12(umwm_domain* dom);
int umwm_hip_sds_d12(umwm_domain* dom);
int umwm_hip_snl_d12(umwm_domain* dom);
int umwm_hip_s_ice(umwm_domain* dom);
/* source (src/umwm_physics.F90:9-86): dts, integration, e <- (e+ef)/2. */
int umwm_hip_source(umwm_domain* dom);
/* propagation, refraction (src/umwm_advection.F90:18,102). */
int umwm_hip_propagation(umwm_domain* dom);
int umwm_hip_refraction(umwm_domain* dom);
/* e = ef (src/umwm_top.F90:113). */
int umwm_hip_update(umwm_domain* dom);
/* stress(option), option "atm" or "ocn" (src/umwm_stress.F90:17).  "ocn"
 * reads ssin, sds and snl of the last source evaluation, as the reference
 * does at output time, so that substep must have run through the operator
 * entry points; after a fused substep it returns UMWM_ERR_STATE.  Outputs
 * (per sea point): tailocnx/y, taux/tauy_ocntop, taux/tauy_ocnbot,
 * taux/tauy_snl, epsx/epsy_ocn, epsx/epsy_atm, taux1..3, tauy1..3. */
int umwm_hip_stress(umwm_domain* dom, const char* option);
/* diag (src/umwm_physics.F90:89-205). */
int umwm_hip_diag(umwm_domain* dom);
/* Marks the end of the first substep (first = .false.,
 * src/umwm_top.F90:139). */
int umwm_hip_end_first(umwm_domain* dom);

/* ---------------------------------------------------------------------- */
/* Fused path: same results as the operator sequence above               */
/* ---------------------------------------------------------------------- */
/*
 * One substep: forcinginterpolate, sin_d12, sds_d12, snl_d12, s_ice,
 * source, propagation, refraction, e = ef, stress('atm'), and on the first
 * substep diag and first = .false. (src/umwm_top.F90:91-141), in four
 * fused kernels.  Returns the new sumt and the dts taken.
 */
int umwm_hip_substep(umwm_domain* dom, double* sumt, double* dts);

/*
 * `n` iterations of the same inner-loop body (src/umwm_top.F90:91-141),
 * launched back to back with no host synchronisation between them; the
 * scalars are read once at the end.  A substep that starts at
 * sumt >= dtg finds the model inactive and does nothing, so a caller may
 * run ahead of the loop test of :89.  Returns sumt and the dts of the
 * last active substep, and `nsub`, the substeps taken since the current
 * dtg began.
 */
int umwm_hip_substeps(umwm_domain* dom, int32_t n, double* sumt, double* dts, int32_t* nsub);
/*
 * umwm_hip_substeps without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same results as umwm_hip_substeps, bit for bit.  The fused
 * kernels' scratch (the frequency prefix of the cos^2 moments and the
 * stress partial sums, 2 x 4 x om x ldi values of the working precision,
 * ldi = im + 1 rounded up to 64) lives in one file-scope device array of
 * 32 MiB per GPU, shared by every call on that GPU and kept until the
 * process ends; a domain needing more returns UMWM_ERR_SIZE (FP64 with
 * om = 37 fits up to about 14,000 sea points, FP32 twice that).  The array
 * keeps the last call's contents and is cleared at the start of each call:
 * calls that overlap on one GPU (two streams, two threads) race on it.
 * The domain itself must already exist (umwm_hip_create allocates it).
 */
int umwm_hip_substeps_noalloc(umwm_domain* dom, int32_t n, double* sumt, double* dts,
                              int32_t* nsub);
/*
 * The inner loop to sumt >= dtg (src/umwm_top.F90:88-150) followed by the
 * per-dtg stokes_drift (when stokes) and diag (:152-155).  `nsub` receives
 * the number of substeps taken.  The first call of a run includes the
 * cold (dts = 0) substep.  The loop runs ahead: it launches as many
 * substeps as the last dts says remain (as umwm_hip_substeps) before it
 * reads sumt back, so the host waits once per run of substeps.
 */
int umwm_hip_run_dtg(umwm_domain* dom, int32_t* nsub);

/* ---------------------------------------------------------------------- */
/* State exchange                                                          */
/* ---------------------------------------------------------------------- */
/*
 * Copies a named field to/from a host double array in the reference's
 * shape (see top of file).  Spectra: e ef ssin sds snl dummy rotl rotr;
 * (om,im) tables: k kdk k3dk k4 oneoverk4 cp0 cg0 invcp0 fkovg logl2overz
 * cothkd sbf sdv sdt snl_arg bf1_renorm bf2_renorm dwn sice; per point:
 * wspd wdir uc vc rhoa rhow rhorat fice ustar cd shelt fcutoff oc
 * physics_time_step taux tauy taux_form tauy_form taux_skin tauy_skin
 * taux_diag tauy_diag tailatmx tailatmy ht mwp mwd mss mwl momx momy
 * cgmxx cgmxy cgmyy dwd dwl dwp dcp0 dcg0 dcp dcg dx dy d ar dxn dxs dye
 * dyw; maps (returned as doubles holding the reference's integers): ii mi
 * ni iw ie is in iiw iie iis iin mask; Stokes: us vs (im,lm), ds (im).
 * `n` is the number of doubles in `buf` and must equal the field size.
 */
int umwm_hip_get_field(umwm_domain* dom, const char* name, double* buf,
                       int64_t n);
int umwm_hip_set_field(umwm_domain* dom, const char* name, const double* buf,
                       int64_t n);
/* ---------------------------------------------------------------------- */
/* Output and restart files                                               */
/* ---------------------------------------------------------------------- */
/*
 * The reference writes NetCDF (umwmout_<time>.nc, umwmrst_<time>.nc);
 * NetCDF is not available here, so these write the same variables (names,
 * dimension names and order, units, single precision) into a raw container
 * described in umwm_amd/csrc/rawio.hpp and read by umwm_amd/rawfile.py.
 *
 * umwm_hip_write_output: output_grid_nc (src/umwm_io.F90:317-854), every
 * diagnostic on the (x, y) grid with NaN on land; the reference calls
 * stress('ocn') first (src/umwm_top.F90:160-170), and so should the caller.
 * umwm_hip_write_restart / umwm_hip_read_restart: restart_write and
 * restart_read (src/umwm_restart.F90:18-145), the spectrum F(f,th,x) and
 * ust(x) of the sea points.  Reading a restart into a domain created with
 * restart = 1 continues the run without the cold first substep.
 */
int umwm_hip_write_output(umwm_domain* dom, const char* path);
int umwm_hip_write_restart(umwm_domain* dom, const char* path);
int umwm_hip_read_restart(umwm_domain* dom, const char* path);

/* Scalars: sumt dts dta dtr dtamin cfllim cgmax mindelx dlnf dth bf1 bf2
 * first nsub.  Settable: sumt dts dtr_temp first, and keep_sources: set to 1
 * (it cannot be unset) before the substep that reaches dtg, it makes the
 * fused path save that substep's E^n and incoming ustar (one more spectrum
 * of device memory), so that stress("ocn") at output time can rebuild the
 * ssin, sds and snl the reference's output step reads (src/umwm_top.F90:
 * 160-170); the save costs one spectrum copy per dtg. */
int umwm_hip_get_scalar(umwm_domain* dom, const char* name, double* v);
int umwm_hip_set_scalar(umwm_domain* dom, const char* name, double v);
/* Sizes in doubles of a named field (0 if unknown). */
int64_t umwm_hip_field_size(const umwm_domain* dom, const char* name);

/* ---------------------------------------------------------------------- */
/* Timing of the fused kernels (HIP events on the domain's stream)        */
/* ---------------------------------------------------------------------- */
/* Enables per-kernel event timing of the fused path (off by default). */
int umwm_hip_set_timing(umwm_domain* dom, int enable);
/* Accumulated milliseconds and launch counts of the four fused kernels
 * and of the surrogate's inference since the last reset: ms[0..4] =
 * rates, integrate, advect, stress, surrogate (network and its inputs);
 * counts[0..4] likewise.  Resets the accumulators when reset != 0. */
int umwm_hip_get_timing(umwm_domain* dom, double ms[5], int64_t counts[5],
                        int reset);
/* Waits for all work queued on the domain's stream. */
int umwm_hip_synchronize(umwm_domain* dom);
/* Tile shape of the fused kernels, tile[5] = {D1, W1, D2, W2, NB}:
 * rates with D1 directions per thread (6, 9, 18 or 36) and W1 waves,
 * integrate with D2 (3, 6, 9, 12 or 18) and W2 waves (W in 1..8,
 * D*W >= pm), and advect prefetching NB directions ahead (2 or 4).
 * Used by the autotuner (umwm_amd/autotune.py). */
int umwm_hip_set_tile(umwm_domain* dom, const int32_t tile[5]);
int umwm_hip_get_tile(const umwm_domain* dom, int32_t tile[5]);

/* ---------------------------------------------------------------------- */
/* Batched ensembles: many independent domains of uneven sizes advance in */
/* one launch per kernel per substep, each with its own dts.  No         */
/* counterpart in the reference (one domain per process).                 */
/* ---------------------------------------------------------------------- */
/* All domains must share the precision; they are borrowed, not owned. */
int umwm_hip_batch_create(umwm_domain* const* doms, int32_t n,
                          umwm_batch** out);
int umwm_hip_batch_destroy(umwm_batch* b);
/* Inner loop of every member to its own sumt >= dtg, replaying one
 * hipGraph per substep when use_graph != 0 (one launch per kernel per pm
 * class of members).  nsub[k] receives member k's substep count (array of
 * n, may be NULL). */
int umwm_hip_batch_run_dtg(umwm_batch* b, int use_graph, int32_t* nsub);
/* n substeps of every member, launched back to back (hipGraph replays
 * when use_graph != 0) with one host wait; a member whose sumt has reached
 * dtg does nothing.  The caller starts each dtg (as umwm_hip_batch_run_dtg
 * does) by setting sumt = 0 after loading the next forcing.  nsub[k]
 * receives member k's substeps taken this dtg (array of n members, may be
 * NULL). */
int umwm_hip_batch_substeps(umwm_batch* b, int32_t n, int use_graph, int32_t* nsub);
/* Tile shape of the batch's fused launches (layout of umwm_hip_set_tile). */
int umwm_hip_batch_tile(const umwm_batch* b, int32_t out[5]);
/* Sets the batch's tile shape (same rules as umwm_hip_set_tile over the
 * largest pm of the batch).  Members are launched in groups of equal pm;
 * each group uses the tile's D1, D2 and NB with W = ceil(pm / D). */
int umwm_hip_batch_set_tile(umwm_batch* b, const int32_t tile[5]);

/* ---------------------------------------------------------------------- */
/* Learned source-term surrogate.  No counterpart in the reference, whose */
/* source terms are closed forms (src/umwm_source_functions.F90:16-105,   */
/* combined in src/umwm_physics.F90:59-68).  A multilayer perceptron maps */
/* one sea point's spectrum (ln e of the om x pm bins), wind (u10, v10)   */
/* and ln depth to the one non-local factor of the combined source rate   */
/* R = ssin - sds - sbf: the directional spreading factor dm^2 = (1 +     */
/* mss_fac mss)^2 of sds_d12 in every bin, learned as ln dm^2.  The fused */
/* kernels form R from it and the closed forms of the bin (ssin, sbf and  */
/* the saturation form cf coth (e k^4)^p of sds).  Weights and            */
/* activations are bfloat16, products accumulate in FP32 on the MFMA      */
/* units; the network is trained on the device from samples the plain     */
/* source kernels produce.                                                */
/* ---------------------------------------------------------------------- */
typedef struct umwm_surrogate umwm_surrogate;   /* opaque */

typedef struct umwm_surrogate_config {
  int32_t om, pm;       /* spectral shape served: om*pm + 3 inputs, om*pm outputs */
  int32_t hidden;       /* width of each hidden layer (512) */
  int32_t layers;       /* hidden layers, GELU (3) */
  int32_t batch;        /* training rows per Adam step, a multiple of 8 */
  int32_t chunk;        /* sea points per inference launch (>= 64) */
  int64_t capacity;     /* training samples held on the device */
  uint64_t seed;        /* weight initialisation and minibatch draws */
  double fmin, fmax;    /* frequency grid of the served domains */
} umwm_surrogate_config;

/* How an attached domain uses the network.  Every audit_stride-th group of
 * 64 sea points (rotating with the substep; 0: none) is audited: the
 * dissipation the network's dm^2 gives is compared with the plain one.  A
 * point falls back to the plain source terms where its wind speed or depth
 * lies outside the training range widened by range_margin (relative),
 * where a normalised spectral input exceeds zmax standard deviations,
 * where in more than resid_tol of its energy-weighted bins the network's
 * dm^2 leaves the bounds the moment prefix sets, [1, (1 + 2 mss_fac m0)^2]
 * widened by a quarter (m0 the zeroth moment below the bin's frequency),
 * or where an audit finds its energy-weighted error above audit_tol.
 * The audits of one substep together also gate the whole domain: when
 * their summed energy-weighted error exceeds audit_tol, every point takes
 * the plain source terms in the next substep (the audits go on, and reopen
 * the gate once the error is back within audit_tol).  Where the network
 * serves, its dm^2 is clamped to those bounds. */
typedef struct umwm_surrogate_policy {
  int32_t audit_stride;
  int32_t reserved0;
  double resid_tol, audit_tol, range_margin, zmax;
} umwm_surrogate_policy;

int umwm_hip_surrogate_create(const umwm_surrogate_config* cfg, umwm_surrogate** out);
int umwm_hip_surrogate_destroy(umwm_surrogate* s);
/* out[0] = inputs, out[1] = outputs, out[2] = hidden width, out[3] =
 * hidden layers, out[4] = samples collected, out[5] = 1 once trained. */
int umwm_hip_surrogate_dims(const umwm_surrogate* s, int32_t out[6]);
/* Layer `layer` (0 .. layers): weights [N][K] row-major (K inputs, N
 * outputs of the layer) and bias [N], FP32 master copies. */
int umwm_hip_surrogate_get_layer(umwm_surrogate* s, int32_t layer, float* w, float* b);
int umwm_hip_surrogate_set_layer(umwm_surrogate* s, int32_t layer, const float* w,
                                 const float* b);
/* Network outputs y (n x outputs) of n raw input rows x (n x inputs),
 * without normalisation. */
int umwm_hip_surrogate_forward(umwm_surrogate* s, const float* x, int32_t n, float* y);
/* The same outputs through the inference kernels a substep uses (the
 * last layer computed transposed, outputs per row of points, bfloat16). */
int umwm_hip_surrogate_infer(umwm_surrogate* s, const float* x, int32_t n, float* y);
/* One forward and backward pass over n (a multiple of 8) raw rows: the
 * mean squared error against t (n x outputs) and the gradients of every
 * layer, gw[l] [N][K] and gb[l] [N] (either may be NULL). */
int umwm_hip_surrogate_gradients(umwm_surrogate* s, const float* x, const float* t, int32_t n,
                                 double* loss, float* const* gw, float* const* gb);
/* Appends one training sample per sea point of the domain at its current
 * state: the plain source kernels (sin_d12, sds_d12, snl_d12) evaluate
 * ssin, sds and sbf on its spectrum.  `total` receives the samples held. */
int umwm_hip_surrogate_collect(umwm_surrogate* s, umwm_domain* dom, int64_t* total);
/* Normalisation over the first ntrain samples (all when <= 0), then
 * `steps` Adam steps on minibatches drawn from them; the mean loss of the
 * first and last few steps (normalised units). */
int umwm_hip_surrogate_train(umwm_surrogate* s, int32_t steps, double lr, int64_t ntrain,
                             double* loss_first, double* loss_last);
/* The normalisation the last training fixed: input means and inverse     */
/* standard deviations (inputs values each), output means and standard    */
/* deviations (outputs values each) of ln dm^2; any pointer may be NULL.  */
int umwm_hip_surrogate_get_norm(umwm_surrogate* s, float* in_mean, float* in_inv_sd,
                                float* out_mean, float* out_sd);
/* Mean loss over collected samples [first, first + n) (held-out rows). */
int umwm_hip_surrogate_eval(umwm_surrogate* s, int64_t first, int64_t n, double* loss);
/* Makes an FP32 domain's fused substeps take the network's rates where the
 * policy allows (NULL surrogate detaches).  The surrogate's buffers serve
 * one domain at a time: domains sharing it run one after another. */
int umwm_hip_surrogate_attach(umwm_domain* dom, umwm_surrogate* s,
                              const umwm_surrogate_policy* policy);
/* The network's spreading factor dm^2 of the last substep (before the
 * kernels clamp it to its bounds), (om, pm, im) doubles in the reference's
 * shape, and each sea point's choice (1: the network's rates) as `use`
 * (im). */
int umwm_hip_surrogate_fields(umwm_domain* dom, double* dm_net, double* use);
/* Statistics since attach or the last reset: out[0] point-substeps,
 * out[1] of which took the network's rates, out[2..4] fell back on range,
 * residual and audit, out[5] audited, out[6] energy-weighted audit error
 * of the dissipation (sum |sds_net - sds| e dk / sum sds e dk over the
 * audited points, sds_net = the clamped network dm^2 times the saturation
 * form). */
int umwm_hip_surrogate_stats(umwm_domain* dom, double out[7], int reset);

#ifdef __cplusplus
}
#endif

#endif /* UMWM_HIP_H */
