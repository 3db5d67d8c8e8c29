/*
 * gitm_vert.h -- C ABI of the MI355X implementation of GITM's vertical
 * dynamics step (HIP, gfx950).
 *
 * The reference path is a set of argument-less Fortran subroutines that talk
 * through module globals:
 *
 *   advance_vertical_all          src/advance_vertical.f90:8-39
 *     advance_vertical(iLon,iLat,iBlock)                 :45-219
 *       advance_vertical_1d_ausm  src/vertical_solver_ausm.f90:26-433
 *       advance_vertical_1d_rusanov src/vertical_solver_rusanov.f90:10-384
 *
 * One call of gitm_vert_advance replaces one call of advance_vertical_all for
 * every column it is given: it gathers each column from the caller's arrays
 * (described by strides, so GITM's own (-1:nLons+2,-1:nLats+2,-1:nAlts+2,..)
 * module arrays can be passed as they are), runs the RK4 vertical step with
 * the solver UseAUSMSolver selects (src/advance_vertical.f90:157-161), and
 * scatters the result back exactly as advance_vertical does (Rho = exp(LogRho),
 * Temperature = Temp/TempUnit, NDensityS = exp(LogNS), NDensity = sum, ...).
 * Plain pointers and sizes only; no C++ or torch types cross this boundary.
 */
#ifndef GITM_VERT_H
#define GITM_VERT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GITM_VERT_NSPECIES 6       /* ModEarth.f90:18 nSpecies            */
#define GITM_VERT_NIONS 10         /* ModEarth.f90:38 nIons (incl. e-)     */
#define GITM_VERT_MAX_NALTS 252    /* levels -1..nAlts+2 fit 256 lanes     */

/* Solver selector: UseAUSMSolver (ModInputs.f90:242, #AUSMSOLVER). */
enum { GITM_VERT_RUSANOV = 0, GITM_VERT_AUSM = 1 };

/* Status codes.  The reference stops the run (stop_gitm) where these are
 * returned; the caller is expected to do the same. */
enum {
  GITM_VERT_OK = 0,
  GITM_VERT_ENONPOSITIVE = 1, /* NDensityS <= 0 entering a column
                                 (advance_vertical.f90:96-100)            */
  GITM_VERT_ESINGULAR = 2,    /* zero row in the friction matrix
                                 (ludcmp.f:14-17)                          */
  GITM_VERT_ENEGION = 3,      /* negative ion density too close to the top
                                 (set_vertical_bcs.Earth.f90:617-620)      */
  GITM_VERT_EINVAL = 4,       /* bad arguments                             */
  GITM_VERT_EUNSUPPORTED = 5, /* a flag this library does not implement
                                 (tides, BC perturbation, UseDamping):
                                 run the Fortran path                      */
  GITM_VERT_EHIP = 6,         /* HIP runtime error                         */
  GITM_VERT_ENOMEM = 7,
  GITM_VERT_ETAPE = 8,        /* adjoint tape too small for the column     */
  GITM_VERT_ELOWTEMP = 9,     /* Temperature < 100 after the message
                                 exchange (exchange_messages_sphere.f90:74-78) */
  GITM_VERT_ESCRATCH = 10     /* the columns need more carries than the
                                 fixed scratch of gitm_vert_advance_noalloc
                                 holds                                     */
};

/* One array of the column working set.  Element (column c, level k,
 * component m) lives at
 *   ptr[ ilon*s_lon + ilat*s_lat + iblk*s_blk + (k - first_level)*s_lev
 *        + m*s_comp ]
 * where c = ilon + n_lons*(ilat + n_lats*iblk) (0-based) and k runs over
 * GITM's level index (-1..nAlts+2).  Strides are in elements (doubles). */
typedef struct gitm_vert_field {
  double *ptr;
  int64_t s_lon, s_lat, s_blk;
  int64_t s_lev;
  int64_t s_comp;
  int32_t first_level; /* -1 for (-1:nAlts+2) arrays, 0 for cMax_GDB      */
  int32_t reserved;
} gitm_vert_field;

typedef struct gitm_vert_columns {
  int32_t n_lons, n_lats, n_blocks; /* columns = n_lons*n_lats*n_blocks     */
  int32_t n_alts;                   /* nAlts (ModSize); levels -1..nAlts+2  */
  /* state, read and written (advance_vertical.f90:105-118, 163-185) */
  gitm_vert_field ndensity_s;        /* NDensityS(.., 1:6, ..)     [m^-3]   */
  gitm_vert_field vertical_velocity; /* VerticalVelocity(.., 1:6, ..)       */
  gitm_vert_field velocity;          /* Velocity(.., 1:3, ..)  E, N, Up     */
  gitm_vert_field temperature;       /* Temperature (divided by TempUnit)   */
  gitm_vert_field rho;               /* Rho                                 */
  /* outputs only; ptr may be NULL */
  gitm_vert_field log_ns;            /* LogNS(.., 1:6, ..)                  */
  gitm_vert_field ndensity;          /* NDensity                            */
  /* column inputs, read only */
  gitm_vert_field temp_unit;         /* TempUnit (calc_rates)               */
  gitm_vert_field mean_major_mass;   /* MeanMajorMass (calc_rates)          */
  gitm_vert_field gamma;             /* Gamma (calc_pressure)               */
  gitm_vert_field cmax;              /* cMax_GDB(.., iUp_, ..), levels 0..nAlts+1 */
  gitm_vert_field eddy;              /* KappaEddyDiffusion                  */
  gitm_vert_field altitude;          /* Altitude_GB                         */
  gitm_vert_field gravity;           /* Gravity_GB                          */
  gitm_vert_field inv_radial_distance; /* InvRadialDistance_GB             */
  gitm_vert_field dalt;              /* dAlt_GB                             */
  gitm_vert_field cos_latitude;      /* CosLatitude(iLat, iBlock); s_lev unused */
  /* ions; used only when use_ion_advection != 0 */
  gitm_vert_field idensity_s;        /* IDensityS(.., 1:10, ..)  read/written */
  gitm_vert_field ivelocity;         /* IVelocity(.., 1:3, ..)  read        */
  gitm_vert_field sza;               /* SZA(iLon,iLat,iBlock)               */
  gitm_vert_field mlat;              /* MLatitude(iLon,iLat,nAlts,iBlock)   */
  /* UseMSIS only: per column, what msis_bcs (init_msis.Earth.f90:367-494)
   * returns for level 0: components 0..5 LogNS(0,1:6) (ln m^-3), 6 Temp(0)
   * in K, 7 and 8 the east and north winds in m/s; s_lev unused.  Only
   * these enter: the BCs overwrite every level -1 value, and the MSIS
   * LogNS(0) of N(4S) and NO, by extrapolation
   * (set_vertical_bcs.Earth.f90:113-137,198-272,364-375). */
  gitm_vert_field msis_bc;
  /* UseGSWMTides / UseWACCMTides only: TidesEast, TidesNorth, TidesTemp
   * (ModTides.f90:29-31, (-1:nLons+2,-1:nLats+2,2,nBlocksMax)), levels -1
   * and 0 (first_level = -1); the east and north winds of levels -1:0 and
   * the temperature in K (GSWM: added to Temp(-1:0); WACCM: replaces it),
   * set_vertical_bcs.Earth.f90:150-159. */
  gitm_vert_field tides_east, tides_north, tides_temp;
} gitm_vert_columns;

/* Run parameters (ModInputs / ModPlanet values the path reads). */
typedef struct gitm_vert_params {
  double dt;                    /* Dt                                        */
  double omega_body;            /* OmegaBodyInput        ModInputs.f90:285   */
  double max_vertical_velocity; /* MaximumVerticalVelocity       :224        */
  double beta_limiter;          /* BetaLimiter                   :104        */
  double min_tec;               /* MinTEC                        :261        */
  double vert_tau;              /* VertTau(:) (ModEarth.f90:325), Rusanov    */
  double mass[GITM_VERT_NSPECIES];              /* Mass(1:6) [kg]           */
  double thermal_diff_coef[GITM_VERT_NSPECIES]; /* ThermalDiffCoefS         */
  int32_t use_neutral_friction;       /* ModInputs.f90:240                  */
  int32_t use_coriolis;               /* :235                               */
  int32_t use_ion_advection;          /* :311                               */
  int32_t use_improved_ion_advection; /* :259                               */
  int32_t use_nighttime_ion_bcs;      /* :260                               */
  int32_t use_msis;                   /* :190; needs c->msis_bc          */
  int32_t use_gswm_tides, use_waccm_tides; /* :195-196; need c->tides_*   */
  /* Must be 0 (otherwise GITM_VERT_EUNSUPPORTED): the user's BC
   * perturbation (user_bc_perturbation, user.f90:176) */
  int32_t use_bc_perturbation;        /* :402                               */
  int32_t during_perturb;             /* :121                               */
  /* UseDamping (:221): the Rusanov step's VertTau profile
   * (vertical_solver_rusanov.f90:581-584).  With the AUSM solver
   * GITM_VERT_EUNSUPPORTED: the reference's assignment there
   * (vertical_solver_ausm.f90:857-860) indexes VertTau with the exhausted
   * loop variable nAlts+3, outside the array, and the AUSM step never reads
   * VertTau, so its effect is whatever that store overwrites. */
  int32_t use_damping;
  int32_t reserved;
} gitm_vert_params;

typedef struct gitm_vert_error {
  int32_t code;   /* GITM_VERT_* of the first failure                      */
  int32_t column; /* first failing column (0-based), -1 if none            */
  int32_t level;  /* level index where known, else -99                     */
  int32_t reserved;
} gitm_vert_error;

typedef struct gitm_vert_handle gitm_vert_handle;

/* Fill `p` with the reference defaults (ModInputs.f90, ModEarth.f90
 * init_planet masses, calc_rates.Earth.f90:83-84 ThermalDiffCoefS). */
void gitm_vert_default_params(gitm_vert_params *p);

/* Create a handle bound to HIP device `device` (the current device when
 * device < 0).  Work is queued on the handle's stream (the null stream
 * until gitm_vert_set_stream). */
int gitm_vert_create(gitm_vert_handle **out, int device);
int gitm_vert_destroy(gitm_vert_handle *h);
int gitm_vert_set_stream(gitm_vert_handle *h, void *hip_stream);

/* Advance every column by n_steps vertical steps (each one a full
 * advance_vertical gather/solve/scatter) with `solver`.  All pointers in
 * `c` are device pointers.  Asynchronous: work is queued on the handle's
 * stream; argument errors are returned at once, errors found while
 * stepping are reported by gitm_vert_get_error. */
int gitm_vert_advance(gitm_vert_handle *h, const gitm_vert_params *p,
                      const gitm_vert_columns *c, int solver, int n_steps);

/* Same, for host arrays (the Fortran drop-in): copies the fields in,
 * advances, copies the written fields back and synchronises.  Returns the
 * stepping status (as gitm_vert_get_error would). */
int gitm_vert_advance_host(gitm_vert_handle *h, const gitm_vert_params *p,
                           const gitm_vert_columns *c, int solver, int n_steps);

/* gitm_vert_advance for callers that cannot allocate device memory (code
 * under hipGraph capture, serving loops): no handle, no hipMalloc/hipFree,
 * no library allocation.  It has no counterpart in the reference.  Its
 * scratch (the error word and the RK4 carries that gitm_vert_advance keeps in
 * its handle) is a fixed 64 MiB device array that every device the library
 * runs on holds until the process ends.  Runs that need more carries
 * return GITM_VERT_ESCRATCH before touching the device: with ion advection
 * more than 8,630 columns at nAlts = 50 or 4,481 at nAlts = 100; without,
 * only columns of nAlts > 188 need carries, at most 963 of them at nAlts =
 * 252.  Runs on the current device, asynchronously on
 * `hip_stream` (NULL: the null stream); the kernels are those of
 * gitm_vert_advance, so results are identical.  Calls that overlap on one
 * device share the scratch and race: serialise them (one stream).
 * err_word: NULL, or one device word that receives ~0 when no column failed,
 * else (column << 32 | code << 16 | (level + 1024)) of the first failure. */
int gitm_vert_advance_noalloc(const gitm_vert_params *p, const gitm_vert_columns *c,
                              int solver, int n_steps, void *hip_stream,
                              unsigned long long *err_word);

/* Synchronise the handle's stream and return (and clear) the first error
 * recorded by the kernels since the last call. */
int gitm_vert_get_error(gitm_vert_handle *h, gitm_vert_error *err);

/* Tile choice: waves per column (1..4) for the next launches; 0 restores
 * the automatic choice (smallest that holds nAlts+4 levels). */
int gitm_vert_set_tile(gitm_vert_handle *h, int waves_per_column);

const char *gitm_vert_strerror(int code);

/* ------------------------------------------------------------------------
 * Column prep and the vertical time step: the per-step inputs the vertical
 * step reads, computed on the device from the state, so that a caller can
 * keep the state in device memory across steps.
 * ------------------------------------------------------------------------ */
typedef struct gitm_vert_prep_params {
  double mass[GITM_VERT_NSPECIES];      /* Mass(1:6) [kg]                     */
  double vibration[GITM_VERT_NSPECIES]; /* Vibration(1:6)  ModEarth.f90:307-312 */
  double cfl;                /* CFL                  ModInputs.f90:123      */
  double fixed_dt;           /* FixedDt, the step before the CFL limits     */
  double dt_end;             /* EndTime - CurrentTime                        */
  double eddy_coef;          /* EddyDiffusionCoef    ModInputs.f90:302       */
  double eddy_p0, eddy_p1;   /* EddyDiffusionPressure0/1           :303-304  */
  int32_t use_ion_advection; /* adds the ion CFL limit (calc_timestep.f90:125-136) */
  int32_t reserved;
} gitm_vert_prep_params;

/* Fill `p` with the reference defaults. */
void gitm_vert_default_prep_params(gitm_vert_prep_params *p);

/* For every column of `c` (device pointers), from NDensityS, Temperature,
 * Rho, VerticalVelocity, dAlt (and IVelocity with use_ion_advection):
 *   NDensity = sum of NDensityS, Pressure = Temperature*Rho, cp and Gamma
 *                                          (calc_pressure.f90:17-23,56-86)
 *   MeanMajorMass, TempUnit                (calc_rates.Earth.f90:34-50)
 *   KappaEddyDiffusion from Pressure       (Earth.f90:376-411)
 *   cMax(iUp_) on levels 0..nAlts+1        (calc_timestep.f90:113-118)
 * written to c->ndensity, pressure (optional), cp (optional),
 * c->gamma, c->mean_major_mass, c->temp_unit, c->eddy and c->cmax, all on
 * levels -1..nAlts+2 except cmax.  *dt_device (one device double) receives
 * min(fixed_dt, dt_end, CFL / max over 1..nAlts of cMax/dAlt over every
 * column) (calc_timestep.f90:85-145; the MPI_AllREDUCE MIN over ranks at
 * :144 stays with the caller).  Asynchronous on the handle's stream.
 * Replaces: calc_pressure (neutral part), calc_rates' mean mass,
 * calc_eddy_diffusion_coefficient and calc_timestep_vertical for the
 * columns given. */
int gitm_vert_prepare(gitm_vert_handle *h, const gitm_vert_prep_params *p,
                      const gitm_vert_columns *c, const gitm_vert_field *pressure,
                      const gitm_vert_field *cp, double *dt_device);

/* ------------------------------------------------------------------------
 * Batches: many ensemble members, each a whole GITM grid of its own
 * (n_lons, n_lats, n_blocks, n_alts and arrays), advanced together.  The
 * reference runs one model per MPI job and has no counterpart; each member
 * is stepped exactly as gitm_vert_advance would step it alone.
 *
 * A batch step is one launch over every member's columns, one lane per
 * level (see gitm_vert_batch_tile).  gitm_vert_batch_create builds the
 * device member table and reserves the scratch memory; afterwards
 * gitm_vert_batch_advance and gitm_vert_batch_prepare only queue kernels on
 * the handle's stream (no allocation, no copy, no synchronisation), so a
 * multi-step loop of them can be captured in a hipGraph.  The member arrays
 * must stay valid (and at the same addresses) for the life of the batch.
 * ------------------------------------------------------------------------ */
typedef struct gitm_vert_batch gitm_vert_batch;

/* Batch of `n_members` members (device pointers, as for gitm_vert_advance).
 * use_ion_advection reserves the ion carries, so that steps with ion
 * advection can run. */
int gitm_vert_batch_create(gitm_vert_handle *h, const gitm_vert_columns *members,
                           int n_members, int use_ion_advection, gitm_vert_batch **out);
int gitm_vert_batch_destroy(gitm_vert_batch *b);

/* Tile of a batch step.  Workgroups have waves_per_workgroup wavefronts
 * (1..4, at least the tallest member's (nAlts+4)/64 rounded up).  A column of
 * more than 64 levels takes a whole workgroup; columns of at most 64 levels
 * take one wavefront each, columns_per_workgroup (1..waves_per_workgroup) of
 * them per workgroup.  carries_in_lds: the RK4 carries in LDS (1), in
 * device memory (0), or (2) in LDS for the members whose columns then still
 * fit the CU as often as the kernel's registers allow (the short columns
 * decide together, for a workgroup's slots).  persistent: 1 = as many
 * workgroups as the GPU holds at
 * once, each taking every gridDim-th task (valid when every column of more
 * than 64 levels needs all waves_per_workgroup wavefronts); 0 = one
 * workgroup per task.  The tile changes the speed, never a result bit. */
typedef struct gitm_vert_batch_tile {
  int32_t waves_per_workgroup;
  int32_t columns_per_workgroup;
  int32_t carries_in_lds;
  int32_t persistent;
} gitm_vert_batch_tile;

/* Sets the tile (NULL: the automatic one, made at creation).  EINVAL for a
 * tile the members do not fit (or whose LDS exceeds a CU's).  May grow the
 * scratch memory (synchronises the handle's stream). */
int gitm_vert_batch_set_tile(gitm_vert_batch *b, const gitm_vert_batch_tile *t);
int gitm_vert_batch_get_tile(const gitm_vert_batch *b, gitm_vert_batch_tile *t);

/* Advance every member by n_steps steps.  dt_device: NULL to use p->dt for
 * all members, or n_members device doubles, member m stepping with
 * dt_device[m] (e.g. from gitm_vert_batch_prepare).  Errors as for
 * gitm_vert_advance, reported through gitm_vert_get_error of the handle with
 * the column numbered across the batch (members in the caller's order); see
 * gitm_vert_batch_locate. */
int gitm_vert_batch_advance(gitm_vert_batch *b, const gitm_vert_params *p, int solver,
                            int n_steps, const double *dt_device);

/* gitm_vert_prepare for every member (no pressure or cp outputs);
 * dt_device[m] receives member m's step. */
int gitm_vert_batch_prepare(gitm_vert_batch *b, const gitm_vert_prep_params *p,
                            double *dt_device);

/* ------------------------------------------------------------------------
 * Adjoint of the AUSM step (reverse-mode differentiation).  The reference
 * has no counterpart; it serves sensitivity studies (e.g. of the upper
 * thermosphere's density to the temperature profile and to heating), driven
 * step by step from checkpoints by the caller (gitm_amd.adjoint in Python).
 * ------------------------------------------------------------------------ */
typedef struct gitm_vert_adjoint_fields {
  gitm_vert_field ndensity_s;        /* d/d NDensityS(.., 1:6, ..)           */
  gitm_vert_field vertical_velocity; /* d/d VerticalVelocity(.., 1:6, ..)    */
  gitm_vert_field velocity;          /* d/d Velocity(.., 1:3, ..)            */
  gitm_vert_field temperature;       /* d/d Temperature                      */
  gitm_vert_field rho;               /* d/d Rho                              */
} gitm_vert_adjoint_fields;

/* Adjoints of the column-prep outputs that the vertical step reads
 * (gitm_vert_prepare writes the values): TempUnit, Gamma, MeanMajorMass,
 * cMax (levels 0..nAlts+1, described as c->cmax) and KappaEddyDiffusion. */
typedef struct gitm_vert_adjoint_prep_fields {
  gitm_vert_field temp_unit, gamma, mean_major_mass, cmax, eddy;
} gitm_vert_adjoint_prep_fields;

/* The adjoint of one AUSM vertical step of every column of `c` (the step's
 * input state, device pointers, not modified): the adjoints `gout` of the
 * step's outputs (NDensityS, VerticalVelocity, Velocity, Temperature and Rho
 * as advance_vertical scatters them, levels -1..nAlts+2) are carried back,
 * and `gin` receives the adjoints of the step's inputs (the same fields
 * before the step).  gin may be gout (in place).  Hand-written: each column
 * recomputes its step, recording every stage's intermediates in the
 * handle's device memory, then reverses the stages' phases on the device
 * (vertical_solver_ausm.f90:26-999,1093-2117, set_vertical_bcs.Earth.f90,
 * calc_neutral_friction.f90:47-106 and the gather/scatter of
 * advance_vertical.f90:84-185).  The column-prep inputs (TempUnit, Gamma,
 * MeanMajorMass, cMax, eddy coefficient) and the geometry are held fixed;
 * gitm_vert_adjoint_all also returns their adjoints.  Neutrals only
 * (use_ion_advection and use_msis must be 0), nAlts <= 188.  Asynchronous
 * on the handle's stream; errors as for gitm_vert_advance. */
int gitm_vert_adjoint(gitm_vert_handle *h, const gitm_vert_params *p,
                      const gitm_vert_columns *c, const gitm_vert_adjoint_fields *gout,
                      const gitm_vert_adjoint_fields *gin);

/* gitm_vert_adjoint, and `gprep` receives the adjoints of the step's
 * column-prep inputs (overwritten, every level). */
int gitm_vert_adjoint_all(gitm_vert_handle *h, const gitm_vert_params *p,
                          const gitm_vert_columns *c, const gitm_vert_adjoint_fields *gout,
                          const gitm_vert_adjoint_fields *gin,
                          const gitm_vert_adjoint_prep_fields *gprep);

/* The adjoint of gitm_vert_prepare (calc_pressure.f90:56-86,
 * calc_rates.Earth.f90:34-50, Earth.f90:376-411, calc_timestep.f90:113-118):
 * the adjoints of its outputs (`gprep`, and of cp in `gcp`, which may be
 * NULL) are carried to NDensityS, Temperature, Rho and VerticalVelocity of
 * `c`'s state and ADDED to `gin` (its velocity field is not used).  The time
 * step (a MIN over columns) is held fixed.  Asynchronous. */
int gitm_vert_prepare_adjoint(gitm_vert_handle *h, const gitm_vert_prep_params *p,
                              const gitm_vert_columns *c,
                              const gitm_vert_adjoint_prep_fields *gprep,
                              const gitm_vert_field *gcp, const gitm_vert_adjoint_fields *gin);

/* The adjoint of a heating source applied after the step, Temperature
 * (1:nAlts) += Dt * Q / (cp * TempUnit) (the form of add_sources.f90:53-66
 * for a heating rate Q), for n_cols contiguous columns of n_alts + 4 levels
 * ([column][level], level -1 first).  From the adjoint `g_temp` of the
 * temperature after the source, with a = g_temp * (Dt / (cp * TempUnit)) on
 * levels 1..nAlts and 0 on the others: g_heat += a, and when g_temp_unit and
 * g_cp are given (both or neither), g_temp_unit -= a * Q / TempUnit and
 * g_cp = -(a * Q / cp) (overwritten, every level).  The temperature's own
 * adjoint passes through unchanged.  The reference has no adjoint; this is
 * the differentiable window's.  Asynchronous. */
int gitm_vert_source_adjoint(gitm_vert_handle *h, long long n_cols, int n_alts, double dt,
                             const double *g_temp, const double *cp, const double *temp_unit,
                             const double *heat, double *g_heat, double *g_temp_unit,
                             double *g_cp);

/* The same adjoint as gitm_vert_adjoint, taken by recomputing the step with
 * a taped number type (every FP64 operation a node of a tape in device
 * memory) and sweeping the tape back with atomic adds: slow (about 170x the
 * forward step) and memory-hungry (about 22 MB per 54-level column), kept
 * as an independent check of the hand-written adjoint.  nAlts <= 124.
 * Synchronous: returns the status (GITM_VERT_ETAPE if a column outgrew its
 * tape). */
int gitm_vert_adjoint_taped(gitm_vert_handle *h, const gitm_vert_params *p,
                            const gitm_vert_columns *c, const gitm_vert_adjoint_fields *gout,
                            const gitm_vert_adjoint_fields *gin);

/* Device memory the adjoints may use (0: the defaults; the hand-written
 * adjoint's record: 10% of the device's memory up to 4 GiB, 271 KB per
 * column at nAlts <= 60; the taped adjoint's tape: 40% up to 96 GiB).
 * Larger grids run in chunks of columns.  A buffer already larger than the
 * new budget is freed. */
int gitm_vert_set_adjoint_memory(gitm_vert_handle *h, size_t bytes);

/* Frees the adjoints' device buffers (they are otherwise kept, and reused,
 * until gitm_vert_destroy). */
int gitm_vert_release_adjoint_memory(gitm_vert_handle *h);

/* ------------------------------------------------------------------------
 * The operator-split sources step's neutral terms (add_sources.f90:53-70,
 * 94-107): thermal conduction (calc_thermal_conduction,
 * calc_conduction.f90:19-125) and viscosity (calc_viscosity.f90:4-189), each
 * three implicit solves of calc_conduction_1d (calc_conduction.f90:132-244)
 * combined to second order, then (with apply) the updates of Temperature,
 * Velocity and VerticalVelocity.  The coefficients are the caller's:
 * KappaTemp from calc_rates, ViscCoef and ViscCoefS from
 * calc_viscosity_coef, cp from calc_pressure (or gitm_vert_prepare).  Ion
 * drag, gravity-wave acceleration and the friction of the sources step are
 * other physics; the caller adds them (add_sources.f90:96-98,110-112).
 * ------------------------------------------------------------------------ */
typedef struct gitm_vert_sources_params {
  double dt;                          /* Dt                                  */
  double mass[GITM_VERT_NSPECIES];    /* Mass(1:6) [kg]                      */
  int32_t use_turbulent_cond;         /* UseTurbulentCond                    */
  int32_t apply;                      /* also update the state               */
} gitm_vert_sources_params;

typedef struct gitm_vert_sources_fields {
  /* inputs, read on levels 0..nAlts+1 (heating terms on 1..nAlts) */
  gitm_vert_field kappa_temp;     /* KappaTemp                               */
  gitm_vert_field visc_coef;      /* ViscCoef                                */
  gitm_vert_field visc_coef_s;    /* ViscCoefS(.., 1:6)                      */
  gitm_vert_field cp;             /* cp                                      */
  gitm_vert_field radial_distance; /* RadialDistance_GB                      */
  gitm_vert_field heating;        /* the bracket add_sources multiplies by Dt
                                     (:54-64), in Temperature units per s;
                                     NULL: none                             */
  gitm_vert_field chem_heating;   /* ChemicalHeatingRate, or NULL           */
  gitm_vert_field user_heating;   /* UserHeatingRate, or NULL               */
  /* outputs on levels 0..nAlts+1; NULL to skip */
  gitm_vert_field conduction;     /* Conduction                              */
  gitm_vert_field viscosity;      /* Viscosity(.., 1:3): the up component is
                                     what calc_viscosity leaves in it (the
                                     last species' full-step solve, :161),
                                     as add_sources adds it                 */
  gitm_vert_field vertical_viscosity; /* VerticalViscosityS(.., 1:6)        */
} gitm_vert_sources_fields;

/* From c's Temperature, TempUnit, Rho, KappaEddyDiffusion, NDensityS,
 * Velocity, VerticalVelocity and Altitude (device pointers) and the fields
 * of `f`.
 * Asynchronous on the handle's stream. */
int gitm_vert_sources(gitm_vert_handle *h, const gitm_vert_sources_params *p,
                      const gitm_vert_columns *c, const gitm_vert_sources_fields *f);

/* The same on host arrays (GITM's own): staged on the device over the
 * levels the step touches and copied back.  Synchronous. */
int gitm_vert_sources_host(gitm_vert_handle *h, const gitm_vert_sources_params *p,
                           const gitm_vert_columns *c, const gitm_vert_sources_fields *f);

/* ------------------------------------------------------------------------
 * The horizontal step: advance_horizontal(iBlock) (advance_horizontal.f90:
 * 58-924) for every block, each altitude 1..nAlts: the RK4 (:154-345) of
 * horizontal_solver (:535-818), MC-limited face values
 * (calc_facevalues_lons/lats, :928-1034) and Rusanov fluxes in longitude
 * and latitude, for Rho, Temperature, Velocity, NDensityS(1:6),
 * VerticalVelocity and, with UseIonAdvection, IDensityS(1:2) (floored at
 * 1e2, electrons re-summed; with UseImprovedIonAdvection the divergence
 * of the ion velocity follows, :432-518).  The halo cells (-1:0, n+1:n+2)
 * are read, never written (the electron re-sum of UseImprovedIonAdvection
 * aside, which covers every cell, :512-516): the caller fills them (exchange_messages_sphere,
 * advance_horizontal.f90:20) and, off the full sphere, applies
 * set_horizontal_bcs.  A negative neutral density after the step (the
 * reference stops, :358-383) is reported as GITM_VERT_ENONPOSITIVE through
 * gitm_vert_get_error.
 * ------------------------------------------------------------------------ */
typedef struct gitm_horiz_params {
  double dt;                          /* Dt                                  */
  double omega_body;                  /* OmegaBodyInput                      */
  double beta_limiter;                /* BetaLimiter                         */
  int32_t use_coriolis;               /* UseCoriolis                         */
  int32_t use_ion_advection;          /* UseIonAdvection                     */
  int32_t use_improved_ion_advection; /* UseImprovedIonAdvection: with ions,
                                         the divergence of IVelocityPar
                                         (:432-518, UAM_Gradient) after the
                                         step; needs the grad_* fields     */
  int32_t use_topography;             /* UseTopography: every horizontal
                                         gradient less dVarDAlt times the
                                         level surface's slope (:581-662,
                                         :870-871, :919-920); needs d_alt,
                                         dalt_dlon, dalt_dlat.  dVarDAlt
                                         reads level k-1 after the step has
                                         written it (:349-429), so the
                                         levels advance in order, one launch
                                         each                              */
  int32_t altitudes_per_workgroup;    /* 0: chosen by the library; 1 with
                                         use_topography                    */
  int32_t reserved;
} gitm_horiz_params;

typedef struct gitm_horiz_fields {
  /* (-1:nLons+2,-1:nLats+2,-1:nAlts+2,nBlocks) geometry (initialize.f90:
   * 198-267), descriptors as the state's */
  gitm_vert_field inv_dlon_gb;     /* InvDLonDist_GB                         */
  gitm_vert_field inv_dlon_fb;     /* InvDLonDist_FB                         */
  gitm_vert_field dlon_fb;         /* dLonDist_FB                            */
  gitm_vert_field inv_dlat_gb;     /* InvDLatDist_GB                         */
  gitm_vert_field inv_dlat_fb;     /* InvDLatDist_FB                         */
  gitm_vert_field dlat_fb;         /* dLatDist_FB                            */
  gitm_vert_field radial_distance; /* RadialDistance_GB                      */
  /* (-1:nLats+2,nBlocksMax), descriptors as cos_latitude's */
  gitm_vert_field latitude;        /* Latitude                               */
  gitm_vert_field tan_latitude;    /* TanLatitude                            */
  /* cMax_GDB(0:nLons+1,0:nLats+1,0:nAlts+1,3,nBlocks): ptr at
   * (1,1,0,iEast_,1), s_comp to iNorth_, first_level 0 */
  gitm_vert_field cmax_horizontal;
  /* output, interior cells of levels 1..nAlts; ptr may be NULL */
  gitm_vert_field temp_source;     /* HorizontalTempSource                   */
  /* UseImprovedIonAdvection only (else ignored): the ion velocity of the
   * divergence, IVelocityPar(..,3,nBlocks) on Earth (IVelocity where
   * NonMagnetic), described as ivelocity; UAM_Gradient's coefficients
   * GradLon{M,0,P}_CB(nLons,nBlocksMax), ptr at (1,1), s_lon 1,
   * s_blk nLons, and GradLat{M,0,P}_CB(nLats,nBlocksMax), ptr at (1,1),
   * s_lat 1, s_blk nLats (initialize.f90:268-292) */
  gitm_vert_field ivelocity_par;
  gitm_vert_field grad_lon_m, grad_lon_0, grad_lon_p;
  gitm_vert_field grad_lat_m, grad_lat_0, grad_lat_p;
  /* use_topography only (else ignored), interior cells of levels 1..nAlts
   * read: dAlt_GB (described as the state), dAltDLon_CB and dAltDLat_CB
   * ((nLons,nLats,nAlts,nBlocks): ptr at (1,1,1,1), first_level 1,
   * initialize.f90:294-300) */
  gitm_vert_field d_alt;
  gitm_vert_field dalt_dlon, dalt_dlat;
} gitm_horiz_fields;

/* Rho, Temperature, Velocity, NDensityS, VerticalVelocity, Gamma,
 * InvRadialDistance_GB, CosLatitude and (with ions) IDensityS, IVelocity
 * come from c (device pointers; c->n_lons * c->n_lats <= 256 cells per
 * block).  Asynchronous on the handle's stream. */
int gitm_horiz_advance(gitm_vert_handle *h, const gitm_horiz_params *p,
                       const gitm_vert_columns *c, const gitm_horiz_fields *f);

/* The same step on host arrays (GITM's own module arrays): every field is
 * staged on the device over the index ranges the step touches, halo cells
 * and ghost levels included, and the written fields (the state, IDensityS,
 * temp_source) are copied back.  Synchronous; returns the step's status as
 * gitm_vert_get_error does. */
int gitm_horiz_advance_host(gitm_vert_handle *h, const gitm_horiz_params *p,
                            const gitm_vert_columns *c, const gitm_horiz_fields *f);

/* ------------------------------------------------------------------------
 * The message exchange around the horizontal step
 * (exchange_messages_sphere.f90:4-95, the block-to-block transfer of
 * ModSphereInterface.f90:245-513) for blocks that share one device.  The
 * blocks form a regular grid, n_blocks_lon wide and n_blocks_lat high, block
 * b at longitude index b % n_blocks_lon and latitude index b / n_blocks_lon
 * (both increasing eastward and northward; n_lons * n_blocks_lon even).
 * Every halo cell (-1:0, n+1:n+2, corners included) of levels 1..nAlts
 * takes the value of the global-grid cell it overlaps: longitudes wrap;
 * past a pole the row is mirrored and the longitude turned by 180 degrees,
 * as the pole packing of srcSphereAB/AB_XFER_util.f90:40-80 does.  Carried:
 * Rho, Temperature, iTemperature, eTemperature, Velocity, VerticalVelocity,
 * NDensityS(1:6) and, with use_ion_advection, IDensityS(1:10) and the ion
 * velocity.  Then the east and north Velocity and ion velocity change sign
 * in the halo rows past a pole, on every level (:30-72), and Temperature <
 * 100 (GITM_VERT_ELOWTEMP) or Rho < 0 (GITM_VERT_ENONPOSITIVE) on the
 * blocks' cells is reported through gitm_vert_get_error (:74-84).
 * ------------------------------------------------------------------------ */
typedef struct gitm_horiz_exchange_params {
  int32_t n_blocks_lon, n_blocks_lat;
  int32_t use_ion_advection;          /* UseIonAdvection                     */
  int32_t reserved;
} gitm_horiz_exchange_params;

typedef struct gitm_horiz_exchange_fields {
  gitm_vert_field i_temperature;   /* iTemperature, described as rho         */
  gitm_vert_field e_temperature;   /* eTemperature                           */
  gitm_vert_field ion_velocity;    /* IVelocityPar (IVelocity where
                                      NonMagnetic), as ivelocity; may be NULL
                                      without ion advection                 */
  gitm_vert_field latitude;        /* Latitude, as cos_latitude             */
} gitm_horiz_exchange_fields;

/* Rho, Temperature, Velocity, VerticalVelocity, NDensityS and (with ions)
 * IDensityS come from c (device pointers).  Asynchronous on the handle's
 * stream. */
int gitm_horiz_exchange(gitm_vert_handle *h, const gitm_horiz_exchange_params *p,
                        const gitm_vert_columns *c, const gitm_horiz_exchange_fields *f);

/* Member and member column of a batch-wide column number. */
int gitm_vert_batch_locate(const gitm_vert_batch *b, int column, int *member,
                           int *member_column);

#ifdef __cplusplus
}
#endif
#endif
