/*
 * gitm_vert_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of GITM's per-column vertical dynamics step, used as the
 * parity checker for the HIP implementation.  Nothing in the product library
 * (gitm_amd/) links, loads or calls this code; only tests/, the smoke check and
 * the CPU-baseline leg of bench.py do.
 *
 * Data layout ("column-contiguous", shared with the fixtures): every per-level
 * field of one column is stored with the level index fastest, levels
 * -1..nAlts+2 at offsets 0..L-1 (L = nAlts+4); multi-component fields are
 * [col][component][level].
 */
#ifndef GITM_VERT_ORACLE_H
#define GITM_VERT_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define OV_NSPECIES 6  /* O, O2, N2, N(4S), NO, He   (ModEarth.f90:12-18)   */
#define OV_NIONS 10    /* O+ NO+ O2+ N2+ N+ O(2D)+ O(2P)+ H+ He+ e-
                          (ModEarth.f90:28-38)                           */
#define OV_NIONS_ADVECT 2 /* ModEarth.f90:39                              */
#define OV_LMAX 260    /* nAlts <= 256                                    */

enum { OV_SOLVER_RUSANOV = 0, OV_SOLVER_AUSM = 1 };

enum {
  OV_OK = 0,
  OV_ERR_NONPOSITIVE_DENSITY = 1, /* advance_vertical.f90:96-100          */
  OV_ERR_SINGULAR_MATRIX = 2,     /* ludcmp.f:14-17                       */
  OV_ERR_NEGATIVE_ION_TOP = 3,    /* set_vertical_bcs.Earth.f90:617-620   */
  OV_ERR_BAD_ARGS = 4
};

typedef struct {
  int n_alts;
  double dt;                  /* Dt                                        */
  double omega_body;          /* OmegaBodyInput       ModInputs.f90:285    */
  double max_vert_vel;        /* MaximumVerticalVelocity ModInputs.f90:224 */
  double beta_limiter;        /* BetaLimiter          ModInputs.f90:104    */
  int use_neutral_friction;   /* ModInputs.f90:240                         */
  int use_coriolis;           /* ModInputs.f90:235                         */
  int use_ion_advection;      /* ModInputs.f90:311                         */
  int use_improved_ion_advection; /* ModInputs.f90:259                     */
  int use_nighttime_ion_bcs;  /* ModInputs.f90:260                         */
  double min_tec;             /* ModInputs.f90:261                         */
  double vert_tau;            /* VertTau(:) (init_planet, ModEarth.f90:325) */
  int use_msis;               /* ModInputs.f90:190: level-0 ghosts from msis */
  int use_gswm_tides;         /* ModInputs.f90:195                         */
  int use_waccm_tides;        /* ModInputs.f90:196                         */
  int use_damping;            /* ModInputs.f90:221 (Rusanov VertTau)       */
} ov_params;

typedef struct {
  int n_cols;
  int n_alts;
  /* state, in/out */
  double *nds;   /* NDensityS        [col][6][L]                       */
  double *vvel;  /* VerticalVelocity [col][6][L]                       */
  double *vel;   /* Velocity (E,N,U) [col][3][L]                       */
  double *temp;  /* Temperature (normalised by TempUnit) [col][L]      */
  double *rho;   /* Rho              [col][L]                          */
  /* outputs */
  double *logns; /* LogNS            [col][6][L]                       */
  double *nden;  /* NDensity         [col][L]                          */
  /* column inputs (prep) */
  const double *tempunit, *mmm, *gamma, *cmax, *eddy; /* [col][L]      */
  const double *alt, *grav, *invr, *dalt;             /* [col][L]      */
  const double *coslat;                               /* [col]         */
  /* ions: may be NULL when use_ion_advection == 0 */
  double *ions;        /* IDensityS incl. electrons [col][10][L]       */
  const double *ivel;  /* IVelocity                 [col][3][L]        */
  const double *sza;   /* SZA                       [col]              */
  const double *mlat;  /* MLatitude(:,:,nAlts,:)    [col]              */
  /* UseMSIS: msis_bcs' level-0 values [col][9]: LogNS(0,1:6), Temp(0) in
   * K, east and north wind */
  const double *msis;
  /* tides: TidesEast, TidesNorth, TidesTemp (K) at levels -1, 0
   * [col][3][2] */
  const double *tides;
} ov_columns;

typedef struct {
  int code;
  int col;
  int level;
} ov_status;

/* Advance every column by n_steps vertical steps with the chosen solver.
 * Returns OV_OK or the first error (status filled). */
int ov_advance(const ov_params *p, const ov_columns *c, int solver,
               int n_steps, ov_status *status);

/* Column-parallel variant (POSIX threads), for the CPU baseline. */
int ov_advance_threaded(const ov_params *p, const ov_columns *c, int solver,
                        int n_steps, int n_threads, ov_status *status);

/* Building blocks exposed for unit tests. */
double ov_limiter_mc(double beta, double d_up, double d_down);
int ov_lu_solve6(double a[6][6], double b[6]); /* ludcmp + lubksb */
void ov_facevalues_4th(int n_alts, const double *dalt_f, const double *inv_dalt_f,
                       double beta, const double *var, double *left,
                       double *right);
void ov_kt_facevalues(int n_alts, const double *dalt_f, const double *inv_dalt_f,
                      double beta, const double *var, double *left_m12,
                      double *right_m12, double *left_p12, double *right_p12);
void ov_species_masses(double mass[OV_NSPECIES]);
void ov_glc_weights(double h1, double h2, double h3, double h4, double w[5]);
void ov_bc_coefs(double h1, double h2, double h3, double h4, int backward,
                 double C[5]);

/* Column prep and the vertical time step (calc_pressure, calc_rates' mean
 * mass, calc_eddy_diffusion_coefficient, calc_timestep_vertical). */
typedef struct {
  double mass[OV_NSPECIES], vibration[OV_NSPECIES];
  double cfl, fixed_dt, dt_end, eddy_coef, eddy_p0, eddy_p1;
  int use_ion_advection;
} ov_prep_params;

/* Per-level inputs in the column-contiguous layout ([col][comp][level],
 * levels -1..nAlts+2): nds (6 comps), temp (normalised), rho, vvel (6),
 * dalt, ivel (3; read only with use_ion_advection).  Outputs (one comp):
 * nden, press, cp, gamma, mmm, tempunit, eddy on every level, cmax on levels
 * 0..nAlts+1 (others untouched).  Returns the time step
 * min(fixed_dt, dt_end, CFL limits of every column). */
double ov_prepare(const ov_prep_params *p, int n_cols, int n_alts, const double *nds,
                  const double *temp, const double *rho, const double *vvel,
                  const double *dalt, const double *ivel, double *nden, double *press,
                  double *cp, double *gamma, double *mmm, double *tempunit,
                  double *eddy, double *cmax);


/* The operator-split sources step's neutral terms (add_sources.f90:53-66,
 * 94-107): thermal conduction (calc_conduction.f90:19-125), viscosity
 * (calc_viscosity.f90:4-189), both through calc_conduction_1d
 * (calc_conduction.f90:132-244), then the updates of Temperature, Velocity
 * and VerticalVelocity.  Layout as ov_columns ([col][comp][level], levels
 * -1..nAlts+2); the coefficient arrays are read on levels 0..nAlts+1. */
typedef struct {
  int n_cols, n_alts;
  double dt;
  int use_turbulent_cond;     /* ModInputs UseTurbulentCond                 */
  int apply;                  /* update the state (add_sources) too         */
  double mass[OV_NSPECIES];
  /* state (updated when apply) */
  double *temp;               /* Temperature (normalised)                  */
  double *vel;                /* Velocity [3]                              */
  double *vvel;               /* VerticalVelocity [6]                      */
  /* inputs */
  const double *tempunit, *rho, *cp, *eddy;
  const double *nds;          /* NDensityS [6]                             */
  const double *kappa_temp;   /* KappaTemp                                 */
  const double *visc_coef;    /* ViscCoef                                  */
  const double *visc_coef_s;  /* ViscCoefS [6]                             */
  const double *alt;          /* Altitude_GB                               */
  const double *raddist;      /* RadialDistance_GB                         */
  const double *heating;      /* the Dt-multiplied heating bracket of
                                 add_sources.f90:54-65 (1/s), or NULL      */
  const double *chem_heating; /* ChemicalHeatingRate, or NULL              */
  const double *user_heating; /* UserHeatingRate, or NULL                  */
  /* outputs, levels 0..nAlts+1 (NULL to skip) */
  double *conduction;         /* Conduction                                */
  double *viscosity;          /* Viscosity [3]; the up component is what
                                 calc_viscosity leaves there              */
  double *vert_viscosity;     /* VerticalViscosityS [6]                     */
} ov_sources;

void ov_sources_step(const ov_sources *s);

/* calc_conduction_1d for one column: quantity on levels -1..nAlts+2 (only
 * 0..nAlts+1 read), diff and mulfac on 0..nAlts+1 (arrays of L with level
 * -1 at index 0), alt and raddist on -1..nAlts+2; result on 0..nAlts+1. */
void ov_conduction_1d(int n_alts, double dt, int neumann, const double *quantity,
                      const double *diff, const double *mulfac, const double *alt,
                      const double *raddist, double *result);

/* The horizontal step (advance_horizontal(iBlock), advance_horizontal.f90:
 * 58-924) over every block and altitude 1..nAlts.  Block layout: state and
 * geometry fields [block][component][level][lat][lon] with level
 * -1..nAlts+2, lat -1..nLats+2, lon -1..nLons+2; cmax [block][3][0..nAlts+1]
 * [0..nLats+1][0..nLons+1] (components east, north, up); latitude, tan_lat
 * and cos_lat [block][-1..nLats+2].  Halo cells are read, never written.
 * Returns OV_OK, OV_ERR_NONPOSITIVE_DENSITY for a negative neutral density
 * after the step (the reference stops, :358-383) or OV_ERR_BAD_ARGS. */
typedef struct {
  int n_lons, n_lats, n_alts, n_blocks;
  double dt, omega_body, beta_limiter;
  int use_coriolis, use_ion_advection;
  /* state, updated on interior cells of levels 1..nAlts */
  double *rho, *temp, *vel, *vvel, *nds; /* vel [3], vvel and nds [6]     */
  double *ions;                          /* IDensityS [10], ion advection */
  /* inputs */
  const double *gamma, *ivel, *invr, *raddist; /* ivel [3]                 */
  const double *inv_dlon_gb, *inv_dlon_fb, *dlon_fb;
  const double *inv_dlat_gb, *inv_dlat_fb, *dlat_fb;
  const double *latitude, *tan_lat, *cos_lat;
  const double *cmax;
  double *temp_source; /* HorizontalTempSource (interior), or NULL        */
  /* UseImprovedIonAdvection (with ions): the divergence of the ion
   * velocity (advance_horizontal.f90:432-518) through UAM_Gradient
   * (gradient.f90:1-86) */
  int use_improved_ion_advection;
  const double *ivel_par;            /* IVelocityPar [3] (Earth: magnetic)  */
  const double *grad_lon;            /* GradLonM/0/P_CB [block][3][nLons]  */
  const double *grad_lat;            /* GradLatM/0/P_CB [block][3][nLats]  */
  /* UseTopography: every horizontal gradient less the altitude derivative
   * (0.5/dAlt)(Var(k+1) - Var(k-1)) times dAltDLon or dAltDLat
   * (advance_horizontal.f90:581-662, 870-871, 919-920).  Var(k-1) is the
   * value the step already gave level k-1 (the reference writes each level
   * back before the next, :349-429), so the levels advance in order. */
  int use_topography;
  const double *dalt;      /* dAlt_GB, as a state field                     */
  const double *dalt_dlon; /* dAltDLon_CB, as a state field (interior read) */
  const double *dalt_dlat; /* dAltDLat_CB, likewise                         */
} oh_grid;

int oh_advance_horizontal(const oh_grid *p, ov_status *status);

/* The message exchange before and after the horizontal step
 * (exchange_messages_sphere.f90:4-95) on one node: every block of a regular
 * grid of n_blocks_lon x n_blocks_lat blocks (block b at longitude
 * b % n_blocks_lon, latitude b / n_blocks_lon, both increasing) gets its
 * halo cells (-1:0, n+1:n+2, corners included) on levels 1..nAlts from the
 * cells of the global grid they overlap: longitudes wrap; past a pole the
 * row is mirrored and the longitude turned by 180 degrees (the pole packing
 * of srcSphereAB/AB_XFER_util.f90:40-80).  Rho, Temperature,
 * iTemperature, eTemperature, Velocity(3), VerticalVelocity(6),
 * NDensityS(1:6) and, with ions, IDensityS(1:10) and the ion velocity
 * (IVelocityPar on Earth).  Then, in the halo rows past a pole, the east and
 * north components of Velocity and of the ion velocity change sign on every
 * level (exchange_messages_sphere.f90:30-72), and Temperature < 100 or
 * Rho < 0 on the blocks' cells is an error (:74-84). */
typedef struct {
  int n_lons, n_lats, n_alts, n_blocks_lon, n_blocks_lat;
  int use_ion_advection;
  double *rho, *temp, *itemp, *etemp, *vel, *vvel, *nds, *ions, *ivel;
  const double *latitude; /* [block][-1..nLats+2] */
} oh_exchange_args;

enum { OV_ERR_LOW_TEMPERATURE = 5 };

int oh_exchange(const oh_exchange_args *a, ov_status *status);

#ifdef __cplusplus
}
#endif
#endif
