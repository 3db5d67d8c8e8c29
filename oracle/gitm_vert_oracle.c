/*
 * gitm_vert_oracle.c -- TEST INFRASTRUCTURE ONLY (the parity checker).
 *
 * Plain-C restatement of GITM's per-column vertical dynamics step:
 *   advance_vertical            src/advance_vertical.f90:45-219
 *   advance_vertical_1d_ausm    src/vertical_solver_ausm.f90:26-433
 *   advance_vertical_1stage_ausm                         :436-999
 *   calc_rusanov_alts_ausm / calc_facevalues_alts_ausm   :1007-1090
 *   calc_all_fluxes_hydro                                :1093-2070
 *   calc_kt_facevalues                                   :2073-2117
 *   advance_vertical_1d_rusanov src/vertical_solver_rusanov.f90:10-384
 *   advance_vertical_1stage (Rusanov)                    :387-686
 *   Limiter_mc                  src/ModLimiter.f90:16-36
 *   set_vertical_bcs            src/set_vertical_bcs.Earth.f90:8-652
 *   calc_neutral_friction       src/calc_neutral_friction.f90:4-108
 *   LU decomposition / back-substitution (Crout, implicit partial pivoting;
 *   textbook algorithm, same pivot rule as src/ludcmp.f / src/lubksb.f).
 *
 * Every floating-point expression keeps the reference's evaluation order
 * (Fortran evaluates left to right; x**2.0 is x*x and x**3.0 is x*x*x, as
 * amdflang emits them).
 * Code that cannot change any output of the step (the sponge, VelocityCap,
 * HydroNT/HydroPressure/HydroRho, the interim NewLogRho update, Grad/Diff of
 * LogPress/LogRho/LogNum, electron ghost values) is omitted.
 *
 * The boundary conditions are restated with UseMSIS (from the caller's
 * level-0 msis_bcs values), the GSWM and WACCM tides (from the caller's
 * TidesEast/North/Temp) and UseDamping for the Rusanov step; the user's BC
 * perturbation is not (the product library rejects it in the same way).
 */
#include "gitm_vert_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define NSP OV_NSPECIES
#define NIO OV_NIONS
#define LM OV_LMAX
#define X(k) ((k) + 1) /* level k (-1..nAlts+2) -> array offset */

static const double kB = 1.3807e-23;          /* ModConstants.f90:15 */
static const double AMU_ = 1.6726e-27;        /* ModConstants.f90:11 */
static const double PI_ = 3.141592653589793;  /* ModConstants.f90:23 */

enum { iO_ = 0, iO2_ = 1, iN2_ = 2, iN4S_ = 3, iNO_ = 4, iHe_ = 5 };
enum { iE_ = 0, iN_ = 1, iU_ = 2 }; /* iEast_, iNorth_, iUp_ (ModGITM.f90) */

/* Diff0 and DiffExp, ModEarth.f90:237-261 (symmetric tables). */
static const double DIFF0_RAW[NSP][NSP] = {
    {0.000, 0.969, 0.969, 0.969, 0.715, 3.440},
    {0.969, 0.000, 0.715, 0.969, 0.715, 3.208},
    {0.969, 0.715, 0.000, 0.969, 0.527, 2.939},
    {0.969, 0.969, 0.969, 0.000, 0.969, 2.939},
    {0.715, 0.715, 0.527, 0.969, 0.000, 2.939},
    {3.440, 3.208, 2.939, 2.939, 2.939, 0.000}};
static const double DIFFEXP[NSP][NSP] = {
    {0.000, 0.774, 0.774, 0.774, 0.750, 0.749},
    {0.774, 0.000, 0.750, 0.774, 0.750, 0.710},
    {0.774, 0.750, 0.000, 0.774, 0.810, 0.718},
    {0.774, 0.774, 0.774, 0.000, 0.774, 0.718},
    {0.750, 0.750, 0.810, 0.774, 0.000, 0.718},
    {0.749, 0.710, 0.718, 0.718, 0.718, 0.000}};

/* ThermalDiffCoefS, calc_rates.Earth.f90:83-84 */
static const double TDC[NSP] = {0.0, 0.0, 0.0, 0.0, 0.0, -0.38};

void ov_species_masses(double m[NSP]) {
  /* init_planet, ModEarth.f90:271-281 */
  double mO = 16.0 * AMU_, mN = 14.0 * AMU_;
  m[iO_] = mO;
  m[iO2_] = 2.0 * mO;
  m[iN2_] = 2.0 * mN;
  m[iN4S_] = mN;
  m[iNO_] = mN + mO;
  m[iHe_] = 4.0 * AMU_;
}

/* ------------------------------------------------------------------ */
/* Column working set (ModVertical.f90:8-50)                           */
/* ------------------------------------------------------------------ */
typedef struct {
  int nA;
  double altG[LM], gravG[LM], invr[LM], dAltC[LM], dAltF[LM], invDAltF[LM];
  double mmm[LM], gam[LM], eddy[LM], cmax[LM];
  double effg[LM];  /* EffectiveGravity, vertical_solver_ausm.f90:582-589 */
  double raddist[LM];
  double centrifugal, coriolis, sza, mlat;
  double ivel[3][LM]; /* IVel: mutated in place by set_vertical_bcs */
  double kp[LM], ku[LM]; /* LiouKpS, Ku, vertical_solver_ausm.f90:1320-1352 */
  double mass[NSP];
  double msis[9]; /* UseMSIS: msis_bcs' LogNS(0,1:6), Temp(0), east, north wind */
  double tides[3][2]; /* TidesEast, TidesNorth, TidesTemp at levels -1, 0 */
  double vtau[LM];    /* VertTau(iAlt) of the Rusanov step */
} ov_col;

typedef struct {
  double logrho[LM], temp[LM];
  double logns[NSP][LM], vv[NSP][LM], vel[3][LM];
  double logins[NIO][LM];
} ov_state;

typedef struct {
  ov_state st, s1, s2, s3, s4, upd, nw, fin;
  ov_col col;
} ov_ws;

static inline double fmax2(double a, double b) { return a > b ? a : b; }
static inline double fmin2(double a, double b) { return a < b ? a : b; }

/* ------------------------------------------------------------------ */
double ov_limiter_mc(double beta, double dUp, double dDown) {
  /* ModLimiter.f90:16-36 */
  if (dUp > 0.0) {
    if (dDown > 0.0)
      return fmin2(fmin2(beta * dUp, beta * dDown), (dUp + dDown) * 0.5);
    return 0.0;
  }
  if (dDown < 0.0)
    return fmax2(fmax2(beta * dUp, beta * dDown), (dUp + dDown) * 0.5);
  return 0.0;
}

/* calc_facevalues_alts_ausm, vertical_solver_ausm.f90:1039-1090 (identical
 * to calc_facevalues_alts_rusanov, vertical_solver_rusanov.f90:726-779).
 * left/right are indexed by face 1..nAlts+1 (face i between cells i-1, i). */
void ov_facevalues_4th(int nA, const double *dF, const double *iF, double beta,
                       const double *v, double *left, double *right) {
  const double F1 = 0.6250000, F2 = 0.0416667;
  double dl[LM];
  double h, up, down;
  int i;
  for (i = 1; i <= nA; i++) {
    h = iF[X(i + 1)] * 2.0;
    up = h * (F1 * (v[X(i + 1)] - v[X(i)]) - F2 * (v[X(i + 2)] - v[X(i - 1)]));
    h = iF[X(i)] * 2.0;
    down = h * (F1 * (v[X(i)] - v[X(i - 1)]) - F2 * (v[X(i + 1)] - v[X(i - 2)]));
    dl[X(i)] = ov_limiter_mc(beta, up, down);
  }
  i = 0;
  up = (v[X(i + 1)] - v[X(i)]) * iF[X(i + 1)];
  down = (v[X(i)] - v[X(i - 1)]) * iF[X(i)];
  dl[X(i)] = ov_limiter_mc(beta, up, down);
  i = nA + 1;
  up = (v[X(i + 1)] - v[X(i)]) * iF[X(i + 1)];
  down = (v[X(i)] - v[X(i - 1)]) * iF[X(i)];
  dl[X(i)] = ov_limiter_mc(beta, up, down);
  for (i = 1; i <= nA + 1; i++) {
    left[i] = v[X(i - 1)] + 0.5 * dl[X(i - 1)] * dF[X(i)];
    right[i] = v[X(i)] - 0.5 * dl[X(i)] * dF[X(i)];
  }
}

/* calc_rusanov_alts_ausm, vertical_solver_ausm.f90:1007-1031 (twin of
 * calc_rusanov_alts_rusanov, vertical_solver_rusanov.f90:694-718).
 * grad/diff indexed by cell 1..nAlts. */
static void rusanov_alts(const ov_col *c, double beta, const double *v,
                         double *grad, double *diff) {
  int nA = c->nA, i;
  double L[LM], R[LM], DF[LM];
  ov_facevalues_4th(nA, c->dAltF, c->invDAltF, beta, v, L, R);
  for (i = 1; i <= nA; i++)
    grad[i] = 0.5 * (L[i + 1] + R[i + 1] - L[i] - R[i]) / c->dAltC[X(i)];
  for (i = 1; i <= nA + 1; i++)
    DF[i] = 0.5 * fmax2(c->cmax[X(i - 1)], c->cmax[X(i)]) * (R[i] - L[i]);
  if (diff)
    for (i = 1; i <= nA; i++) diff[i] = (DF[i + 1] - DF[i]) / c->dAltC[X(i)];
}

/* calc_kt_facevalues, vertical_solver_ausm.f90:2073-2117.
 * Outputs indexed by cell 1..nAlts. */
void ov_kt_facevalues(int nA, const double *dF, const double *iF, double beta,
                      const double *v, double *LMo, double *RMo, double *LPo,
                      double *RPo) {
  double dl[LM];
  int i;
  for (i = 0; i <= nA + 1; i++) {
    double up = (v[X(i + 1)] - v[X(i)]) * iF[X(i + 1)];
    double down = (v[X(i)] - v[X(i - 1)]) * iF[X(i)];
    dl[X(i)] = ov_limiter_mc(beta, up, down);
  }
  for (i = 1; i <= nA; i++) {
    LMo[i] = v[X(i - 1)] + 0.5 * dl[X(i - 1)] * dF[X(i - 1)];
    RMo[i] = v[X(i)] - 0.5 * dl[X(i)] * dF[X(i)];
    LPo[i] = v[X(i)] + 0.5 * dl[X(i)] * dF[X(i)];
    RPo[i] = v[X(i + 1)] - 0.5 * dl[X(i + 1)] * dF[X(i + 1)];
  }
}

/* ------------------------------------------------------------------ */
/* 6x6 LU with implicit partial pivoting + substitution.               */
/* Same pivot rule as src/ludcmp.f:1-60: rows scaled by 1/max|a_ij|,    */
/* the last row attaining the largest scaled |pivot| wins (>=), a zero  */
/* pivot becomes 1e-20; forward substitution skips leading zeros of the */
/* permuted right-hand side as src/lubksb.f:1-28 does.                  */
/* ------------------------------------------------------------------ */
int ov_lu_solve6(double a[6][6], double b[6]) {
  const int n = 6;
  double scale[6];
  int perm[6];
  int i, j, k, piv = 0;
  for (i = 0; i < n; i++) {
    double big = 0.0;
    for (j = 0; j < n; j++)
      if (fabs(a[i][j]) > big) big = fabs(a[i][j]);
    if (big == 0.0) return OV_ERR_SINGULAR_MATRIX;
    scale[i] = 1. / big;
  }
  for (j = 0; j < n; j++) {
    double best = 0.0;
    for (i = 0; i < j; i++) {
      double s = a[i][j];
      for (k = 0; k < i; k++) s = s - a[i][k] * a[k][j];
      a[i][j] = s;
    }
    for (i = j; i < n; i++) {
      double s = a[i][j];
      for (k = 0; k < j; k++) s = s - a[i][k] * a[k][j];
      a[i][j] = s;
      double t = scale[i] * fabs(s);
      if (t >= best) {
        piv = i;
        best = t;
      }
    }
    if (j != piv) {
      for (k = 0; k < n; k++) {
        double t = a[piv][k];
        a[piv][k] = a[j][k];
        a[j][k] = t;
      }
      scale[piv] = scale[j];
    }
    perm[j] = piv;
    if (a[j][j] == 0.) a[j][j] = 1.0e-20;
    if (j != n - 1) {
      double r = 1. / a[j][j];
      for (i = j + 1; i < n; i++) a[i][j] = a[i][j] * r;
    }
  }
  /* substitution */
  int first = -1;
  for (i = 0; i < n; i++) {
    int ip = perm[i];
    double s = b[ip];
    b[ip] = b[i];
    if (first >= 0) {
      for (j = first; j < i; j++) s = s - a[i][j] * b[j];
    } else if (s != 0.) {
      first = i;
    }
    b[i] = s;
  }
  for (i = n - 1; i >= 0; i--) {
    double s = b[i];
    for (j = i + 1; j < n; j++) s = s - a[i][j] * b[j];
    b[i] = s / a[i][i];
  }
  return OV_OK;
}

/* calc_neutral_friction, calc_neutral_friction.f90:4-108.
 * vel/ns/glc indexed [species][cell 1..nAlts]; nt, temp by cell. */
static int neutral_friction(const ov_col *c, double dtin, double vel[NSP][LM],
                            const double *nt, double ns[NSP][LM],
                            double glc[NSP][LM], const double *temp) {
  int i, is, js;
  for (i = 1; i <= c->nA; i++) {
    double V[NSP], C[NSP][NSP], M[NSP][NSP], eddyc[NSP];
    for (is = 0; is < NSP; is++) V[is] = vel[is][i];
    memset(C, 0, sizeof(C));
    for (is = 0; is < NSP; is++) {
      double kTOverM = kB * temp[i] / c->mass[is];
      double denscale = 1.0 / nt[i];
      for (js = 0; js < NSP; js++) {
        if (js == is) continue;
        double d0 = 1.0e17 * DIFF0_RAW[is][js];
        double TempDij =
            (1.0e-04) * (d0 * pow(temp[i], DIFFEXP[is][js])) / (nt[i] * (1.0e-06));
        C[is][js] = kTOverM * denscale * ns[js][i] / TempDij;
      }
      eddyc[is] = -1.0 * c->eddy[X(i)] * glc[is][i];
    }
    for (is = 0; is < NSP; is++)
      for (js = 0; js < NSP; js++) M[is][js] = -dtin * C[is][js];
    for (is = 0; is < NSP; is++) {
      double s = 0.0;
      for (js = 0; js < NSP; js++) s = s + C[is][js];
      M[is][is] = 1.0 + dtin * s;
    }
    int rc = ov_lu_solve6(M, V);
    if (rc) return rc;
    for (is = 0; is < NSP; is++) vel[is][i] = V[is] + eddyc[is];
  }
  return OV_OK;
}

/* Weights of the 5-point non-uniform derivative at the centre point, with
 * h1..h4 the spacings between the five points (vertical_solver_ausm.f90
 * :657-676). */
void ov_glc_weights(double h1, double h2, double h3, double h4, double w[5]) {
  double MeshH2 = h2 + h1;
  double MeshH3 = h3 + h2 + h1;
  double MeshH4 = h4 + h3 + h2 + h1;
  w[0] = (h2 * h3 * (h3 + h4)) / (h1 * MeshH2 * MeshH3 * MeshH4);
  w[1] = -1.0 * (MeshH2 * h3 * (h3 + h4)) / (h1 * h2 * (h2 + h3) * (h2 + h3 + h4));
  w[3] = MeshH2 * h2 * (h4 + h3) / (MeshH3 * (h2 + h3) * h3 * h4);
  w[4] = -1.0 * MeshH2 * h2 * h3 / (MeshH4 * (h2 + h3 + h4) * (h3 + h4) * h4);
  w[2] = (h2 * h3 * (h3 + h4) + MeshH2 * h3 * (h3 + h4) - MeshH2 * h2 * (h3 + h4) -
          MeshH2 * h2 * h3) /
         (MeshH2 * h2 * h3 * (h3 + h4));
}

/* 5-point non-uniform gradient of ln(rho_s/rho),
 * vertical_solver_ausm.f90:649-688 (vertical_solver_rusanov.f90:510-545). */
static void grad_log_cons(const ov_col *c, const ov_state *u, double ns[NSP][LM],
                          const double *rho, double glc[NSP][LM]) {
  int nA = c->nA, i, k, sp;
  double lc[NSP][LM];
  for (sp = 0; sp < NSP; sp++)
    for (k = -1; k <= nA + 2; k++)
      lc[sp][X(k)] = log(c->mass[sp] * ns[sp][X(k)] / rho[X(k)]);
  (void)u;
  for (i = 1; i <= nA; i++) {
    double C[5];
    ov_glc_weights(c->dAltF[X(i - 1)], c->dAltF[X(i)], c->dAltF[X(i + 1)],
                   c->dAltF[X(i + 2)], C);
    for (sp = 0; sp < NSP; sp++)
      glc[sp][i] = C[0] * lc[sp][X(i - 2)] + C[1] * lc[sp][X(i - 1)] +
                   C[2] * lc[sp][X(i)] + C[3] * lc[sp][X(i + 1)] +
                   C[4] * lc[sp][X(i + 2)];
  }
}

/* ------------------------------------------------------------------ */
/* set_vertical_bcs, set_vertical_bcs.Earth.f90:8-652                   */
/* ------------------------------------------------------------------ */
static void bc_fwd_coefs(const double *dF, int k, double C[5]) {
  /* :228-248 */
  double h1 = dF[X(k + 2)], h2 = dF[X(k + 3)], h3 = dF[X(k + 4)], h4 = dF[X(k + 5)];
  double H1 = h1, H2 = h1 + h2, H3 = h1 + h2 + h3, H4 = h1 + h2 + h3 + h4;
  C[0] = -1.0 * (H2 * H3 * H4 + H1 * H3 * H4 + H1 * H2 * H4 + H1 * H2 * H3) /
         (H1 * H2 * H3 * H4);
  C[1] = 1.0 * (H2 * H3 * H4) / (h1 * h2 * (h2 + h3) * (h2 + h3 + h4));
  C[2] = -1.0 * (H1 * H3 * H4) / (H2 * h2 * h3 * (h3 + h4));
  C[3] = 1.0 * (H1 * H2 * H4) / (H3 * (h3 + h2) * h3 * h4);
  C[4] = -1.0 * (H1 * H2 * H3) / (H4 * (h2 + h3 + h4) * (h3 + h4) * h4);
}

static void bc_bwd_coefs(const double *dF, int k, double C[5]) {
  /* :505-525 */
  double h1 = dF[X(k - 1)], h2 = dF[X(k - 2)], h3 = dF[X(k - 3)], h4 = dF[X(k - 4)];
  double H1 = h1, H2 = h1 + h2, H3 = h1 + h2 + h3, H4 = h1 + h2 + h3 + h4;
  C[0] = 1.0 * (H2 * H3 * H4 + H1 * H3 * H4 + H1 * H2 * H4 + H1 * H2 * H3) /
         (H1 * H2 * H3 * H4);
  C[1] = -1.0 * (H2 * H3 * H4) / (h1 * h2 * (h2 + h3) * (h2 + h3 + h4));
  C[2] = 1.0 * (H1 * H3 * H4) / (H2 * h2 * h3 * (h3 + h4));
  C[3] = -1.0 * (H1 * H2 * H4) / (H3 * (h3 + h2) * h3 * h4);
  C[4] = 1.0 * (H1 * H2 * H3) / (H4 * (h2 + h3 + h4) * (h3 + h4) * h4);
}

void ov_bc_coefs(double h1, double h2, double h3, double h4, int backward,
                 double C[5]) {
  /* forward (:228-248) and backward (:505-525) one-sided weights; the
   * point being filled is at distance h1 from the first of the five */
  double dF[LM];
  if (!backward) {
    dF[X(2)] = h1; dF[X(3)] = h2; dF[X(4)] = h3; dF[X(5)] = h4;
    bc_fwd_coefs(dF, 0, C);
  } else {
    dF[X(9)] = h1; dF[X(8)] = h2; dF[X(7)] = h3; dF[X(6)] = h4;
    bc_bwd_coefs(dF, 10, C);
  }
}

static int set_vertical_bcs(ov_col *c, const ov_params *p, ov_state *s) {
  const int nA = c->nA;
  const int photo[NSP] = {0, 0, 0, 1, 1, 0}; /* :89-91 */
  const double *dF = c->dAltF;
  double NS[NSP][LM];
  double C[5];
  int sp, k, d, is;
  double *T = s->temp;

  for (sp = 0; sp < NSP; sp++) /* :93 */
    for (k = -1; k <= nA + 2; k++) NS[sp][X(k)] = exp(s->logns[sp][X(k)]);

  if (p->use_msis) {
    /* :113-137: msis_bcs at levels -1 and 0.  The caller supplies level 0;
     * every level -1 value set here is overwritten below (Temp(-1) = Temp(0),
     * the hydrostatic or extrapolated LogNS(-1), Vel(-1) from levels 0..4,
     * LogRho(-1) from the species), so level 0's values stand in for both. */
    for (k = -1; k <= 0; k++) {
      for (sp = 0; sp < NSP; sp++) s->logns[sp][X(k)] = c->msis[sp];
      T[X(k)] = c->msis[6];
      s->vel[iE_][X(k)] = c->msis[7];
      s->vel[iN_][X(k)] = c->msis[8];
    }
  } else {
    /* Bottom, UseMsis = F: :138-143 */
    for (k = -1; k <= 0; k++) {
      s->vel[iE_][X(k)] = 0.0;
      s->vel[iN_][X(k)] = 0.0;
    }
  }
  /* :145-148 */
  for (k = -1; k <= 0; k++) {
    s->vel[iU_][X(k)] = 0.0;
    for (sp = 0; sp < NSP; sp++) s->vv[sp][X(k)] = 0.0;
  }
  /* :150-154 */
  if (p->use_gswm_tides)
    for (k = -1; k <= 0; k++) {
      s->vel[iE_][X(k)] = c->tides[0][k + 1];
      s->vel[iN_][X(k)] = c->tides[1][k + 1];
      T[X(k)] = c->tides[2][k + 1] + T[X(k)];
    }
  /* :155-159 */
  if (p->use_waccm_tides)
    for (k = -1; k <= 0; k++) {
      s->vel[iE_][X(k)] = c->tides[0][k + 1];
      s->vel[iN_][X(k)] = c->tides[1][k + 1];
      T[X(k)] = c->tides[2][k + 1];
    }
  /* :192-196 */
  T[X(-1)] = T[X(0)];
  for (d = 0; d < 3; d++) s->vel[d][X(-1)] = s->vel[d][X(0)];

  for (sp = 0; sp < NSP; sp++) {
    if (!photo[sp]) { /* :198-223 */
      k = -1;
      double MeanGravity = -0.5 * (c->effg[X(k)] + c->effg[X(k + 1)]);
      double MeanTemp = 0.5 * (T[X(k + 1)] + T[X(k)]);
      double MeanMass = 0.5 * (c->mmm[X(k + 1)] + c->mmm[X(k)]);
      double ISH = MeanGravity * MeanMass / (MeanTemp * kB);
      NS[sp][X(k)] = NS[sp][X(k + 1)] * (T[X(k + 1)] / T[X(k)]) *
                     exp(1.0 * ISH * dF[X(k)]);
      s->logns[sp][X(k)] = log(NS[sp][X(k)]);
      if (sp == iO_) s->logns[sp][X(k)] = s->logns[sp][X(0)];
    } else { /* :226-272 */
      for (k = 0; k >= -1; k--) {
        double *ln = s->logns[sp];
        bc_fwd_coefs(dF, k, C);
        double dLogNS = ln[X(k + 1)] * C[0] + ln[X(k + 2)] * C[1] +
                        ln[X(k + 3)] * C[2] + ln[X(k + 4)] * C[3] +
                        ln[X(k + 5)] * C[4];
        ln[X(k)] = ln[X(k + 1)] - dF[X(k + 1)] * dLogNS;
        if (dLogNS <= 0.0) ln[X(k)] = ln[X(k + 1)];
        if (sp == iNO_) {
          if (ln[X(k)] < ln[X(k + 1)] - 0.333) ln[X(k)] = ln[X(k + 1)] - 0.333;
        }
      }
    }
  }

  for (k = 0; k >= -1; k--) { /* :277-378 */
    bc_fwd_coefs(dF, k, C);
    if (p->use_ion_advection) {
      for (is = 0; is < NIO - 1; is++) s->logins[is][X(k)] = s->logins[is][X(k + 1)];
      for (d = 0; d < 3; d++) {
        double *iv = c->ivel[d];
        double dVV = iv[X(k + 1)] * C[0] + iv[X(k + 2)] * C[1] +
                     iv[X(k + 3)] * C[2] + iv[X(k + 4)] * C[3] +
                     iv[X(k + 5)] * C[4];
        iv[X(k)] = iv[X(k + 1)] - dF[X(k + 1)] * dVV;
      }
    }
    for (sp = 0; sp < NSP; sp++) {
      double *vv = s->vv[sp];
      double dVV = vv[X(k + 1)] * C[0] + vv[X(k + 2)] * C[1] +
                   vv[X(k + 3)] * C[2] + vv[X(k + 4)] * C[3] +
                   vv[X(k + 5)] * C[4];
      vv[X(k)] = vv[X(k + 1)] - dF[X(k + 1)] * dVV;
    }
    for (sp = 0; sp < NSP; sp++) NS[sp][X(k)] = exp(s->logns[sp][X(k)]);
    double SumRho = 0.0;
    for (sp = 0; sp < NSP; sp++) SumRho = SumRho + c->mass[sp] * NS[sp][X(k)];
    s->logrho[X(k)] = log(SumRho);
    s->vel[iU_][X(k)] = 0.0;
    for (sp = 0; sp < NSP; sp++)
      s->vel[iU_][X(k)] = s->vel[iU_][X(k)] + NS[sp][X(k)] * c->mass[sp] *
                                                  s->vv[sp][X(k)] /
                                                  exp(s->logrho[X(k)]);
    if (k == -1 && p->use_msis) { /* :364-375 */
      for (d = iE_; d <= iN_; d++) {
        double *v = s->vel[d];
        double dVV = v[X(k + 1)] * C[0] + v[X(k + 2)] * C[1] + v[X(k + 3)] * C[2] +
                     v[X(k + 4)] * C[3] + v[X(k + 5)] * C[4];
        v[X(k)] = v[X(k + 1)] - dF[X(k + 1)] * dVV;
      }
    }
  }

  for (sp = 0; sp < NSP; sp++) { /* :382-395 */
    if (photo[sp] && s->vv[sp][X(1)] > 0.0 && sp != iO_) {
      s->vv[sp][X(0)] = -1.0 * s->vv[sp][X(1)];
      s->vv[sp][X(-1)] = -1.0 * s->vv[sp][X(1)];
    }
  }

  /* Top: :420-486 */
  for (k = nA + 1; k <= nA + 2; k++) {
    s->vel[iE_][X(k)] = s->vel[iE_][X(nA)];
    s->vel[iN_][X(k)] = s->vel[iN_][X(nA)];
  }
  if (p->use_ion_advection) {
    for (k = nA + 1; k <= nA + 2; k++) {
      c->ivel[iE_][X(k)] = c->ivel[iE_][X(nA)];
      c->ivel[iN_][X(k)] = c->ivel[iN_][X(nA)];
    }
  }
  for (sp = 0; sp < NSP; sp++) {
    double *vv = s->vv[sp];
    vv[X(nA + 1)] = vv[X(nA)];
    vv[X(nA + 2)] = vv[X(nA)];
    if (photo[sp] && vv[X(nA)] < 0.0) {
      vv[X(nA + 1)] = -vv[X(nA)];
      vv[X(nA + 2)] = -vv[X(nA - 1)];
    }
  }
  if (s->vel[iU_][X(nA)] < 0.0) {
    s->vel[iU_][X(nA + 1)] = -s->vel[iU_][X(nA)];
    s->vel[iU_][X(nA + 2)] = -s->vel[iU_][X(nA)];
  } else {
    s->vel[iU_][X(nA + 1)] = s->vel[iU_][X(nA)];
    s->vel[iU_][X(nA + 2)] = s->vel[iU_][X(nA)];
  }
  if (p->use_ion_advection) {
    double *iu = c->ivel[iU_];
    if (iu[X(nA)] < 0.0) {
      iu[X(nA + 1)] = -iu[X(nA)];
      iu[X(nA + 2)] = -iu[X(nA)];
    } else {
      iu[X(nA + 1)] = iu[X(nA)];
      iu[X(nA + 2)] = iu[X(nA)];
    }
  }
  T[X(nA + 1)] = T[X(nA)];
  T[X(nA + 2)] = T[X(nA)];
  /* LogRho(nAlts+1:nAlts+2) extrapolated at :465-467 is overwritten by the
   * species sum at :598-604 before anything reads it. */
  for (sp = 0; sp < NSP; sp++) {
    for (k = nA + 1; k <= nA + 2; k++) {
      double MeanGravity = -0.5 * (c->effg[X(k)] + c->effg[X(k - 1)]);
      double MeanTemp = 0.5 * (T[X(k - 1)] + T[X(k)]);
      double ISH = MeanGravity * c->mass[sp] / (MeanTemp * kB);
      NS[sp][X(k)] = NS[sp][X(k - 1)] * (T[X(k - 1)] / T[X(k)]) *
                     exp(-ISH * dF[X(k)]);
      s->logns[sp][X(k)] = log(NS[sp][X(k)]);
    }
  }

  if (p->use_ion_advection) { /* :488-596 */
    int plasma = 0;
    if (p->use_nighttime_ion_bcs) {
      if (c->sza > PI_ / 2 && ((c->mlat > 30.0 && c->mlat < 70.0) ||
                               (c->mlat > -60.0 && c->mlat < -30.0)))
        plasma = 1;
    }
    double tec = 0.0;
    for (k = 1; k <= nA; k++) {
      double v = p->use_improved_ion_advection ? s->logins[0][X(k)]
                                               : exp(s->logins[0][X(k)]);
      tec = tec + dF[X(k)] * v;
    }
    tec = tec / 1e16;
    for (k = nA + 1; k <= nA + 2; k++) {
      bc_bwd_coefs(dF, k, C);
      for (is = 0; is < NIO - 1; is++) {
        double *li = s->logins[is];
        double n0, n1, n2, n3, n4, n5;
        if (p->use_improved_ion_advection) {
          /* check_for_negative_densities, :608-649 */
          double mn = li[X(k - 5)];
          int j;
          for (j = k - 4; j <= k; j++)
            if (li[X(j)] < mn) mn = li[X(j)];
          if (mn < 0.0) {
            if (li[X(k - 6)] < 0.0) return OV_ERR_NEGATIVE_ION_TOP;
            for (j = k - 5; j <= k - 1; j++) {
              if (li[X(j)] < 0.0) {
                if (li[X(j - 1)] > 0.0 && li[X(j + 1)] > 0.0)
                  li[X(j)] = (li[X(j - 1)] + li[X(j + 1)]) / 2.0;
                else
                  li[X(j)] = li[X(j - 1)] * 0.95;
              }
            }
            if (li[X(k)] < 0.0) li[X(k)] = li[X(k - 1)] * 0.95;
          }
          n0 = log(li[X(k)]);
          n1 = log(li[X(k - 1)]);
          n2 = log(li[X(k - 2)]);
          n3 = log(li[X(k - 3)]);
          n4 = log(li[X(k - 4)]);
          n5 = log(li[X(k - 5)]);
        } else {
          n0 = li[X(k)];
          n1 = li[X(k - 1)];
          n2 = li[X(k - 2)];
          n3 = li[X(k - 3)];
          n4 = li[X(k - 4)];
          n5 = li[X(k - 5)];
        }
        double dn = C[0] * n1 + C[1] * n2 + C[2] * n3 + C[3] * n4 + C[4] * n5;
        if (!(tec < p->min_tec && plasma)) {
          if (dn > -0.001 * n0 / dF[X(nA)]) dn = -0.001 * n0 / dF[X(nA)];
          n0 = n1 + dn * dF[X(k)];
          li[X(k)] = p->use_improved_ion_advection ? exp(n0) : n0;
        }
      }
    }
  }

  for (k = nA + 1; k <= nA + 2; k++) { /* :598-604 */
    double SumRho = 0.0;
    for (sp = 0; sp < NSP; sp++)
      SumRho = SumRho + c->mass[sp] * exp(s->logns[sp][X(k)]);
    s->logrho[X(k)] = log(SumRho);
  }
  return OV_OK;
}

/* ------------------------------------------------------------------ */
/* calc_all_fluxes_hydro, vertical_solver_ausm.f90:1093-2070            */
/* ------------------------------------------------------------------ */
typedef struct {
  double rhosflux[NSP][LM], momsflux[NSP][LM], eflux[LM], momflux[3][LM];
} ov_fluxes;

static void ausm_fluxes(const ov_col *c, double beta, double dtin,
                        double rhos[NSP][LM], double ps[NSP][LM],
                        double hps[NSP][LM], const ov_state *u,
                        int sub[NSP][LM], ov_fluxes *F) {
  const int nA = c->nA;
  const double *dF = c->dAltF, *iF = c->invDAltF;
  const double MInf = 1.0e-19, LiouBeta = 1.0 / 8.0; /* :1312-1313 */
  int i, sp, d;
  static __thread double lg[LM];
  static __thread double RSLm[NSP][LM], RSRm[NSP][LM], RSLp[NSP][LM], RSRp[NSP][LM];
  static __thread double HLm[NSP][LM], HRm[NSP][LM], HLp[NSP][LM], HRp[NSP][LM];
  static __thread double PSLm[NSP][LM], PSRm[NSP][LM], PSLp[NSP][LM], PSRp[NSP][LM];
  static __thread double VLm[NSP][LM], VRm[NSP][LM], VLp[NSP][LM], VRp[NSP][LM];
  static __thread double GLm[3][LM], GRm[3][LM], GLp[3][LM], GRp[3][LM];
  double RLm[LM], RRm[LM], RLp[LM], RRp[LM];
  double PLm[LM], PRm[LM], PLp[LM], PRp[LM];
  double gLm[LM], gRm[LM], gLp[LM], gRp[LM];

  for (sp = 0; sp < NSP; sp++) { /* :1356-1436 */
    int k;
    for (k = -1; k <= nA + 2; k++) lg[X(k)] = log(rhos[sp][X(k)]);
    ov_kt_facevalues(nA, dF, iF, beta, lg, RSLm[sp], RSRm[sp], RSLp[sp], RSRp[sp]);
    for (k = -1; k <= nA + 2; k++) lg[X(k)] = log(hps[sp][X(k)]);
    ov_kt_facevalues(nA, dF, iF, beta, lg, HLm[sp], HRm[sp], HLp[sp], HRp[sp]);
    for (k = -1; k <= nA + 2; k++) lg[X(k)] = log(ps[sp][X(k)]);
    ov_kt_facevalues(nA, dF, iF, beta, lg, PSLm[sp], PSRm[sp], PSLp[sp], PSRp[sp]);
    for (i = 1; i <= nA; i++) {
      RSLm[sp][i] = exp(RSLm[sp][i]);
      RSRm[sp][i] = exp(RSRm[sp][i]);
      RSLp[sp][i] = exp(RSLp[sp][i]);
      RSRp[sp][i] = exp(RSRp[sp][i]);
      HLm[sp][i] = exp(HLm[sp][i]);
      HRm[sp][i] = exp(HRm[sp][i]);
      HLp[sp][i] = exp(HLp[sp][i]);
      HRp[sp][i] = exp(HRp[sp][i]);
      PSLm[sp][i] = exp(PSLm[sp][i]);
      PSRm[sp][i] = exp(PSRm[sp][i]);
      PSLp[sp][i] = exp(PSLp[sp][i]);
      PSRp[sp][i] = exp(PSRp[sp][i]);
    }
    ov_kt_facevalues(nA, dF, iF, beta, u->vv[sp], VLm[sp], VRm[sp], VLp[sp], VRp[sp]);
  }

  for (i = 1; i <= nA; i++) { /* :1459-1493 */
    RLm[i] = RRm[i] = RLp[i] = RRp[i] = 0.0;
    PLm[i] = PRm[i] = PLp[i] = PRp[i] = 0.0;
    for (sp = 0; sp < NSP; sp++) {
      RLm[i] = RLm[i] + RSLm[sp][i];
      RRm[i] = RRm[i] + RSRm[sp][i];
      RLp[i] = RLp[i] + RSLp[sp][i];
      RRp[i] = RRp[i] + RSRp[sp][i];
    }
    for (sp = 0; sp < NSP; sp++) {
      PLm[i] = PLm[i] + PSLm[sp][i];
      PRm[i] = PRm[i] + PSRm[sp][i];
      PLp[i] = PLp[i] + PSLp[sp][i];
      PRp[i] = PRp[i] + PSRp[sp][i];
    }
  }
  for (d = 0; d < 3; d++) /* :1496-1503 */
    ov_kt_facevalues(nA, dF, iF, beta, u->vel[d], GLm[d], GRm[d], GLp[d], GRp[d]);
  for (i = 1; i <= nA; i++) { /* :1505-1532 */
    GLm[iU_][i] = GRm[iU_][i] = GLp[iU_][i] = GRp[iU_][i] = 0.0;
    for (sp = 0; sp < NSP; sp++) {
      GLm[iU_][i] = GLm[iU_][i] + RSLm[sp][i] * VLm[sp][i] / RLm[i];
      GRm[iU_][i] = GRm[iU_][i] + RSRm[sp][i] * VRm[sp][i] / RRm[i];
      GLp[iU_][i] = GLp[iU_][i] + RSLp[sp][i] * VLp[sp][i] / RLp[i];
      GRp[iU_][i] = GRp[iU_][i] + RSRp[sp][i] * VRp[sp][i] / RRp[i];
    }
  }
  ov_kt_facevalues(nA, dF, iF, beta, c->gam, gLm, gRm, gLp, gRp); /* :1535 */

  for (i = 1; i <= nA; i++) {
    /* :1538-1562 bulk energies */
#define SQ(x) ((x) * (x))
#define EN(G, P, R, V)                                                     \
  ((1.0 / (G[i] - 1.0)) * P[i] +                                            \
   0.5 * R[i] * (SQ(V[iU_][i]) + SQ(V[iE_][i]) + SQ(V[iN_][i])))
#define HE(G, P, R, V)                                                     \
  (0.5 * (SQ(V[iU_][i]) + SQ(V[iE_][i]) + SQ(V[iN_][i])) +                  \
   (G[i] / (G[i] - 1.0)) * P[i] / R[i])
    double ELm = EN(gLm, PLm, RLm, GLm), ERm = EN(gRm, PRm, RRm, GRm);
    double ELp = EN(gLp, PLp, RLp, GLp), ERp = EN(gRp, PRp, RRp, GRp);
    /* :1569-1615 enthalpies, :1619-1655 Liou sound speed */
    double HLm_ = HE(gLm, PLm, RLm, GLm), HRm_ = HE(gRm, PRm, RRm, GRm);
    double HLp_ = HE(gLp, PLp, RLp, GLp), HRp_ = HE(gRp, PRp, RRp, GRp);
    double SubCs;
    SubCs = sqrt(2.0 * ((gLm[i] - 1.0) / (gLm[i] + 1.0)) * HLm_);
    double CSLm = SQ(SubCs) / fmax2(SubCs, GLm[iU_][i]);
    SubCs = sqrt(2.0 * ((gRm[i] - 1.0) / (gRm[i] + 1.0)) * HRm_);
    double CSRm = SQ(SubCs) / fmax2(SubCs, -1.0 * GRm[iU_][i]);
    double CSm = fmin2(CSLm, CSRm);
    SubCs = sqrt(2.0 * ((gLp[i] - 1.0) / (gLp[i] + 1.0)) * HLp_);
    double CSLp = SQ(SubCs) / fmax2(SubCs, GLp[iU_][i]);
    SubCs = sqrt(2.0 * ((gRp[i] - 1.0) / (gRp[i] + 1.0)) * HRp_);
    double CSRp = SQ(SubCs) / fmax2(SubCs, -1.0 * GRp[iU_][i]);
    double CSp = fmin2(CSLp, CSRp);
#undef EN
#undef HE

    double NVm[NSP], NVp[NSP], NPm[NSP], NPp[NSP];
    double FRm[NSP], FRp[NSP], FMm[NSP], FMp[NSP];
    for (sp = 0; sp < NSP; sp++) {
      /* :1659-1680 means */
      double MPSp = 0.5 * (PSLp[sp][i] + PSRp[sp][i]);
      double MPSm = 0.5 * (PSLm[sp][i] + PSRm[sp][i]);
      double MHPSp = 0.5 * (HLp[sp][i] + HRp[sp][i]);
      double MHPSm = 0.5 * (HLm[sp][i] + HRm[sp][i]);
      double MRSp = 0.5 * (RSLp[sp][i] + RSRp[sp][i]);
      double MRSm = 0.5 * (RSLm[sp][i] + RSRm[sp][i]);
      /* :1689-1740 Mach numbers */
      double MLm = VLm[sp][i] / CSm, MRm = VRm[sp][i] / CSm;
      double MLp = VLp[sp][i] / CSp, MRp = VRp[sp][i] / CSp;
      double M2m = 0.5 * (SQ(MLm) + SQ(MRm));
      double M2p = 0.5 * (SQ(MLp) + SQ(MRp));
      double MZm = sqrt(fmin2(1.0, fmax2(M2m, MInf)));
      double MZp = sqrt(fmin2(1.0, fmax2(M2p, MInf)));
      double FAm = MZm * (2.0 - MZm), FAp = MZp * (2.0 - MZp);
      /* :1742-1819 split Mach polynomials */
      double MF4PLm, MF4NRm, MF4PLp, MF4NRp;
      if (fabs(MLm) >= 1.0) MF4PLm = 0.5 * (MLm + fabs(MLm));
      else MF4PLm = 0.25 * SQ(MLm + 1.0) * (1.0 - 16.0 * LiouBeta * (-0.25 * SQ(MLm - 1.0)));
      if (fabs(MRm) >= 1.0) MF4NRm = 0.5 * (MRm - fabs(MRm));
      else MF4NRm = -0.25 * SQ(MRm - 1.0) * (1.0 + 16.0 * LiouBeta * (0.25 * SQ(MRm + 1.0)));
      if (fabs(MLp) >= 1.0) MF4PLp = 0.5 * (MLp + fabs(MLp));
      else MF4PLp = 0.25 * SQ(MLp + 1.0) * (1.0 - 16.0 * LiouBeta * (-0.25 * SQ(MLp - 1.0)));
      if (fabs(MRp) >= 1.0) MF4NRp = 0.5 * (MRp - fabs(MRp));
      else MF4NRp = -0.25 * SQ(MRp - 1.0) * (1.0 + 16.0 * LiouBeta * (0.25 * SQ(MRp + 1.0)));
      /* :1822-1869 pressure-diffusion Mach term */
      double kp = c->kp[X(i)], ku = c->ku[X(i)], zeta = 1.0;
      double MPm, MPp;
      if (!sub[sp][X(i)]) {
        MPp = kp * fmax2((1.0 - M2p), 0.0) * (PSRp[sp][i] - PSLp[sp][i]) /
              (MRSp * CSp * (FAp * CSp + zeta * c->dAltC[X(i)] / dtin));
        MPm = kp * fmax2((1.0 - M2m), 0.0) * (PSRm[sp][i] - PSLm[sp][i]) /
              (MRSm * CSm * (FAm * CSm + zeta * c->dAltC[X(i)] / dtin));
      } else {
        MPm = kp * fmax2((1.0 - M2m), 0.0) *
              ((PSRm[sp][i] - HRm[sp][i]) - (PSLm[sp][i] - HLm[sp][i])) /
              (MRSm * CSm * (FAm * CSm + zeta * c->dAltC[X(i)] / dtin));
        MPp = kp * fmax2((1.0 - M2p), 0.0) *
              ((PSRp[sp][i] - HRp[sp][i]) - (PSLp[sp][i] - HLp[sp][i])) /
              (MRSp * CSp * (FAp * CSp + zeta * c->dAltC[X(i)] / dtin));
      }
      /* :1872-1895 numerical velocity */
      NVm[sp] = CSm * (MF4PLm + MF4NRm - MPm);
      NVp[sp] = CSp * (MF4PLp + MF4NRp - MPp);
      /* :1899-1934 numerical pressure */
      if (!sub[sp][X(i)]) {
        NPp[sp] = (MPSp) - 0.5 * ku * CSp * ((RSRp[sp][i]) * VRp[sp][i] - (RSLp[sp][i]) * VLp[sp][i]);
        NPm[sp] = (MPSm) - 0.5 * ku * CSm * ((RSRm[sp][i]) * VRm[sp][i] - (RSLm[sp][i]) * VLm[sp][i]);
      } else {
        NPp[sp] = (MPSp - MHPSp) - 0.5 * ku * CSp * ((RSRp[sp][i]) * VRp[sp][i] - (RSLp[sp][i]) * VLp[sp][i]);
        NPm[sp] = (MPSm - MHPSm) - 0.5 * ku * CSm * ((RSRm[sp][i]) * VRm[sp][i] - (RSLm[sp][i]) * VLm[sp][i]);
      }
      /* :1937-1970 upwinded species fluxes */
      if (NVp[sp] >= 0.0) {
        FRp[sp] = RSLp[sp][i] * NVp[sp];
        FMp[sp] = RSLp[sp][i] * VLp[sp][i] * NVp[sp];
      } else {
        FRp[sp] = RSRp[sp][i] * NVp[sp];
        FMp[sp] = RSRp[sp][i] * VRp[sp][i] * NVp[sp];
      }
      if (NVm[sp] >= 0.0) {
        FRm[sp] = RSLm[sp][i] * NVm[sp];
        FMm[sp] = RSLm[sp][i] * VLm[sp][i] * NVm[sp];
      } else {
        FRm[sp] = RSRm[sp][i] * NVm[sp];
        FMm[sp] = RSRm[sp][i] * VRm[sp][i] * NVm[sp];
      }
    }
    /* :1973-2026 bulk numerical velocity, energy and momentum fluxes */
    double BNVp = 0.0, BNVm = 0.0;
    for (sp = 0; sp < NSP; sp++) {
      BNVp = BNVp + (RSLp[sp][i] + RSRp[sp][i]) * NVp[sp] / (RLp[i] + RRp[i]);
      BNVm = BNVm + (RSLm[sp][i] + RSRm[sp][i]) * NVm[sp] / (RLm[i] + RRm[i]);
    }
    double EFp, EFm, MOp[3], MOm[3];
    if (BNVp >= 0.0) {
      EFp = (ELp + PLp[i]) * BNVp;
      for (d = 0; d < 3; d++) MOp[d] = RLp[i] * GLp[d][i] * BNVp;
    } else {
      EFp = (ERp + PRp[i]) * BNVp;
      for (d = 0; d < 3; d++) MOp[d] = RRp[i] * GRp[d][i] * BNVp;
    }
    if (BNVm >= 0.0) {
      EFm = (ELm + PLm[i]) * BNVm;
      for (d = 0; d < 3; d++) MOm[d] = RLm[i] * GLm[d][i] * BNVm;
    } else {
      EFm = (ERm + PRm[i]) * BNVm;
      for (d = 0; d < 3; d++) MOm[d] = RRm[i] * GRm[d][i] * BNVm;
    }
    /* :2029-2068 spherical divergence */
    double LeftRadius = 0.5 * (c->raddist[X(i)] + c->raddist[X(i - 1)]);
    double RightRadius = 0.5 * (c->raddist[X(i)] + c->raddist[X(i + 1)]);
    double AP = SQ(RightRadius), AM = SQ(LeftRadius);
    /* amdflang lowers r**3.0 (integral constant exponent) to r*r*r */
    double Vol = (1.0 / 3.0) * (RightRadius * RightRadius * RightRadius -
                                LeftRadius * LeftRadius * LeftRadius);
    for (sp = 0; sp < NSP; sp++) {
      F->rhosflux[sp][i] = ((AP)*FRp[sp] - (AM)*FRm[sp]) / Vol;
      F->momsflux[sp][i] = ((AP)*FMp[sp] - (AM)*FMm[sp]) / Vol;
      F->momsflux[sp][i] = F->momsflux[sp][i] + (NPp[sp] - NPm[sp]) / c->dAltC[X(i)];
    }
    F->eflux[i] = ((AP)*EFp - (AM)*EFm) / Vol;
    for (d = 0; d < 3; d++) F->momflux[d][i] = (AP * MOp[d] - AM * MOm[d]) / Vol;
#undef SQ
  }
}

/* ------------------------------------------------------------------ */
/* advance_vertical_1stage_ausm, vertical_solver_ausm.f90:436-999       */
/* `out` holds a copy of `u` on entry (the driver's New* = * copies).   */
/* ------------------------------------------------------------------ */
static int ausm_stage(const ov_col *c, const ov_params *p, double dtin,
                      const ov_state *u, ov_state *out) {
  const int nA = c->nA;
  const double *m = c->mass;
  int i, k, sp, is;
  static __thread double NS[NSP][LM], RhoS[NSP][LM], PS[NSP][LM], HNS[NSP][LM];
  static __thread double HPS[NSP][LM], MomS[NSP][LM], NewRhoS[NSP][LM], NewNS[NSP][LM];
  static __thread double Dev[NSP][LM], glc[NSP][LM], nvel[NSP][LM];
  static __thread int sub[NSP][LM];
  static __thread ov_fluxes F;
  double Rho[LM], NT[LM], Press[LM], TE[LM], Mom[3][LM];
  double NewRho[LM], NewNT[LM];
  double GradTemp[LM], DiffVE[LM], DiffVN[LM], gdum[LM];

  for (sp = 0; sp < NSP; sp++) /* :591 */
    for (k = -1; k <= nA + 2; k++) NS[sp][X(k)] = exp(u->logns[sp][X(k)]);
  for (k = -1; k <= nA + 2; k++) Rho[X(k)] = exp(u->logrho[X(k)]);
  for (k = -1; k <= nA + 2; k++) { /* :600-609 */
    NT[X(k)] = 0.0;
    for (sp = 0; sp < NSP; sp++) NT[X(k)] = NT[X(k)] + NS[sp][X(k)];
    Press[X(k)] = NT[X(k)] * kB * u->temp[X(k)];
  }
  /* :611-646: only GradTemp, DiffVel(E,N) and the ion terms are live */
  rusanov_alts(c, p->beta_limiter, u->temp, GradTemp, NULL);
  rusanov_alts(c, p->beta_limiter, u->vel[iE_], gdum, DiffVE);
  rusanov_alts(c, p->beta_limiter, u->vel[iN_], gdum, DiffVN);

  grad_log_cons(c, u, NS, Rho, glc); /* :649-688 */

  /* :694-751 hydrostatic background per species */
  for (sp = 0; sp < NSP; sp++) {
    for (k = -1; k <= nA + 2; k++) HNS[sp][X(k)] = NS[sp][X(k)];
    for (k = 1; k <= nA + 2; k++) {
      double MeanMass = m[sp];
      double MeanGravity = 0.5 * (c->effg[X(k)] + c->effg[X(k - 1)]);
      double MeanTemp = 0.5 * (u->temp[X(k)] + u->temp[X(k - 1)]);
      double ISH = -1.0 * MeanMass * MeanGravity / (kB * MeanTemp);
      HNS[sp][X(k)] = HNS[sp][X(k - 1)] * (u->temp[X(k - 1)] / u->temp[X(k)]) *
                      exp(-1.0 * c->dAltF[X(k)] * ISH);
    }
  }
  for (sp = 0; sp < NSP; sp++) {
    k = -1;
    double MeanMass = 0.5 * (c->mmm[X(-1)] + c->mmm[X(0)]);
    double MeanGravity = 0.5 * (c->effg[X(k + 1)] + c->effg[X(k)]);
    double MeanTemp = 0.5 * (u->temp[X(k + 1)] + u->temp[X(k)]);
    double ISH = -1.0 * MeanMass * MeanGravity / (kB * MeanTemp);
    HNS[sp][X(k)] = HNS[sp][X(k + 1)] * (u->temp[X(k + 1)] / u->temp[X(k)]) *
                    exp(c->dAltF[X(k)] * ISH);
  }
  for (k = -1; k <= nA + 2; k++) { /* :752-773 */
    for (sp = 0; sp < NSP; sp++) {
      HPS[sp][X(k)] = HNS[sp][X(k)] * kB * u->temp[X(k)];
      double hrhos = m[sp] * HNS[sp][X(k)];
      RhoS[sp][X(k)] = m[sp] * NS[sp][X(k)];
      NewRhoS[sp][X(k)] = RhoS[sp][X(k)];
      MomS[sp][X(k)] = m[sp] * NS[sp][X(k)] * u->vv[sp][X(k)];
      PS[sp][X(k)] = NS[sp][X(k)] * u->temp[X(k)] * kB;
      /* :775-792 */
      double ratio = fabs(RhoS[sp][X(k)] - hrhos) / RhoS[sp][X(k)];
      Dev[sp][X(k)] = RhoS[sp][X(k)] - hrhos;
      sub[sp][X(k)] = !(fabs(ratio) > 1.0);
    }
    TE[X(k)] = Press[X(k)] / (c->gam[X(k)] - 1.0) +
               0.5 * Rho[X(k)] *
                   (u->vel[iU_][X(k)] * u->vel[iU_][X(k)] +
                    u->vel[iN_][X(k)] * u->vel[iN_][X(k)] +
                    u->vel[iE_][X(k)] * u->vel[iE_][X(k)]);
    for (int d = 0; d < 3; d++) Mom[d][X(k)] = Rho[X(k)] * u->vel[d][X(k)];
  }

  ausm_fluxes(c, p->beta_limiter, dtin, RhoS, PS, HPS, u, sub, &F); /* :798 */

  for (i = 1; i <= nA; i++) { /* :806-829 */
    for (sp = 0; sp < NSP; sp++) {
      NewRhoS[sp][X(i)] = RhoS[sp][X(i)] - dtin * (F.rhosflux[sp][i]);
      out->logns[sp][X(i)] = log(NewRhoS[sp][X(i)] / m[sp]);
    }
  }
  if (p->use_ion_advection) {
    double GradLI[LM], DiffLI[LM];
    for (is = 0; is < OV_NIONS_ADVECT; is++) {
      rusanov_alts(c, p->beta_limiter, u->logins[is], GradLI, DiffLI);
      for (i = 1; i <= nA; i++)
        out->logins[is][X(i)] = out->logins[is][X(i)] -
                                dtin * (c->ivel[iU_][X(i)] * GradLI[i]) +
                                dtin * DiffLI[i];
    }
  }
  for (k = -1; k <= nA + 2; k++) { /* :836-847 */
    NewRho[X(k)] = 0.0;
    NewNT[X(k)] = 0.0;
    for (sp = 0; sp < NSP; sp++) {
      NewNS[sp][X(k)] = NewRhoS[sp][X(k)] / m[sp];
      NewRho[X(k)] = NewRho[X(k)] + NewRhoS[sp][X(k)];
      NewNT[X(k)] = NewNT[X(k)] + NewNS[sp][X(k)];
    }
  }
  for (i = 1; i <= nA; i++) { /* :862-900 */
    for (sp = 0; sp < NSP; sp++) {
      double nm;
      if (!sub[sp][X(i)])
        nm = MomS[sp][X(i)] - dtin * (F.momsflux[sp][i]) +
             dtin * RhoS[sp][X(i)] * c->effg[X(i)];
      else
        nm = MomS[sp][X(i)] - dtin * (F.momsflux[sp][i]) +
             dtin * Dev[sp][X(i)] * c->effg[X(i)];
      double w = nm / NewRhoS[sp][X(i)];
      w = w + dtin * (u->vel[iN_][X(i)] * u->vel[iN_][X(i)] +
                      u->vel[iE_][X(i)] * u->vel[iE_][X(i)]) *
                  c->invr[X(i)];
      if (p->use_coriolis) w = w + dtin * (c->coriolis * u->vel[iE_][X(i)]);
      w = w - dtin * (TDC[sp] * kB * GradTemp[i]) / m[sp];
      nvel[sp][i] = w;
    }
  }
  if (p->use_neutral_friction) { /* :902-913 */
    static __thread double nsn[NSP][LM];
    double ntn[LM], tt[LM];
    for (i = 1; i <= nA; i++) {
      ntn[i] = NewNT[X(i)];
      tt[i] = u->temp[X(i)];
      for (sp = 0; sp < NSP; sp++) nsn[sp][i] = NewNS[sp][X(i)];
    }
    int rc = neutral_friction(c, dtin, nvel, ntn, nsn, glc, tt);
    if (rc) return rc;
  }
  for (sp = 0; sp < NSP; sp++)
    for (i = 1; i <= nA; i++) out->vv[sp][X(i)] = nvel[sp][i];
  for (k = -1; k <= nA + 2; k++) /* :916-923 */
    for (sp = 0; sp < NSP; sp++) {
      out->vv[sp][X(k)] = fmax2(-p->max_vert_vel, out->vv[sp][X(k)]);
      out->vv[sp][X(k)] = fmin2(p->max_vert_vel, out->vv[sp][X(k)]);
    }
  for (k = -1; k <= nA + 2; k++) { /* :925-941 */
    out->vel[iU_][X(k)] = 0.0;
    for (sp = 0; sp < NSP; sp++)
      out->vel[iU_][X(k)] = out->vel[iU_][X(k)] +
                            out->vv[sp][X(k)] * NewRhoS[sp][X(k)] / NewRho[X(k)];
  }
  for (i = 1; i <= nA; i++) { /* :944-986 */
    double hf = 0.35 * exp(-1.0 * (c->altG[X(i)] - c->altG[X(0)]) / (45.0e+03));
    double nmE = Mom[iE_][X(i)] - dtin * (F.momflux[iE_][i]);
    out->vel[iE_][X(i)] = nmE / NewRho[X(i)];
    out->vel[iE_][X(i)] = out->vel[iE_][X(i)] + hf * dtin * DiffVE[i];
    double nmN = Mom[iN_][X(i)] - dtin * (F.momflux[iN_][i]);
    out->vel[iN_][X(i)] = nmN / NewRho[X(i)];
    out->vel[iN_][X(i)] = out->vel[iN_][X(i)] + hf * dtin * DiffVN[i];
    double nte = TE[X(i)] - dtin * F.eflux[i] +
                 dtin * Rho[X(i)] * u->vel[iU_][X(i)] * c->effg[X(i)];
    double vu = out->vel[iU_][X(i)], vn = out->vel[iN_][X(i)], ve = out->vel[iE_][X(i)];
    double np = (nte - 0.5 * NewRho[X(i)] * (vu * vu + vn * vn + ve * ve)) *
                (c->gam[X(i)] - 1.0);
    out->temp[X(i)] = np / (kB * NewNT[X(i)]);
  }
  for (i = 1; i <= nA; i++) { /* :994-997 */
    double s = 0.0;
    for (sp = 0; sp < NSP; sp++) s = s + m[sp] * exp(out->logns[sp][X(i)]);
    out->logrho[X(i)] = log(s);
  }
  return OV_OK;
}

/* ------------------------------------------------------------------ */
/* advance_vertical_1stage (Rusanov), vertical_solver_rusanov.f90:387-686 */
/* ------------------------------------------------------------------ */
static int rusanov_stage(const ov_col *c, const ov_params *p, const ov_state *u,
                         ov_state *out) {
  const int nA = c->nA;
  const double *m = c->mass;
  const double dt = p->dt;
  int i, k, sp, is;
  static __thread double NS[NSP][LM], glc[NSP][LM], nvel[NSP][LM], nss[NSP][LM];
  static __thread double GLNS[NSP][LM], DLNS[NSP][LM], GVV[NSP][LM], DVV[NSP][LM];
  double Rho[LM], NT[LM];
  double GT[LM], DT[LM], GV[3][LM], DV[3][LM], DivVel[LM];

  for (sp = 0; sp < NSP; sp++) /* :456-460 */
    for (k = -1; k <= nA + 2; k++) NS[sp][X(k)] = exp(u->logns[sp][X(k)]);
  for (k = -1; k <= nA + 2; k++) {
    Rho[X(k)] = exp(u->logrho[X(k)]);
    double s = 0.0;
    for (sp = 0; sp < NSP; sp++) s = s + NS[sp][X(k)];
    NT[X(k)] = exp(log(s));
  }
  rusanov_alts(c, p->beta_limiter, u->temp, GT, DT); /* :464-473 */
  for (int d = 0; d < 3; d++) rusanov_alts(c, p->beta_limiter, u->vel[d], GV[d], DV[d]);
  for (i = 1; i <= nA; i++) /* :476 */
    DivVel[i] = GV[iU_][i] + 2 * u->vel[iU_][X(i)] * c->invr[X(i)];
  for (sp = 0; sp < NSP; sp++) { /* :479-491 */
    rusanov_alts(c, p->beta_limiter, u->logns[sp], GLNS[sp], DLNS[sp]);
    rusanov_alts(c, p->beta_limiter, u->vv[sp], GVV[sp], DVV[sp]);
  }
  grad_log_cons(c, u, NS, Rho, glc); /* :500-545 */

  for (i = 1; i <= nA; i++) { /* :550-614 */
    for (sp = 0; sp < NSP; sp++) {
      double divvv = GVV[sp][i] + 2 * u->vv[sp][X(i)] * c->invr[X(i)];
      out->logns[sp][X(i)] = u->logns[sp][X(i)] -
                             dt * (divvv + u->vv[sp][X(i)] * GLNS[sp][i]) +
                             dt * DLNS[sp][i];
    }
    out->vel[iU_][X(i)] = 0.0;
    for (sp = 0; sp < NSP; sp++) {
      double vv = u->vv[sp][X(i)];
      double vn = u->vel[iN_][X(i)], ve = u->vel[iE_][X(i)];
      double w = vv -
                 dt * (vv * GVV[sp][i] - (vn * vn + ve * ve) * c->invr[X(i)] +
                       u->temp[X(i)] * GLNS[sp][i] * kB / m[sp] +
                       GT[i] * kB / m[sp] - c->gravG[X(i)]) +
                 dt * DVV[sp][i] - vv / c->vtau[X(i)];
      if (p->use_coriolis)
        w = w + dt * (c->centrifugal / c->invr[X(i)] + c->coriolis * ve);
      w = w - dt * (TDC[sp] * kB * GT[i]) / m[sp];
      nvel[sp][i] = w;
    }
  }
  if (p->use_ion_advection) {
    double GradLI[LM], DiffLI[LM];
    for (is = 0; is < OV_NIONS_ADVECT; is++) {
      rusanov_alts(c, p->beta_limiter, u->logins[is], GradLI, DiffLI);
      for (i = 1; i <= nA; i++)
        out->logins[is][X(i)] = out->logins[is][X(i)] -
                                dt * (c->ivel[iU_][X(i)] * GradLI[i]) +
                                dt * DiffLI[i];
    }
  }
  if (p->use_neutral_friction) { /* :632-641 (old densities) */
    double ntn[LM], tt[LM];
    for (i = 1; i <= nA; i++) {
      ntn[i] = NT[X(i)];
      tt[i] = u->temp[X(i)];
      for (sp = 0; sp < NSP; sp++) nss[sp][i] = NS[sp][X(i)];
    }
    int rc = neutral_friction(c, dt, nvel, ntn, nss, glc, tt);
    if (rc) return rc;
  }
  for (i = 1; i <= nA; i++) { /* :643-658 */
    for (sp = 0; sp < NSP; sp++) {
      double w = fmax2(-p->max_vert_vel, nvel[sp][i]);
      w = fmin2(p->max_vert_vel, w);
      out->vv[sp][X(i)] = w;
      out->vel[iU_][X(i)] = out->vel[iU_][X(i)] +
                            w * (m[sp] * NS[sp][X(i)] / Rho[X(i)]);
    }
  }
  for (i = 1; i <= nA; i++) { /* :660-679 */
    double vu = u->vel[iU_][X(i)];
    out->vel[iE_][X(i)] = out->vel[iE_][X(i)] - dt * vu * GV[iE_][i] + dt * DV[iE_][i];
    out->vel[iN_][X(i)] = out->vel[iN_][X(i)] - dt * vu * GV[iN_][i] + dt * DV[iN_][i];
    out->temp[X(i)] = out->temp[X(i)] -
                      dt * (vu * GT[i] + (c->gam[X(i)] - 1.0) * (u->temp[X(i)] * DivVel[i])) +
                      dt * DT[i];
  }
  for (i = 1; i <= nA; i++) { /* :680-684 */
    double s = 0.0;
    for (sp = 0; sp < NSP; sp++) s = s + m[sp] * exp(out->logns[sp][X(i)]);
    out->logrho[X(i)] = log(s);
  }
  return OV_OK;
}

/* ------------------------------------------------------------------ */
/* RK4 drivers                                                          */
/* ------------------------------------------------------------------ */

/* advance_vertical_1d_ausm, vertical_solver_ausm.f90:26-433 */
static int ausm_rk4(ov_ws *w, const ov_params *p) {
  ov_col *c = &w->col;
  const int L = c->nA + 4;
  const int nA = c->nA;
  const double Dt = p->dt;
  int rc;
  ov_state *st = &w->st, *s1 = &w->s1, *s2 = &w->s2, *s3 = &w->s3, *s4 = &w->s4;
  ov_state *upd = &w->upd, *nw = &w->nw, *fin = &w->fin;

  if ((rc = set_vertical_bcs(c, p, st))) return rc; /* :135 */
  *s1 = *st;                                         /* :146-151 */
  *nw = *st;
  if ((rc = ausm_stage(c, p, 0.5 * Dt, st, nw))) return rc; /* :160-162 */
  {
    int k, f;
    for (k = 0; k < L; k++) { /* :167-189: Stage2 = Stage1 + (New - Y) */
      s2->logrho[k] = s1->logrho[k] + (nw->logrho[k] - st->logrho[k]);
      s2->temp[k] = s1->temp[k] + (nw->temp[k] - st->temp[k]);
      for (f = 0; f < NSP; f++) {
        s2->logns[f][k] = s1->logns[f][k] + (nw->logns[f][k] - st->logns[f][k]);
        s2->vv[f][k] = s1->vv[f][k] + (nw->vv[f][k] - st->vv[f][k]);
      }
      for (f = 0; f < 3; f++) s2->vel[f][k] = s1->vel[f][k] + (nw->vel[f][k] - st->vel[f][k]);
      for (f = 0; f < NIO; f++)
        s2->logins[f][k] = s1->logins[f][k] + (nw->logins[f][k] - st->logins[f][k]);
    }
  }
  const ov_state *stages[3] = {s2, s3, s4};
  ov_state *next[3] = {s3, s4, NULL};
  const double dts[3] = {0.5 * Dt, Dt, 0.5 * Dt}; /* :211, :280, :374 */
  for (int j = 0; j < 3; j++) {
    *upd = *stages[j];
    if ((rc = set_vertical_bcs(c, p, upd))) return rc; /* :213,:283,:355 */
    *st = *upd;
    *nw = *upd;
    if ((rc = ausm_stage(c, p, dts[j], st, nw))) return rc;
    if (next[j]) {
      ov_state *sn = next[j];
      int k, f;
      for (k = 0; k < L; k++) {
        sn->logrho[k] = s1->logrho[k] + (nw->logrho[k] - st->logrho[k]);
        sn->temp[k] = s1->temp[k] + (nw->temp[k] - st->temp[k]);
        for (f = 0; f < NSP; f++) {
          sn->logns[f][k] = s1->logns[f][k] + (nw->logns[f][k] - st->logns[f][k]);
          sn->vv[f][k] = s1->vv[f][k] + (nw->vv[f][k] - st->vv[f][k]);
        }
        for (f = 0; f < 3; f++) sn->vel[f][k] = s1->vel[f][k] + (nw->vel[f][k] - st->vel[f][k]);
        for (f = 0; f < NIO; f++)
          sn->logins[f][k] = s1->logins[f][k] + (nw->logins[f][k] - st->logins[f][k]);
      }
    }
  }
  /* :381-417 final combination on 1..nAlts, ghosts from Stage1 */
  *fin = *s1;
  {
    const double third = 1.0 / 3.0;
    int k, f;
#define COMB(F) third * (-1.0 * s1->F + s2->F + 2.0 * s3->F + s4->F + (nw->F - st->F))
    for (k = X(1); k <= X(nA); k++) {
      fin->logrho[k] = COMB(logrho[k]);
      fin->temp[k] = COMB(temp[k]);
      for (f = 0; f < NSP; f++) {
        fin->logns[f][k] = COMB(logns[f][k]);
        fin->vv[f][k] = COMB(vv[f][k]);
      }
      for (f = 0; f < 3; f++) fin->vel[f][k] = COMB(vel[f][k]);
      for (f = 0; f < NIO; f++) fin->logins[f][k] = COMB(logins[f][k]);
    }
#undef COMB
  }
  if ((rc = set_vertical_bcs(c, p, fin))) return rc; /* :419 */
  *st = *fin;
  return OV_OK;
}

/* advance_vertical_1d_rusanov, vertical_solver_rusanov.f90:10-384 */
static int rusanov_rk4(ov_ws *w, const ov_params *p) {
  ov_col *c = &w->col;
  const int L = c->nA + 4;
  int rc, j, k, f;
  ov_state *st = &w->st, *orig = &w->s1, *upd = &w->upd, *nw = &w->nw, *fin = &w->fin;
  ov_state *K[4] = {&w->s2, &w->s3, &w->s4, &w->fin};

  if ((rc = set_vertical_bcs(c, p, st))) return rc; /* :108 */
  *orig = *st;
  for (j = 0; j < 4; j++) {
    if (j > 0) {
      /* Y + 0.5*K1, Y + 0.5*K2, Y + K3 (:139-169, :221-251, :293-323) */
      const ov_state *kp = K[j - 1];
      const double fac = (j < 3) ? 0.5 : 1.0;
      for (k = 0; k < L; k++) {
        if (j < 3) {
          upd->logrho[k] = orig->logrho[k] + fac * kp->logrho[k];
          upd->temp[k] = orig->temp[k] + fac * kp->temp[k];
          for (f = 0; f < NSP; f++) {
            upd->logns[f][k] = orig->logns[f][k] + fac * kp->logns[f][k];
            upd->vv[f][k] = orig->vv[f][k] + fac * kp->vv[f][k];
          }
          for (f = 0; f < 3; f++) upd->vel[f][k] = orig->vel[f][k] + fac * kp->vel[f][k];
          for (f = 0; f < NIO; f++) upd->logins[f][k] = orig->logins[f][k] + fac * kp->logins[f][k];
        } else {
          upd->logrho[k] = orig->logrho[k] + kp->logrho[k];
          upd->temp[k] = orig->temp[k] + kp->temp[k];
          for (f = 0; f < NSP; f++) {
            upd->logns[f][k] = orig->logns[f][k] + kp->logns[f][k];
            upd->vv[f][k] = orig->vv[f][k] + kp->vv[f][k];
          }
          for (f = 0; f < 3; f++) upd->vel[f][k] = orig->vel[f][k] + kp->vel[f][k];
          for (f = 0; f < NIO; f++) upd->logins[f][k] = orig->logins[f][k] + kp->logins[f][k];
        }
      }
      if ((rc = set_vertical_bcs(c, p, upd))) return rc;
      *st = *upd;
    }
    *nw = *st;
    if ((rc = rusanov_stage(c, p, st, nw))) return rc;
    ov_state *kk = K[j]; /* K = New - Y */
    for (k = 0; k < L; k++) {
      kk->logrho[k] = nw->logrho[k] - st->logrho[k];
      kk->temp[k] = nw->temp[k] - st->temp[k];
      for (f = 0; f < NSP; f++) {
        kk->logns[f][k] = nw->logns[f][k] - st->logns[f][k];
        kk->vv[f][k] = nw->vv[f][k] - st->vv[f][k];
      }
      for (f = 0; f < 3; f++) kk->vel[f][k] = nw->vel[f][k] - st->vel[f][k];
      for (f = 0; f < NIO; f++) kk->logins[f][k] = nw->logins[f][k] - st->logins[f][k];
    }
  }
  /* :345-369 Final = Y + (K1 + 2K2 + 2K3 + K4)/6 on all levels.  K4 lives in
   * `fin`, so combine into `nw`. */
  {
    const double sixth = 1.0 / 6.0;
    const ov_state *k1 = K[0], *k2 = K[1], *k3 = K[2], *k4 = K[3];
#define RK(F) orig->F + sixth * (k1->F + 2.0 * k2->F + 2.0 * k3->F + k4->F)
    for (k = 0; k < L; k++) {
      nw->logrho[k] = RK(logrho[k]);
      nw->temp[k] = RK(temp[k]);
      for (f = 0; f < NSP; f++) {
        nw->logns[f][k] = RK(logns[f][k]);
        nw->vv[f][k] = RK(vv[f][k]);
      }
      for (f = 0; f < 3; f++) nw->vel[f][k] = RK(vel[f][k]);
      for (f = 0; f < NIO; f++) nw->logins[f][k] = RK(logins[f][k]);
    }
#undef RK
  }
  (void)fin;
  if ((rc = set_vertical_bcs(c, p, nw))) return rc; /* :371 */
  *st = *nw;
  return OV_OK;
}

/* ------------------------------------------------------------------ */
/* advance_vertical (gather / dispatch / scatter), advance_vertical.f90 */
/* ------------------------------------------------------------------ */
static void static_column(ov_col *c, const ov_params *p, const ov_columns *cs,
                          int col) {
  const int nA = cs->n_alts, L = nA + 4;
  const size_t o = (size_t)col * L;
  int k;
  c->nA = nA;
  ov_species_masses(c->mass);
  for (k = 0; k < L; k++) {
    c->altG[k] = cs->alt[o + k];
    c->gravG[k] = cs->grav[o + k];
    c->invr[k] = cs->invr[o + k];
    c->dAltC[k] = cs->dalt[o + k];
    c->mmm[k] = cs->mmm[o + k];
    c->gam[k] = cs->gamma[o + k];
    c->eddy[k] = cs->eddy[o + k];
    c->cmax[k] = cs->cmax[o + k];
  }
  /* :103-104 */
  c->centrifugal = (cs->coslat[col] * p->omega_body) * (cs->coslat[col] * p->omega_body);
  c->coriolis = 2 * cs->coslat[col] * p->omega_body;
  c->sza = cs->sza ? cs->sza[col] : 0.0;
  c->mlat = cs->mlat ? cs->mlat[col] : 0.0;
  if (p->use_msis)
    for (k = 0; k < 9; k++) c->msis[k] = cs->msis[(size_t)col * 9 + k];
  if (p->use_gswm_tides || p->use_waccm_tides)
    for (k = 0; k < 6; k++) c->tides[k / 2][k % 2] = cs->tides[(size_t)col * 6 + k];
  /* VertTau: ModEarth.f90:325, or with UseDamping vertical_solver_rusanov.f90
   * :581-584 (the AUSM step never reads it) */
  for (k = 0; k < L; k++)
    c->vtau[k] = p->use_damping ? 15 - (1 - exp(-1.0 * c->altG[k] / 1000.0 / 40.0)) * 5.0
                                : p->vert_tau;
  /* :149-151 */
  for (k = 0; k <= nA + 2; k++) c->dAltF[X(k)] = c->altG[X(k)] - c->altG[X(k - 1)];
  c->dAltF[X(-1)] = c->dAltF[X(0)];
  for (k = 0; k < L; k++) c->invDAltF[k] = 1.0 / c->dAltF[k];
  for (k = 0; k < L; k++) {
    c->effg[k] = c->gravG[k] + c->centrifugal / c->invr[k]; /* ausm :582-589 */
    c->raddist[k] = 1.0 / c->invr[k];                         /* ausm :595-597 */
  }
  /* :1312-1352 Kp / Ku blends (the same for every species) */
  {
    const double LogMaxKu = log(0.50), LogMinKu = log(0.25);
    const double LogMaxKp = log(1.0e-09), LogMinKp = log(1.0e-19);
    const double KpWidth = 15.0e+03, KpAltMidPoint = 160.0e+03;
    for (k = 1; k <= nA; k++) {
      double t = (1.0 + tanh((c->altG[X(k)] - KpAltMidPoint) / KpWidth));
      double lkp = LogMinKp + 0.5 * (LogMaxKp - LogMinKp) * t;
      double lku = LogMinKu + 0.5 * (LogMaxKu - LogMinKu) * t;
      c->kp[X(k)] = exp(lkp);
      c->ku[X(k)] = exp(lku);
    }
  }
  if (cs->ivel) {
    int d;
    for (d = 0; d < 3; d++)
      for (k = 0; k < L; k++) c->ivel[d][k] = cs->ivel[((size_t)col * 3 + d) * L + k];
  }
}

static int advance_column(ov_ws *w, const ov_params *p, const ov_columns *cs,
                          int col, int solver, int n_steps, ov_status *status) {
  const int nA = cs->n_alts, L = nA + 4;
  const size_t o = (size_t)col * L;
  ov_col *c = &w->col;
  ov_state *st = &w->st;
  int k, sp, d, step, rc;

  static_column(c, p, cs, col);
  for (step = 0; step < n_steps; step++) {
    /* :96-100 */
    for (sp = 0; sp < NSP; sp++)
      for (k = 0; k < L; k++)
        if (cs->nds[((size_t)col * NSP + sp) * L + k] <= 0.0) {
          status->code = OV_ERR_NONPOSITIVE_DENSITY;
          status->col = col;
          status->level = k - 2;
          return OV_ERR_NONPOSITIVE_DENSITY;
        }
    memset(st, 0, sizeof(*st));
    for (k = 0; k < L; k++) { /* :105-116 */
      st->logrho[k] = log(cs->rho[o + k]);
      for (d = 0; d < 3; d++) st->vel[d][k] = cs->vel[((size_t)col * 3 + d) * L + k];
      st->temp[k] = cs->temp[o + k] * cs->tempunit[o + k];
      for (sp = 0; sp < NSP; sp++) {
        st->logns[sp][k] = log(cs->nds[((size_t)col * NSP + sp) * L + k]);
        st->vv[sp][k] = cs->vvel[((size_t)col * NSP + sp) * L + k];
      }
    }
    if (p->use_ion_advection) { /* :124-130 */
      for (sp = 0; sp < NIO - 1; sp++)
        for (k = 0; k < L; k++) {
          double v = cs->ions[((size_t)col * NIO + sp) * L + k];
          st->logins[sp][k] = p->use_improved_ion_advection ? v : log(v);
        }
      if (cs->ivel) /* IVel is re-gathered every call (:120-122) */
        for (d = 0; d < 3; d++)
          for (k = 0; k < L; k++) c->ivel[d][k] = cs->ivel[((size_t)col * 3 + d) * L + k];
    }
    rc = (solver == OV_SOLVER_AUSM) ? ausm_rk4(w, p) : rusanov_rk4(w, p);
    if (rc) {
      status->code = rc;
      status->col = col;
      status->level = -99;
      return rc;
    }
    for (k = 0; k < L; k++) { /* :163-185 */
      cs->rho[o + k] = exp(st->logrho[k]);
      for (d = 0; d < 3; d++) cs->vel[((size_t)col * 3 + d) * L + k] = st->vel[d][k];
      cs->temp[o + k] = st->temp[k] / cs->tempunit[o + k];
      double nsum = 0.0;
      for (sp = 0; sp < NSP; sp++) {
        size_t j = ((size_t)col * NSP + sp) * L + k;
        if (cs->logns) cs->logns[j] = st->logns[sp][k];
        cs->vvel[j] = st->vv[sp][k];
        cs->nds[j] = exp(st->logns[sp][k]);
        nsum = nsum + cs->nds[j];
      }
      if (cs->nden) cs->nden[o + k] = nsum;
    }
    if (p->use_ion_advection) { /* :187-212 */
      for (sp = 0; sp < NIO - 1; sp++) {
        for (k = 0; k < L; k++) {
          size_t j = ((size_t)col * NIO + sp) * L + k;
          if (p->use_improved_ion_advection) {
            cs->ions[j] = st->logins[sp][k];
            if (k >= X(1) && k <= X(nA) && cs->ions[j] < 1) cs->ions[j] = 1.0;
          } else {
            cs->ions[j] = exp(st->logins[sp][k]);
          }
        }
      }
      for (k = 0; k < L; k++) {
        double e = 0.0;
        for (sp = 0; sp < NIO - 1; sp++) e = e + cs->ions[((size_t)col * NIO + sp) * L + k];
        cs->ions[((size_t)col * NIO + (NIO - 1)) * L + k] = e;
      }
    }
  }
  return OV_OK;
}

static int check_args(const ov_params *p, const ov_columns *c) {
  if (!p || !c) return OV_ERR_BAD_ARGS;
  if (c->n_alts < 6 || c->n_alts + 4 > LM || p->n_alts != c->n_alts) return OV_ERR_BAD_ARGS;
  if (p->use_ion_advection && (!c->ions || !c->ivel)) return OV_ERR_BAD_ARGS;
  return OV_OK;
}

int ov_advance(const ov_params *p, const ov_columns *c, int solver, int n_steps,
               ov_status *status) {
  ov_status dummy;
  int col, rc;
  if (!status) status = &dummy;
  status->code = OV_OK;
  status->col = -1;
  status->level = 0;
  if ((rc = check_args(p, c))) return status->code = rc;
  ov_ws *w = (ov_ws *)malloc(sizeof(ov_ws));
  if (!w) return status->code = OV_ERR_BAD_ARGS;
  for (col = 0; col < c->n_cols; col++) {
    rc = advance_column(w, p, c, col, solver, n_steps, status);
    if (rc) break;
  }
  free(w);
  return rc;
}

typedef struct {
  const ov_params *p;
  const ov_columns *c;
  int solver, n_steps, c0, c1, rc;
  ov_status st;
} ov_job;

static void *ov_thread(void *arg) {
  ov_job *j = (ov_job *)arg;
  ov_ws *w = (ov_ws *)malloc(sizeof(ov_ws));
  j->rc = OV_OK;
  j->st.code = OV_OK;
  if (!w) {
    j->rc = OV_ERR_BAD_ARGS;
    return NULL;
  }
  for (int col = j->c0; col < j->c1; col++) {
    j->rc = advance_column(w, j->p, j->c, col, j->solver, j->n_steps, &j->st);
    if (j->rc) break;
  }
  free(w);
  return NULL;
}

int ov_advance_threaded(const ov_params *p, const ov_columns *c, int solver,
                        int n_steps, int n_threads, ov_status *status) {
  ov_status dummy;
  int rc, t;
  if (!status) status = &dummy;
  status->code = OV_OK;
  status->col = -1;
  status->level = 0;
  if ((rc = check_args(p, c))) return status->code = rc;
  if (n_threads < 1) n_threads = 1;
  if (n_threads > c->n_cols) n_threads = c->n_cols > 0 ? c->n_cols : 1;
  ov_job *jobs = (ov_job *)calloc((size_t)n_threads, sizeof(ov_job));
  pthread_t *th = (pthread_t *)calloc((size_t)n_threads, sizeof(pthread_t));
  if (!jobs || !th) {
    free(jobs);
    free(th);
    return status->code = OV_ERR_BAD_ARGS;
  }
  for (t = 0; t < n_threads; t++) {
    jobs[t].p = p;
    jobs[t].c = c;
    jobs[t].solver = solver;
    jobs[t].n_steps = n_steps;
    jobs[t].c0 = (int)((long)c->n_cols * t / n_threads);
    jobs[t].c1 = (int)((long)c->n_cols * (t + 1) / n_threads);
    pthread_create(&th[t], NULL, ov_thread, &jobs[t]);
  }
  rc = OV_OK;
  for (t = 0; t < n_threads; t++) {
    pthread_join(th[t], NULL);
    if (jobs[t].rc && rc == OV_OK) {
      rc = jobs[t].rc;
      *status = jobs[t].st;
    }
  }
  free(jobs);
  free(th);
  return rc;
}


/* ------------------------------------------------------------------ */
/* Column prep and the vertical time step                              */
/* ------------------------------------------------------------------ */
/* Bit for bit with the reference build's calc_pressure, calc_rates,
   calc_eddy_diffusion_coefficient and calc_timestep_vertical
   (oracle/ref_prep_harness.f90, tests/test_prep_reference.py). */
double ov_prepare(const ov_prep_params *p, int n_cols, int n_alts, const double *nds,
                  const double *temp, const double *rho, const double *vvel,
                  const double *dalt, const double *ivel, double *nden, double *press,
                  double *cp, double *gamma, double *mmm, double *tempunit,
                  double *eddy, double *cmax) {
  const int L = n_alts + 4;
  double dt_local = 1.0e32; /* calc_timestep.f90:105 */
  for (int col = 0; col < n_cols; col++) {
    const size_t o = (size_t)col * L;
    for (int k = 0; k < L; k++) {
      double n[OV_NSPECIES];
      for (int s = 0; s < OV_NSPECIES; s++) n[s] = nds[((size_t)col * OV_NSPECIES + s) * L + k];
      /* calc_pressure.f90:17-23 */
      double N = 0.0;
      for (int s = 0; s < OV_NSPECIES; s++) N = N + n[s];
      nden[o + k] = N;
      const double P = temp[o + k] * rho[o + k];
      press[o + k] = P;
      /* calc_pressure.f90:58-83 */
      double c = 0.0, g = 0.0;
      for (int s = 0; s < OV_NSPECIES; s++) {
        c = c + (p->vibration[s] - 2) * n[s] * (kB / p->mass[s]);
        g = g + n[s] / (p->vibration[s] - 2);
      }
      cp[o + k] = c / (2.0 * N);
      g = g * 2.0 / N + 1;
      gamma[o + k] = g;
      /* calc_rates.Earth.f90:34-50 */
      double m = 0.0;
      for (int s = 0; s < OV_NSPECIES; s++) m = m + p->mass[s] * n[s] / (N + 1.0);
      if (m == 0) m = p->mass[0];
      mmm[o + k] = m;
      tempunit[o + k] = m / kB;
      /* Earth.f90:395-405 */
      double e = 0.0;
      if (P > p->eddy_p0)
        e = p->eddy_coef;
      else if (P > p->eddy_p1)
        e = p->eddy_coef * (P - p->eddy_p1) / (p->eddy_p0 - p->eddy_p1);
      eddy[o + k] = e;
    }
    /* calc_timestep.f90:113-136 */
    double rmax = -1.0, imax = -1.0;
    for (int k = 0; k <= n_alts + 1; k++) {
      double w = 0.0;
      for (int s = 0; s < OV_NSPECIES; s++) {
        const double a = fabs(vvel[((size_t)col * OV_NSPECIES + s) * L + X(k)]);
        if (a > w) w = a;
      }
      const double c = w + sqrt(gamma[o + X(k)] * temp[o + X(k)]);
      cmax[o + X(k)] = c;
      if (k >= 1 && k <= n_alts) {
        const double r = c / dalt[o + X(k)];
        if (r > rmax) rmax = r;
        if (p->use_ion_advection) {
          const double cm = fabs(ivel[((size_t)col * 3 + 2) * L + X(k)]) +
                            sqrt(gamma[o + X(k)] * temp[o + X(k)]);
          const double ri = cm / dalt[o + X(k)];
          if (ri > imax) imax = ri;
        }
      }
    }
    dt_local = fmin2(dt_local, p->cfl / rmax);
    if (p->use_ion_advection) dt_local = fmin2(dt_local, p->cfl / imax);
  }
  dt_local = fmin2(dt_local, p->dt_end); /* :142 */
  return fmin2(p->fixed_dt, dt_local);   /* main.f90:62, calc_timestep.f90:147 */
}


/* ------------------------------------------------------------------ */
/* Sources step: conduction and viscosity (calc_conduction.f90,        */
/* calc_viscosity.f90) and the neutral updates of add_sources.f90      */
/* ------------------------------------------------------------------ */
void ov_conduction_1d(int nA, double DtIn, int NeuBCS, const double *Quantity,
                      const double *Diff, const double *MulFac, const double *Alt,
                      const double *RadDist, double *dTdt) {
  /* calc_conduction.f90:132-244; arrays indexed by level + 1 (X()) */
  double r2[LM], di[LM], m[LM], du[LM], dl[LM], r[LM], du12[LM], du22[LM];
  double a[LM], b[LM], c[LM], d[LM], cp[LM], dp[LM], temp[LM], told[LM];
  int i;
  if (nA < 1 || nA + 4 > LM) return;
  for (i = 0; i <= nA + 1; i++) {
    told[X(i)] = Quantity[X(i)];
    r2[X(i)] = RadDist[X(i)] * RadDist[X(i)];
    di[X(i)] = Diff[X(i)] * r2[X(i)];
    m[X(i)] = DtIn / (MulFac[X(i)] * r2[X(i)]);
    du[X(i)] = Alt[X(i + 1)] - Alt[X(i)];
    dl[X(i)] = Alt[X(i)] - Alt[X(i - 1)];
    r[X(i)] = du[X(i)] / dl[X(i)];
    du12[X(i)] = du[X(i)] * du[X(i)] * (1 + r[X(i)]) * (1 + r[X(i)]);
    du22[X(i)] = 0.5 * (dl[X(i)] * du[X(i)] + du[X(i)] * du[X(i)]);
  }
  /* :188-194 (dl reused) */
  for (i = 1; i <= nA; i++)
    dl[X(i)] = di[X(i + 1)] - r[X(i)] * r[X(i)] * di[X(i - 1)] -
               (1 - r[X(i)] * r[X(i)]) * di[X(i)];
  dl[X(0)] = dl[X(1)];
  dl[X(nA + 1)] = dl[X(nA)];
  for (i = 0; i <= nA + 1; i++) { /* :198-206 */
    a[X(i)] = di[X(i)] / du22[X(i)] * r[X(i)] - dl[X(i)] / du12[X(i)] * r[X(i)] * r[X(i)];
    c[X(i)] = di[X(i)] / du22[X(i)] + dl[X(i)] / du12[X(i)];
    b[X(i)] = -1 / m[X(i)] - di[X(i)] / du22[X(i)] * (1 + r[X(i)]) -
              dl[X(i)] / du12[X(i)] * (1 - r[X(i)] * r[X(i)]);
    d[X(i)] = -told[X(i)] / m[X(i)];
  }
  d[X(0)] = -told[X(0)]; /* :210-213 */
  a[X(0)] = 0.0;
  b[X(0)] = -1.0;
  c[X(0)] = 0.0;
  i = nA + 1;
  if (NeuBCS) { /* :221-230 */
    a[X(i)] = 1.0 * (r[X(i)] * (1.0 + r[X(i)]) * di[X(i)] * m[X(i)] / du22[X(i)]);
    b[X(i)] = -1.0 * (1.0 + r[X(i)] * (1 + r[X(i)]) * di[X(i)] * m[X(i)] / du22[X(i)]);
    c[X(i)] = 0.0;
    d[X(i)] = -told[X(i)];
  } else {
    a[X(i)] = 0.0;
    b[X(i)] = -1.0;
    c[X(i)] = 0.0;
    d[X(i)] = -told[X(i)];
  }
  cp[X(0)] = c[X(0)] / b[X(0)]; /* :237-247 */
  for (i = 1; i <= nA + 1; i++) cp[X(i)] = c[X(i)] / (b[X(i)] - cp[X(i - 1)] * a[X(i)]);
  dp[X(0)] = d[X(0)] / b[X(0)];
  for (i = 1; i <= nA + 1; i++)
    dp[X(i)] = (d[X(i)] - dp[X(i - 1)] * a[X(i)]) / (b[X(i)] - cp[X(i - 1)] * a[X(i)]);
  temp[X(nA + 1)] = dp[X(nA + 1)];
  for (i = nA; i >= 0; i--) temp[X(i)] = dp[X(i)] - cp[X(i)] * temp[X(i + 1)];
  for (i = 0; i <= nA + 1; i++) dTdt[X(i)] = temp[X(i)] - Quantity[X(i)];
}

/* three solves (Dt/2, Dt/2 from the first result, Dt) and the second-order
 * combination (2H - F) - Q, as calc_thermal_conduction and calc_viscosity
 * do; `unit` scales the solved quantity (TempUnit for the temperature, NULL
 * for the winds); returns the last solve's result in `last` */
static void two_stage(int nA, double Dt, const double *Q, const double *unit,
                      const double *diff, const double *mulfac, const double *alt,
                      const double *rd, double *inc, double *last) {
  double q[LM], res[LM], s1[LM], h[LM], f[LM];
  int i;
  for (i = 0; i < nA + 4; i++) q[i] = unit ? Q[i] * unit[i] : Q[i];
  ov_conduction_1d(nA, Dt / 2.0, 1, q, diff, mulfac, alt, rd, res);
  for (i = 0; i <= nA + 1; i++)
    s1[X(i)] = unit ? Q[X(i)] + res[X(i)] / unit[X(i)] : Q[X(i)] + res[X(i)];
  s1[X(-1)] = s1[X(nA + 2)] = 0.0; /* not read */
  for (i = 0; i < nA + 4; i++) q[i] = unit ? s1[i] * unit[i] : s1[i];
  ov_conduction_1d(nA, Dt / 2.0, 1, q, diff, mulfac, alt, rd, res);
  for (i = 0; i <= nA + 1; i++)
    h[X(i)] = unit ? s1[X(i)] + res[X(i)] / unit[X(i)] : s1[X(i)] + res[X(i)];
  for (i = 0; i < nA + 4; i++) q[i] = unit ? Q[i] * unit[i] : Q[i];
  ov_conduction_1d(nA, Dt, 1, q, diff, mulfac, alt, rd, res);
  for (i = 0; i <= nA + 1; i++) {
    f[X(i)] = unit ? Q[X(i)] + res[X(i)] / unit[X(i)] : Q[X(i)] + res[X(i)];
    inc[X(i)] = (2.0 * h[X(i)] - f[X(i)]) - Q[X(i)];
    if (last) last[X(i)] = res[X(i)];
  }
}

void ov_sources_step(const ov_sources *s) {
  const int nA = s->n_alts, L = nA + 4;
  if (nA < 1 || L > LM) return;
  for (int col = 0; col < s->n_cols; col++) {
    const size_t o = (size_t)col * L;
    double diff[LM], mulfac[LM], cond[LM], visc[3][LM], vvisc[NSP][LM], last[LM];
    const double *alt = s->alt + o, *rd = s->raddist + o;
    const double *rho = s->rho + o, *eddy = s->eddy + o;
    int i, d, sp;
    for (i = 0; i < L; i++) diff[i] = mulfac[i] = 0.0;
    /* calc_thermal_conduction, calc_conduction.f90:48-123 */
    for (i = 0; i <= nA + 1; i++) {
      const double prandtl =
          s->use_turbulent_cond ? eddy[X(i)] * rho[X(i)] * s->cp[o + X(i)] : 0.0;
      mulfac[X(i)] = rho[X(i)] * s->cp[o + X(i)];
      diff[X(i)] = s->kappa_temp[o + X(i)] + prandtl;
    }
    two_stage(nA, s->dt, s->temp + o, s->tempunit + o, diff, mulfac, alt, rd, cond, NULL);
    /* calc_viscosity.f90:53-117, east and north */
    for (d = 0; d < 2; d++) {
      for (i = 0; i <= nA + 1; i++) {
        mulfac[X(i)] = rho[X(i)];
        diff[X(i)] = s->visc_coef[o + X(i)] + rho[X(i)] * eddy[X(i)];
      }
      two_stage(nA, s->dt, s->vel + (size_t)(col * 3 + d) * L, NULL, diff, mulfac, alt, rd,
                visc[d], NULL);
    }
    /* :119-186, each species; Viscosity(:,:,:,iUp_) keeps the last
     * species' full-step result (:161) */
    for (sp = 0; sp < NSP; sp++) {
      const double *n = s->nds + (size_t)(col * NSP + sp) * L;
      const double *vc = s->visc_coef_s + (size_t)(col * NSP + sp) * L;
      for (i = 0; i <= nA + 1; i++) {
        diff[X(i)] = (4.0 / 3.0) * vc[X(i)] + s->mass[sp] * n[X(i)] * eddy[X(i)];
        mulfac[X(i)] = s->mass[sp] * n[X(i)];
      }
      two_stage(nA, s->dt, s->vvel + (size_t)(col * NSP + sp) * L, NULL, diff, mulfac, alt,
                rd, vvisc[sp], last);
    }
    for (i = 0; i <= nA + 1; i++) visc[2][X(i)] = last[X(i)];
    if (s->conduction)
      for (i = 0; i <= nA + 1; i++) s->conduction[o + X(i)] = cond[X(i)];
    if (s->viscosity)
      for (d = 0; d < 3; d++)
        for (i = 0; i <= nA + 1; i++) s->viscosity[(size_t)(col * 3 + d) * L + X(i)] = visc[d][X(i)];
    if (s->vert_viscosity)
      for (sp = 0; sp < NSP; sp++)
        for (i = 0; i <= nA + 1; i++)
          s->vert_viscosity[(size_t)(col * NSP + sp) * L + X(i)] = vvisc[sp][X(i)];
    if (!s->apply) continue;
    /* add_sources.f90:53-70 */
    double *T = s->temp + o;
    for (i = 1; i <= nA; i++) {
      double t = T[X(i)] + s->dt * (s->heating ? s->heating[o + X(i)] : 0.0);
      if (s->chem_heating) t = t + s->chem_heating[o + X(i)];
      if (s->user_heating) t = t + s->user_heating[o + X(i)];
      T[X(i)] = t;
    }
    for (i = 0; i <= nA + 1; i++) T[X(i)] = T[X(i)] + cond[X(i)];
    /* :94-107 (ion drag, gravity waves and the sources-step friction are
     * other physics, added by the caller) */
    for (d = 0; d < 3; d++) {
      double *v = s->vel + (size_t)(col * 3 + d) * L;
      for (i = 0; i <= nA + 1; i++) v[X(i)] = v[X(i)] + visc[d][X(i)];
    }
    for (sp = 0; sp < NSP; sp++) {
      double *w = s->vvel + (size_t)(col * NSP + sp) * L;
      for (i = 0; i <= nA + 1; i++) w[X(i)] = w[X(i)] + vvisc[sp][X(i)];
    }
  }
}
