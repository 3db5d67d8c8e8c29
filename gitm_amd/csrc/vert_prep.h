// vert_prep.h -- the per-step inputs of the vertical step, computed on the
// device from the state: calc_pressure's NDensity, Pressure, cp and Gamma
// (calc_pressure.f90:17-23,56-86), calc_rates' MeanMajorMass and TempUnit
// (calc_rates.Earth.f90:34-50), calc_eddy_diffusion_coefficient
// (Earth.f90:376-411) and calc_timestep_vertical's cMax(iUp_) and CFL limit
// (calc_timestep.f90:85-147).
//
// One workgroup per column, one lane per level, as the step kernel; a level's
// values need nothing from its neighbours, and the column's CFL limit is a
// maximum over its levels (a workgroup reduction).  Operations keep the
// reference's order; min and max are exact, so the result does not depend on
// the order of the reductions.
#pragma once

#include <hip/hip_runtime.h>

#include "gitm_vert.h"
#include "vert_column.h"

namespace gv {

struct PrepParams {
  double mass[NSP], vib2[NSP];  // Vibration - 2
  double kb_over_m[NSP];        // Boltzmanns_Constant / Mass
  double cfl, eddy_coef, eddy_p0, eddy_p1;
  int n_alts, ions;
};

// Column prep of one level; returns the level's CFL rates (cMax/dAlt and, with
// ions, the ion one) through rate/irate (0 where they do not apply).
__device__ __forceinline__ void prep_level(const PrepParams &p, const gitm_vert_columns &io,
                                           const gitm_vert_field &pf,
                                           const gitm_vert_field &cpf, int ci, int cj,
                                           int cb, int k, double &rate, double &irate) {
  const int nA = p.n_alts;
  const long long bn = col_base(io.ndensity_s, ci, cj, cb);
  double n[NSP];
#pragma unroll
  for (int s = 0; s < NSP; s++) n[s] = *fld(io.ndensity_s, bn, k, s);
  const double T = *fld(io.temperature, col_base(io.temperature, ci, cj, cb), k, 0);
  const double rho = *fld(io.rho, col_base(io.rho, ci, cj, cb), k, 0);
  double N = 0.0;
#pragma unroll
  for (int s = 0; s < NSP; s++) N = N + n[s];
  const double P = T * rho;
  double c = 0.0, g = 0.0, m = 0.0;
#pragma unroll
  for (int s = 0; s < NSP; s++) {
    c = c + p.vib2[s] * n[s] * p.kb_over_m[s];
    g = g + n[s] / p.vib2[s];
  }
  c = c / (2.0 * N);
  g = g * 2.0 / N + 1;
#pragma unroll
  for (int s = 0; s < NSP; s++) m = m + p.mass[s] * n[s] / (N + 1.0);
  if (m == 0) m = p.mass[0];
  double e = 0.0;
  if (P > p.eddy_p0)
    e = p.eddy_coef;
  else if (P > p.eddy_p1)
    e = p.eddy_coef * (P - p.eddy_p1) / (p.eddy_p0 - p.eddy_p1);
  if (io.ndensity.ptr) *fld(io.ndensity, col_base(io.ndensity, ci, cj, cb), k, 0) = N;
  if (pf.ptr) *fld(pf, col_base(pf, ci, cj, cb), k, 0) = P;
  if (cpf.ptr) *fld(cpf, col_base(cpf, ci, cj, cb), k, 0) = c;
  *fld(io.gamma, col_base(io.gamma, ci, cj, cb), k, 0) = g;
  *fld(io.mean_major_mass, col_base(io.mean_major_mass, ci, cj, cb), k, 0) = m;
  *fld(io.temp_unit, col_base(io.temp_unit, ci, cj, cb), k, 0) = m / KB;
  *fld(io.eddy, col_base(io.eddy, ci, cj, cb), k, 0) = e;
  rate = 0.0;
  irate = 0.0;
  if (k >= 0 && k <= nA + 1) {
    const long long bw = col_base(io.vertical_velocity, ci, cj, cb);
    double w = 0.0;
#pragma unroll
    for (int s = 0; s < NSP; s++) w = fmax(w, fabs(*fld(io.vertical_velocity, bw, k, s)));
    const double cs = sqrt(g * T);
    const double cm = w + cs;
    *fld(io.cmax, col_base(io.cmax, ci, cj, cb), k, 0) = cm;
    if (k >= 1 && k <= nA) {
      const double da = *fld(io.dalt, col_base(io.dalt, ci, cj, cb), k, 0);
      rate = cm / da;
      if (p.ions) {
        const double iu =
            fabs(*fld(io.ivelocity, col_base(io.ivelocity, ci, cj, cb), k, 2));
        irate = (iu + cs) / da;
      }
    }
  }
}

// The adjoint of prep_level: the adjoints of the level's prep outputs
// (gf: TempUnit, Gamma, MeanMajorMass, cMax on levels 0..nAlts+1, the eddy
// coefficient; gcp: cp, may have a NULL ptr) are carried to NDensityS,
// Temperature, Rho and VerticalVelocity and added to gin.
__device__ __forceinline__ void prep_level_adjoint(const PrepParams &p,
                                                   const gitm_vert_columns &io,
                                                   const gitm_vert_adjoint_prep_fields &gf,
                                                   const gitm_vert_field &gcp,
                                                   const gitm_vert_adjoint_fields &gin, int ci,
                                                   int cj, int cb, int k) {
  const int nA = p.n_alts;
  const long long bn = col_base(io.ndensity_s, ci, cj, cb);
  double n[NSP];
#pragma unroll
  for (int s = 0; s < NSP; s++) n[s] = *fld(io.ndensity_s, bn, k, s);
  const double T = *fld(io.temperature, col_base(io.temperature, ci, cj, cb), k, 0);
  const double rho = *fld(io.rho, col_base(io.rho, ci, cj, cb), k, 0);
  double N = 0.0;
#pragma unroll
  for (int s = 0; s < NSP; s++) N = N + n[s];
  const double P = T * rho;
  double c = 0.0, g = 0.0, m = 0.0;
#pragma unroll
  for (int s = 0; s < NSP; s++) {
    c = c + p.vib2[s] * n[s] * p.kb_over_m[s];
    g = g + n[s] / p.vib2[s];
  }
  c = c / (2.0 * N);
  g = g * 2.0 / N + 1;
#pragma unroll
  for (int s = 0; s < NSP; s++) m = m + p.mass[s] * n[s] / (N + 1.0);
  const bool mfloor = m == 0;  // m = Mass(1), a constant
  if (mfloor) m = p.mass[0];
  double lm = *fld(gf.mean_major_mass, col_base(gf.mean_major_mass, ci, cj, cb), k, 0) +
              *fld(gf.temp_unit, col_base(gf.temp_unit, ci, cj, cb), k, 0) / KB;
  if (mfloor) lm = 0.0;
  double lg = *fld(gf.gamma, col_base(gf.gamma, ci, cj, cb), k, 0);
  const double lc = gcp.ptr ? *fld(gcp, col_base(gcp, ci, cj, cb), k, 0) : 0.0;
  const double le = *fld(gf.eddy, col_base(gf.eddy, ci, cj, cb), k, 0);
  double lT = 0.0, lrho = 0.0, lw[NSP] = {0, 0, 0, 0, 0, 0};
  if (k >= 0 && k <= nA + 1) {
    // cMax = max_s |w_s| + sqrt(Gamma Temperature)
    const double lcm = *fld(gf.cmax, col_base(gf.cmax, ci, cj, cb), k, 0);
    const long long bw = col_base(io.vertical_velocity, ci, cj, cb);
    double w = 0.0;
    int arg = -1;
#pragma unroll
    for (int s = 0; s < NSP; s++) {
      const double a = fabs(*fld(io.vertical_velocity, bw, k, s));
      if (a > w) {
        w = a;
        arg = s;
      }
    }
    if (arg >= 0) lw[arg] = *fld(io.vertical_velocity, bw, k, arg) < 0.0 ? -lcm : lcm;
    const double cs = sqrt(g * T);
    lg += lcm * 0.5 * T / cs;
    lT += lcm * 0.5 * g / cs;
  }
  // the eddy coefficient, linear in Pressure between the two thresholds
  if (!(P > p.eddy_p0) && P > p.eddy_p1) {
    const double lP = le * p.eddy_coef / (p.eddy_p0 - p.eddy_p1);
    lT += lP * rho;
    lrho += lP * T;
  }
  const long long gn = col_base(gin.ndensity_s, ci, cj, cb);
#pragma unroll
  for (int s = 0; s < NSP; s++) {
    const double ln = lg * (2.0 / (N * p.vib2[s]) - (g - 1.0) / N) +
                      lm * (p.mass[s] - m) / (N + 1.0) +
                      lc * (p.vib2[s] * p.kb_over_m[s] / (2.0 * N) - c / N);
    *fld(gin.ndensity_s, gn, k, s) += ln;
  }
  *fld(gin.temperature, col_base(gin.temperature, ci, cj, cb), k, 0) += lT;
  *fld(gin.rho, col_base(gin.rho, ci, cj, cb), k, 0) += lrho;
  if (k >= 0 && k <= nA + 1) {
    const long long gw = col_base(gin.vertical_velocity, ci, cj, cb);
#pragma unroll
    for (int s = 0; s < NSP; s++) *fld(gin.vertical_velocity, gw, k, s) += lw[s];
  }
}

}  // namespace gv
