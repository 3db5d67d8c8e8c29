// vert_params.h -- host-side checks and conversion of gitm_vert_params /
// gitm_vert_columns into the kernel constants (shared by the HIP library and
// the host build of the column kernels).
#pragma once

#include <math.h>
#include <string.h>

#include "gitm_vert.h"
#include "vert_column.h"

namespace gv {

inline void default_params(gitm_vert_params *p) {
  memset(p, 0, sizeof(*p));
  p->dt = 2.0;
  p->omega_body = 2.0 * 3.141592653589793 / (24.0 * 3600.0);  // ModInputs.f90:285
  p->max_vertical_velocity = 1000.0;                           // :224
  p->beta_limiter = 2.0;                                       // :104
  p->min_tec = 2.0;                                            // :261
  p->vert_tau = 1.0e9;                                         // ModEarth.f90:325
  // init_planet, ModEarth.f90:271-281
  const double amu = 1.6726e-27, mo = 16.0 * amu, mn = 14.0 * amu;
  p->mass[0] = mo;
  p->mass[1] = 2.0 * mo;
  p->mass[2] = 2.0 * mn;
  p->mass[3] = mn;
  p->mass[4] = mn + mo;
  p->mass[5] = 4.0 * amu;
  p->thermal_diff_coef[5] = -0.38;  // calc_rates.Earth.f90:83-84
  p->use_neutral_friction = 1;
  p->use_coriolis = 1;
  p->use_ion_advection = 0;
  p->use_improved_ion_advection = 1;
  p->use_nighttime_ion_bcs = 1;
}

inline bool field_ok(const gitm_vert_field &f) { return f.ptr != nullptr; }

// Returns GITM_VERT_OK or the error code for unusable arguments.
inline int check_args(const gitm_vert_params *p, const gitm_vert_columns *c,
                      int solver, int n_steps) {
  if (!p || !c) return GITM_VERT_EINVAL;
  if (solver != GITM_VERT_AUSM && solver != GITM_VERT_RUSANOV) return GITM_VERT_EINVAL;
  if (n_steps < 0) return GITM_VERT_EINVAL;
  if (c->n_alts < 10 || c->n_alts > GITM_VERT_MAX_NALTS) return GITM_VERT_EINVAL;
  if (c->n_lons < 1 || c->n_lats < 1 || c->n_blocks < 1) return GITM_VERT_EINVAL;
  if ((long long)c->n_lons * c->n_lats * c->n_blocks > 0x7fffffffLL) return GITM_VERT_EINVAL;
  if (p->use_bc_perturbation || p->during_perturb) return GITM_VERT_EUNSUPPORTED;
  if (p->use_damping && solver == GITM_VERT_AUSM) return GITM_VERT_EUNSUPPORTED;
  if (p->use_msis && !field_ok(c->msis_bc)) return GITM_VERT_EINVAL;
  if ((p->use_gswm_tides || p->use_waccm_tides) &&
      !(field_ok(c->tides_east) && field_ok(c->tides_north) && field_ok(c->tides_temp)))
    return GITM_VERT_EINVAL;
  if (!(p->dt > 0.0) || !(p->vert_tau != 0.0)) return GITM_VERT_EINVAL;
  for (int s = 0; s < NSP; s++)
    if (!(p->mass[s] > 0.0)) return GITM_VERT_EINVAL;
  const gitm_vert_field *req[] = {&c->ndensity_s, &c->vertical_velocity, &c->velocity,
                                  &c->temperature, &c->rho, &c->temp_unit,
                                  &c->mean_major_mass, &c->gamma, &c->cmax, &c->eddy,
                                  &c->altitude, &c->gravity, &c->inv_radial_distance,
                                  &c->dalt, &c->cos_latitude};
  for (const gitm_vert_field *f : req)
    if (!field_ok(*f)) return GITM_VERT_EINVAL;
  if (c->cmax.first_level < -1 || c->cmax.first_level > 0) return GITM_VERT_EINVAL;
  if (p->use_ion_advection) {
    if (!field_ok(c->idensity_s) || !field_ok(c->ivelocity)) return GITM_VERT_EINVAL;
    if (p->use_nighttime_ion_bcs && (!field_ok(c->sza) || !field_ok(c->mlat)))
      return GITM_VERT_EINVAL;
  }
  return GITM_VERT_OK;
}

inline KParams make_kparams(const gitm_vert_params *p, const gitm_vert_columns *c,
                            int solver, int n_steps, unsigned long long *err) {
  KParams k;
  memset(&k, 0, sizeof(k));
  k.dt = p->dt;
  k.beta = p->beta_limiter;
  k.maxvv = p->max_vertical_velocity;
  k.min_tec = p->min_tec;
  k.vert_tau = p->vert_tau;
  k.omega = p->omega_body;
  for (int s = 0; s < NSP; s++) {
    k.mass[s] = p->mass[s];
    k.tdc[s] = p->thermal_diff_coef[s];
    k.lmass[s] = log(p->mass[s]);
    k.imass[s] = 1.0 / p->mass[s];
  }
  // alog(MinKpS) etc., vertical_solver_ausm.f90:1317-1332
  k.log_min_kp = log(1.0e-19);
  k.log_max_kp = log(1.0e-09);
  k.log_min_ku = log(0.25);
  k.log_max_ku = log(0.50);
  k.n_alts = c->n_alts;
  k.n_steps = n_steps;
  k.solver = solver;
  k.friction = p->use_neutral_friction != 0;
  k.coriolis = p->use_coriolis != 0;
  k.ions = p->use_ion_advection != 0;
  k.improved_ions = p->use_improved_ion_advection != 0;
  k.nighttime_ion_bcs = p->use_nighttime_ion_bcs != 0;
  k.msis = p->use_msis != 0;
  // WACCM's assignments follow GSWM's and replace all of them (:150-159)
  k.tides = p->use_waccm_tides ? 2 : p->use_gswm_tides ? 1 : 0;
  k.damping = p->use_damping != 0;
  k.scr_fields = 0;  // set with the tile (scratch_fields)
  k.err = err;
  return k;
}

// Workgroup tiles: wavefronts per column, one lane per level.
inline bool valid_waves(int w) { return w >= 1 && w <= 4; }
inline bool tile_fits(int w, int n_alts) { return 64 * w >= n_alts + 4; }
// Smallest tile holding the nAlts+4 levels.
inline int auto_waves(int n_alts) { return (n_alts + 4 + 63) / 64; }

inline void decode_error(unsigned long long key, gitm_vert_error *e) {
  if (key == ~0ULL) {
    e->code = GITM_VERT_OK;
    e->column = -1;
    e->level = 0;
  } else {
    e->column = (int)(key >> 32);
    e->code = (int)((key >> 16) & 0xffff);
    e->level = (int)(key & 0xffff) - 1024;
  }
  e->reserved = 0;
}

}  // namespace gv
