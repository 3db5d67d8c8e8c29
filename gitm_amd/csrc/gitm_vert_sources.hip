// gitm_vert_sources.hip -- the operator-split sources step's neutral terms
// (gitm_vert_sources of include/gitm_vert.h): thermal conduction
// (calc_conduction.f90:19-125) and viscosity (calc_viscosity.f90:4-189),
// each three implicit solves of calc_conduction_1d
// (calc_conduction.f90:132-244) and the second-order combination
// (2H - F) - Q, then the updates of add_sources.f90:53-70,94-107.
//
// Mapping: one workgroup per column, one lane per level (blockDim =
// nAlts+4), as the vertical step.  The lanes stage the column in LDS, one
// level each with coalesced loads: per quantity (0 the temperature, 1-2 the
// east and north winds, 3-8 the species' vertical winds) the diffusion
// coefficient and MulFac times r^2 and the solved quantity, and the
// non-uniform grid factors.  The three tridiagonal (Thomas) solves of each
// quantity are serial in altitude: the Dt/2 and Dt solves from the state
// run at once on 18 lanes, then the second Dt/2 solve from the first's
// result on 9.  The lanes then write the increments and the updated state,
// again one level each.  Operations keep the reference's order: the
// results equal the oracle's bit for bit.
#include "vert_gpu.h"

namespace {

struct SrcParams {
  double dt;
  double mass[gv::NSP];
  int use_turbulent_cond, apply;
  int n_alts, n_cols;
};

struct SrcFields {
  gitm_vert_field kappa_temp, visc_coef, visc_coef_s, cp, raddist, heating, chem_heating,
      user_heating, conduction, viscosity, vvisc;
};

constexpr int NQ = 3 + gv::NSP;  // temperature, two winds, six species winds

// LDS record of one level (doubles): per quantity di = Diff r^2, mr = MulFac
// r^2 and the solved input (the quantity, times TempUnit for the
// temperature); the grid factors; the Thomas coefficients of the two
// concurrent solves (A: Dt/2, C: Dt); the first stage's result (then the
// increment), the full step's value F, and Velocity(up)'s increment.
enum {
  R_DI = 0, R_MR = R_DI + NQ, R_QIN = R_MR + NQ,
  R_ALT = R_QIN + NQ, R_R, R_DU12, R_DU22, R_UNIT,
  R_CPA = R_UNIT + 1, R_DPA = R_CPA + NQ, R_CPC = R_DPA + NQ, R_DPC = R_CPC + NQ,
  R_S1 = R_DPC + NQ, R_F = R_S1 + NQ, R_VUP = R_F + NQ,
  R_LEN = R_VUP + 1  // odd: 64-bit LDS accesses stay conflict-free
};
static_assert(R_LEN % 2 == 1, "odd record length");

struct SrcCol {
  const SrcParams &p;
  const gitm_vert_columns &io;
  const SrcFields &f;
  double *lds;
  int ci, cj, cb;

  __device__ double &rec(int k, int a) const { return lds[(long long)(k + 1) * R_LEN + a]; }
  __device__ double at(const gitm_vert_field &g, int k, int comp = 0) const {
    return *gv::fld(g, gv::col_base(g, ci, cj, cb), k, comp);
  }
  __device__ double *ptr(const gitm_vert_field &g, int k, int comp = 0) const {
    return gv::fld(g, gv::col_base(g, ci, cj, cb), k, comp);
  }
  __device__ double quantity(int q, int k) const {
    if (q == 0) return at(io.temperature, k);
    if (q < 3) return at(io.velocity, k, q - 1);
    return at(io.vertical_velocity, k, q - 3);
  }

  // one Thomas solve of quantity q (calc_conduction.f90:176-247) on levels
  // 0..nAlts+1: input from the record slot `in` (scaled by the unit for the
  // temperature when `scale`), coefficients to the slots cpa/dpa; returns
  // nothing, the back substitution is done by the caller through `emit`.
  template <class Emit>
  __device__ void solve(int q, double dts, int in_slot, bool scale, int cps, int dps,
                        Emit emit) const {
    const int nA = p.n_alts;
    auto input = [&](int k) {
      const double v = rec(k, in_slot + q);
      return scale ? v * rec(k, R_UNIT) : v;
    };
    // level 0: a = 0, b = -1, c = 0, d = -told (:210-213)
    double cpv = 0.0 / -1.0;
    double dpv = (-input(0)) / -1.0;
    rec(0, cps + q) = cpv;
    rec(0, dps + q) = dpv;
    for (int i = 1; i <= nA + 1; i++) {
      const double di0 = rec(i, R_DI + q), r = rec(i, R_R), du22 = rec(i, R_DU22);
      const double m = dts / rec(i, R_MR + q);
      const double told = input(i);
      double a, b, c, d;
      if (i <= nA) {
        const double du12 = rec(i, R_DU12);
        const double dl = rec(i + 1, R_DI + q) - r * r * rec(i - 1, R_DI + q) -
                          (1 - r * r) * di0;  // :188-190
        a = di0 / du22 * r - dl / du12 * r * r;
        c = di0 / du22 + dl / du12;
        b = -1 / m - di0 / du22 * (1 + r) - dl / du12 * (1 - r * r);
        d = -told / m;
      } else {  // Neumann top, :221-230
        a = 1.0 * (r * (1.0 + r) * di0 * m / du22);
        b = -1.0 * (1.0 + r * (1 + r) * di0 * m / du22);
        c = 0.0;
        d = -told;
      }
      const double cpn = c / (b - cpv * a);
      const double dpn = (d - dpv * a) / (b - cpv * a);
      rec(i, cps + q) = cpn;
      rec(i, dps + q) = dpn;
      cpv = cpn;
      dpv = dpn;
    }
    double tnext = dpv;  // temp(nAlts+1) = dp(nAlts+1)
    emit(nA + 1, tnext - input(nA + 1));
    for (int i = nA; i >= 0; i--) {
      tnext = rec(i, dps + q) - rec(i, cps + q) * tnext;
      emit(i, tnext - input(i));  // dTdt_cond (:249-250)
    }
  }
};

template <int NT>
__global__ void __launch_bounds__(NT)
    sources_kernel(SrcParams p, gitm_vert_columns io, SrcFields f) {
  extern __shared__ double lds[];
  const int col = (int)blockIdx.x, j = (int)threadIdx.x, k = j - 1;
  const int nA = p.n_alts;
  const int nlb = io.n_lons * io.n_lats;
  SrcCol c{p, io, f, lds, col % io.n_lons, (col / io.n_lons) % io.n_lats, col / nlb};
  // stage the level (coefficients on levels 0..nAlts+1)
  c.rec(k, R_ALT) = c.at(io.altitude, k);
  const double unit = c.at(io.temp_unit, k);
  c.rec(k, R_UNIT) = unit;
  if (k >= 0 && k <= nA + 1) {
    const double rd = c.at(f.raddist, k);
    const double r2 = rd * rd;
    const double rho = c.at(io.rho, k), eddy = c.at(io.eddy, k);
    {  // calc_conduction.f90:59-73
      const double cpk = c.at(f.cp, k);
      const double prandtl = p.use_turbulent_cond ? eddy * rho * cpk : 0.0;
      c.rec(k, R_DI + 0) = (c.at(f.kappa_temp, k) + prandtl) * r2;
      c.rec(k, R_MR + 0) = (rho * cpk) * r2;
    }
    const double vd = (c.at(f.visc_coef, k) + rho * eddy) * r2;  // calc_viscosity.f90:66-67
    c.rec(k, R_DI + 1) = vd;
    c.rec(k, R_DI + 2) = vd;
    c.rec(k, R_MR + 1) = rho * r2;
    c.rec(k, R_MR + 2) = rho * r2;
#pragma unroll
    for (int s = 0; s < gv::NSP; s++) {  // :125-128
      const double n = c.at(io.ndensity_s, k, s);
      const double mn = p.mass[s] * n;
      c.rec(k, R_DI + 3 + s) = ((4.0 / 3.0) * c.at(f.visc_coef_s, k, s) + mn * eddy) * r2;
      c.rec(k, R_MR + 3 + s) = (p.mass[s] * n) * r2;
    }
#pragma unroll
    for (int q = 0; q < NQ; q++) c.rec(k, R_QIN + q) = c.quantity(q, k);
  }
  __syncthreads();
  if (k >= 0 && k <= nA + 1) {  // grid factors (:181-186)
    const double altm = c.rec(k - 1, R_ALT), alt0 = c.rec(k, R_ALT),
                 altp = c.rec(k + 1, R_ALT);
    const double du = altp - alt0, dl = alt0 - altm;
    const double r = du / dl;
    c.rec(k, R_R) = r;
    c.rec(k, R_DU12) = du * du * (1 + r) * (1 + r);
    c.rec(k, R_DU22) = 0.5 * (dl * du + du * du);
  }
  __syncthreads();
  // the first Dt/2 solve (A) and the Dt solve (C) from the state at once
  // (columns of fewer than 18 levels take some in turn)
  for (int task = j; task < 2 * NQ; task += (int)blockDim.x) {
    const int q = task % NQ;
    const bool isT = q == 0;
    if (task < NQ) {
      c.solve(q, p.dt / 2.0, R_QIN, isT, R_CPA, R_DPA, [&](int i, double res) {
        const double q0 = c.rec(i, R_QIN + q);
        c.rec(i, R_S1 + q) = isT ? q0 + res / c.rec(i, R_UNIT) : q0 + res;
      });
    } else {
      c.solve(q, p.dt, R_QIN, isT, R_CPC, R_DPC, [&](int i, double res) {
        const double q0 = c.rec(i, R_QIN + q);
        c.rec(i, R_F + q) = isT ? q0 + res / c.rec(i, R_UNIT) : q0 + res;
        // Viscosity(:,:,:,iUp_) keeps the last species' full-step result
        // (calc_viscosity.f90:161)
        if (q == NQ - 1) c.rec(i, R_VUP) = res;
      });
    }
  }
  __syncthreads();
  // the second Dt/2 solve (B) from the first's result, and the combination
  // (2H - F) - Q (calc_conduction.f90:117-121, calc_viscosity.f90:110-114)
  for (int q = j; q < NQ; q += (int)blockDim.x) {
    const bool isT = q == 0;
    c.solve(q, p.dt / 2.0, R_S1, isT, R_CPA, R_DPA, [&](int i, double res) {
      const double s1 = c.rec(i, R_S1 + q);
      const double h = isT ? s1 + res / c.rec(i, R_UNIT) : s1 + res;
      c.rec(i, R_S1 + q) = (2.0 * h - c.rec(i, R_F + q)) - c.rec(i, R_QIN + q);
    });
  }
  __syncthreads();
  if (k < 0 || k > nA + 1) return;
  // the increments and add_sources' updates (:53-70, :94-114), one level per
  // lane
  const double cond = c.rec(k, R_S1 + 0);
  if (f.conduction.ptr) *c.ptr(f.conduction, k) = cond;
  if (f.viscosity.ptr) {
    *c.ptr(f.viscosity, k, 0) = c.rec(k, R_S1 + 1);
    *c.ptr(f.viscosity, k, 1) = c.rec(k, R_S1 + 2);
    *c.ptr(f.viscosity, k, 2) = c.rec(k, R_VUP);
  }
  if (f.vvisc.ptr)
#pragma unroll
    for (int s = 0; s < gv::NSP; s++) *c.ptr(f.vvisc, k, s) = c.rec(k, R_S1 + 3 + s);
  if (!p.apply) return;
  double tt = c.rec(k, R_QIN + 0);
  if (k >= 1 && k <= nA) {
    tt = tt + p.dt * (f.heating.ptr ? c.at(f.heating, k) : 0.0);
    if (f.chem_heating.ptr) tt = tt + c.at(f.chem_heating, k);
    if (f.user_heating.ptr) tt = tt + c.at(f.user_heating, k);
  }
  *c.ptr(io.temperature, k) = tt + cond;
  *c.ptr(io.velocity, k, 0) = c.rec(k, R_QIN + 1) + c.rec(k, R_S1 + 1);
  *c.ptr(io.velocity, k, 1) = c.rec(k, R_QIN + 2) + c.rec(k, R_S1 + 2);
  {
    double *vu = c.ptr(io.velocity, k, 2);
    *vu = *vu + c.rec(k, R_VUP);
  }
#pragma unroll
  for (int s = 0; s < gv::NSP; s++)
    *c.ptr(io.vertical_velocity, k, s) = c.rec(k, R_QIN + 3 + s) + c.rec(k, R_S1 + 3 + s);
}

}  // namespace

extern "C" int gitm_vert_sources(gitm_vert_handle *h, const gitm_vert_sources_params *sp,
                                 const gitm_vert_columns *c,
                                 const gitm_vert_sources_fields *sf) {
  if (!h || !sp || !c || !sf) return GITM_VERT_EINVAL;
  if (c->n_alts < 10 || c->n_alts > GITM_VERT_MAX_NALTS) return GITM_VERT_EINVAL;
  if (c->n_lons < 1 || c->n_lats < 1 || c->n_blocks < 1) return GITM_VERT_EINVAL;
  const long long ncols = (long long)c->n_lons * c->n_lats * c->n_blocks;
  if (ncols > 0x7fffffffLL) return GITM_VERT_EINVAL;
  const gitm_vert_field *req[] = {&c->temperature, &c->temp_unit, &c->rho, &c->eddy,
                                  &c->ndensity_s, &c->velocity, &c->vertical_velocity,
                                  &c->altitude, &sf->kappa_temp, &sf->visc_coef,
                                  &sf->visc_coef_s, &sf->cp, &sf->radial_distance};
  for (const gitm_vert_field *g : req)
    if (!g->ptr) return GITM_VERT_EINVAL;
  if (!(sp->dt > 0.0)) return GITM_VERT_EINVAL;
  for (int s = 0; s < gv::NSP; s++)
    if (!(sp->mass[s] > 0.0)) return GITM_VERT_EINVAL;
  GV_HIP(hipSetDevice(h->device));
  SrcParams p;
  p.dt = sp->dt;
  for (int s = 0; s < gv::NSP; s++) p.mass[s] = sp->mass[s];
  p.use_turbulent_cond = sp->use_turbulent_cond != 0;
  p.apply = sp->apply != 0;
  p.n_alts = c->n_alts;
  p.n_cols = (int)ncols;
  SrcFields f;
  f.kappa_temp = sf->kappa_temp;
  f.visc_coef = sf->visc_coef;
  f.visc_coef_s = sf->visc_coef_s;
  f.cp = sf->cp;
  f.raddist = sf->radial_distance;
  f.heating = sf->heating;
  f.chem_heating = sf->chem_heating;
  f.user_heating = sf->user_heating;
  f.conduction = sf->conduction;
  f.viscosity = sf->viscosity;
  f.vvisc = sf->vertical_viscosity;
  const size_t lds = (size_t)R_LEN * (size_t)(c->n_alts + 4) * sizeof(double);
  if (lds > (size_t)160 * 1024) return GITM_VERT_EUNSUPPORTED;  // nAlts > 231
  const dim3 grid((unsigned)ncols), block(c->n_alts + 4);
  switch (gv::auto_waves(c->n_alts)) {
    case 1: hipLaunchKernelGGL(sources_kernel<64>, grid, block, lds, h->stream, p, *c, f); break;
    case 2: hipLaunchKernelGGL(sources_kernel<128>, grid, block, lds, h->stream, p, *c, f); break;
    case 3: hipLaunchKernelGGL(sources_kernel<192>, grid, block, lds, h->stream, p, *c, f); break;
    case 4: hipLaunchKernelGGL(sources_kernel<256>, grid, block, lds, h->stream, p, *c, f); break;
    default: return GITM_VERT_EINVAL;
  }
  GV_HIP(hipGetLastError());
  return GITM_VERT_OK;
}
