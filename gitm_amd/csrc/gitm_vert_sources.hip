// gitm_vert_sources.hip -- the operator-split sources step's neutral terms
// (gitm_vert_sources of include/gitm_vert.h): thermal conduction
// (calc_conduction.f90:19-125) and viscosity (calc_viscosity.f90:4-189),
// each three implicit solves of calc_conduction_1d
// (calc_conduction.f90:132-244) and the second-order combination
// (2H - F) - Q, then the updates of add_sources.f90:53-70,94-107.
//
// Mapping: one workgroup per column, one lane per level (blockDim =
// nAlts+4), as the vertical step.  The tridiagonal (Thomas) solves are
// serial in altitude, so everything that is not is taken out of them: the
// lanes stage the column one level each (coalesced loads), form every
// solve's coefficients one level each, and the serial lanes (one per solve
// and quantity: 0 the temperature, 1-2 the east and north winds, 3-8 the
// species' vertical winds) run only the recurrences.  The two Dt/2 solves
// have the same matrix, so the second reuses the first's pivots and cp and
// runs only its dp recurrence.  Operations keep the reference's order: the
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

// LDS record of one level (doubles), per quantity q:
//   R_A, R_C   the tridiagonal a and c, shared by the three solves (they do
//              not depend on the step); at the top level, where a does,
//              R_C holds the Dt solve's a (c is 0 there);
//   R_BA       b of the Dt/2 solves, then the pivots b - cp a, which the
//              two Dt/2 solves share (same a, b and c, so the same cp);
//   R_BC       b of the Dt solve, then its cp; then the second Dt/2
//              solve's right side, its dp and its solution;
//   R_DA, R_DC d of the first Dt/2 solve and of the Dt solve, then their
//              dp, then their solutions;
//   R_CPA      the Dt/2 solves' cp.  Until the coefficients are formed this
//              region holds the diffusion coefficients times r^2 (one slot
//              for the two winds, whose coefficients are equal) and the
//              altitude.
enum {
  R_A = 0, R_C = R_A + NQ, R_BA = R_C + NQ, R_BC = R_BA + NQ, R_DA = R_BC + NQ,
  R_DC = R_DA + NQ, R_CPA = R_DC + NQ, R_LEN = R_CPA + NQ,
  R_DI = R_CPA, R_ALT = R_CPA + NQ - 1
};
static_assert(R_LEN % 2 == 1, "odd record length: 64-bit LDS accesses stay conflict-free");
constexpr int di_slot(int q) { return q == 0 ? 0 : (q < 3 ? 1 : q - 1); }
static_assert(di_slot(NQ - 1) < NQ - 1, "diffusion slots below the altitude's");

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
};

// Phases of one column (barriers between them):
//  1. stage: one lane per level -1..nAlts+2 loads its level; the diffusion
//     coefficients and the altitude go to LDS, the rest stays in registers;
//  2. coefficients: one lane per level 0..nAlts+1 forms the grid factors
//     and every solve's a, b, c and d (calc_conduction.f90:181-230);
//  3. the forward sweeps and back substitutions of the first Dt/2 solve and
//     the Dt solve, one lane per (solve, quantity), serial in altitude
//     (:232-247);
//  4. one lane per level: the first stage's value, the full step's, and
//     the second Dt/2 solve's right side;
//  5. the second Dt/2 solve on the first's pivots and cp: only its dp
//     recurrence and back substitution, one lane per quantity;
//  6. one lane per level: the combination (2H - F) - Q (:117-121,
//     calc_viscosity.f90:110-114), the increments and add_sources' updates.
template <int NT>
__global__ void __launch_bounds__(NT)
    sources_kernel(SrcParams p, gitm_vert_columns io, SrcFields f) {
  extern __shared__ double lds[];
  const int col = xcd_column((int)blockIdx.x, (int)gridDim.x), j = (int)threadIdx.x, k = j - 1;
  const int nA = p.n_alts;
  const int nlb = io.n_lons * io.n_lats;
  const bool lev = k >= 0 && k <= nA + 1;
  SrcCol c{p, io, f, lds, col % io.n_lons, (col / io.n_lons) % io.n_lats, col / nlb};
  const double dtA = p.dt / 2.0, dtC = p.dt;
  double qin[NQ], mr[NQ], unit = 1.0;
  // 1. stage
  c.rec(k, R_ALT) = c.at(io.altitude, k);
  if (lev) {
    unit = c.at(io.temp_unit, k);
    const double rd = c.at(f.raddist, k);
    const double r2 = rd * rd;
    const double rho = c.at(io.rho, k), eddy = c.at(io.eddy, k);
    {  // calc_conduction.f90:59-73
      const double cpk = c.at(f.cp, k);
      const double prandtl = p.use_turbulent_cond ? eddy * rho * cpk : 0.0;
      c.rec(k, R_DI + di_slot(0)) = (c.at(f.kappa_temp, k) + prandtl) * r2;
      mr[0] = (rho * cpk) * r2;
    }
    // calc_viscosity.f90:66-67
    c.rec(k, R_DI + di_slot(1)) = (c.at(f.visc_coef, k) + rho * eddy) * r2;
    mr[1] = rho * r2;
    mr[2] = rho * r2;
#pragma unroll
    for (int s = 0; s < gv::NSP; s++) {  // :125-128
      const double n = c.at(io.ndensity_s, k, s);
      const double mn = p.mass[s] * n;
      c.rec(k, R_DI + di_slot(3 + s)) =
          ((4.0 / 3.0) * c.at(f.visc_coef_s, k, s) + mn * eddy) * r2;
      mr[3 + s] = (p.mass[s] * n) * r2;
    }
#pragma unroll
    for (int q = 0; q < NQ; q++) qin[q] = c.quantity(q, k);
  }
  __syncthreads();
  // 2. coefficients; they read only R_CPA's region and write the others
  if (lev) {
    const double altm = c.rec(k - 1, R_ALT), alt0 = c.rec(k, R_ALT),
                 altp = c.rec(k + 1, R_ALT);
    const double du = altp - alt0, dl0 = alt0 - altm;  // :181-186
    const double r = du / dl0;
    const double du12 = du * du * (1 + r) * (1 + r);
    const double du22 = 0.5 * (dl0 * du + du * du);
#pragma unroll
    for (int q = 0; q < NQ; q++) {
      const double told = q == 0 ? qin[0] * unit : qin[q];
      const double di0 = c.rec(k, R_DI + di_slot(q));
      if (k == 0) {  // a = 0, b = -1, c = 0, d = -told (:210-213)
        c.rec(k, R_A + q) = 0.0;
        c.rec(k, R_C + q) = 0.0;
        c.rec(k, R_BA + q) = -1.0;
        c.rec(k, R_BC + q) = -1.0;
        c.rec(k, R_DA + q) = -told;
        c.rec(k, R_DC + q) = -told;
      } else if (k <= nA) {
        const double mA = dtA / mr[q], mC = dtC / mr[q];
        const double dl = c.rec(k + 1, R_DI + di_slot(q)) -
                          r * r * c.rec(k - 1, R_DI + di_slot(q)) -
                          (1 - r * r) * di0;  // :188-190
        c.rec(k, R_A + q) = di0 / du22 * r - dl / du12 * r * r;
        c.rec(k, R_C + q) = di0 / du22 + dl / du12;
        c.rec(k, R_BA + q) = -1 / mA - di0 / du22 * (1 + r) - dl / du12 * (1 - r * r);
        c.rec(k, R_BC + q) = -1 / mC - di0 / du22 * (1 + r) - dl / du12 * (1 - r * r);
        c.rec(k, R_DA + q) = -told / mA;
        c.rec(k, R_DC + q) = -told / mC;
      } else {  // Neumann top, :221-230
        const double mA = dtA / mr[q], mC = dtC / mr[q];
        c.rec(k, R_A + q) = 1.0 * (r * (1.0 + r) * di0 * mA / du22);
        c.rec(k, R_C + q) = 1.0 * (r * (1.0 + r) * di0 * mC / du22);
        c.rec(k, R_BA + q) = -1.0 * (1.0 + r * (1 + r) * di0 * mA / du22);
        c.rec(k, R_BC + q) = -1.0 * (1.0 + r * (1 + r) * di0 * mC / du22);
        c.rec(k, R_DA + q) = -told;
        c.rec(k, R_DC + q) = -told;
      }
    }
  }
  __syncthreads();
  // 3. the first Dt/2 solve (A) and the Dt solve (C), at once (columns of
  // fewer than 18 levels take some in turn)
  for (int task = j; task < 2 * NQ; task += (int)blockDim.x) {
    const int q = task % NQ;
    const bool isA = task < NQ;
    // b's slot takes the pivot, then (for C, where cs == bs) cp
    const int bs = (isA ? R_BA : R_BC) + q, ds = (isA ? R_DA : R_DC) + q,
              cs = (isA ? R_CPA : R_BC) + q;
    double cpv = 0.0 / -1.0;
    double dpv = c.rec(0, ds) / -1.0;
    c.rec(0, cs) = cpv;
    c.rec(0, ds) = dpv;
    for (int i = 1; i <= nA; i++) {
      const double a = c.rec(i, R_A + q), cc = c.rec(i, R_C + q);
      const double den = c.rec(i, bs) - cpv * a;
      const double cpn = cc / den;
      const double dpn = (c.rec(i, ds) - dpv * a) / den;
      c.rec(i, bs) = den;
      c.rec(i, cs) = cpn;
      c.rec(i, ds) = dpn;
      cpv = cpn;
      dpv = dpn;
    }
    const int top = nA + 1;
    const double a = c.rec(top, (isA ? R_A : R_C) + q);
    const double den = c.rec(top, bs) - cpv * a;
    double t = (c.rec(top, ds) - dpv * a) / den;  // temp(nAlts+1) = dp(nAlts+1)
    c.rec(top, bs) = den;
    c.rec(top, ds) = t;
    for (int i = nA; i >= 0; i--) {
      t = c.rec(i, ds) - c.rec(i, cs) * t;
      c.rec(i, ds) = t;
    }
  }
  __syncthreads();
  // 4. the first stage's value s1 and the full step's fv (dTdt_cond,
  // :249-250), and the second Dt/2 solve's right side
  double s1[NQ], fv[NQ], vup = 0.0;
  if (lev) {
#pragma unroll
    for (int q = 0; q < NQ; q++) {
      const double in = q == 0 ? qin[0] * unit : qin[q];
      const double ra = c.rec(k, R_DA + q) - in, rc = c.rec(k, R_DC + q) - in;
      s1[q] = q == 0 ? qin[0] + ra / unit : qin[q] + ra;
      fv[q] = q == 0 ? qin[0] + rc / unit : qin[q] + rc;
      // Viscosity(:,:,:,iUp_) keeps the last species' full-step result
      // (calc_viscosity.f90:161)
      if (q == NQ - 1) vup = rc;
      const double told = q == 0 ? s1[0] * unit : s1[q];
      c.rec(k, R_BC + q) = (k >= 1 && k <= nA) ? -told / (dtA / mr[q]) : -told;
    }
  }
  __syncthreads();
  // 5. the second Dt/2 solve (B): A's pivots and cp, its own dp
  for (int q = j; q < NQ; q += (int)blockDim.x) {
    double dpv = c.rec(0, R_BC + q) / -1.0;
    c.rec(0, R_BC + q) = dpv;
    for (int i = 1; i <= nA + 1; i++) {
      dpv = (c.rec(i, R_BC + q) - dpv * c.rec(i, R_A + q)) / c.rec(i, R_BA + q);
      c.rec(i, R_BC + q) = dpv;
    }
    double t = dpv;
    for (int i = nA; i >= 0; i--) {
      t = c.rec(i, R_BC + q) - c.rec(i, R_CPA + q) * t;
      c.rec(i, R_BC + q) = t;
    }
  }
  __syncthreads();
  if (!lev) return;
  // 6. the combination, the increments and add_sources' updates
  // (:53-70, :94-114), one level per lane
  double out[NQ];
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    const double in = q == 0 ? s1[0] * unit : s1[q];
    const double rb = c.rec(k, R_BC + q) - in;
    const double h = q == 0 ? s1[0] + rb / unit : s1[q] + rb;
    out[q] = (2.0 * h - fv[q]) - qin[q];
  }
  const double cond = out[0];
  if (f.conduction.ptr) *c.ptr(f.conduction, k) = cond;
  if (f.viscosity.ptr) {
    *c.ptr(f.viscosity, k, 0) = out[1];
    *c.ptr(f.viscosity, k, 1) = out[2];
    *c.ptr(f.viscosity, k, 2) = vup;
  }
  if (f.vvisc.ptr)
#pragma unroll
    for (int s = 0; s < gv::NSP; s++) *c.ptr(f.vvisc, k, s) = out[3 + s];
  if (!p.apply) return;
  double tt = qin[0];
  if (k >= 1 && k <= nA) {
    tt = tt + p.dt * (f.heating.ptr ? c.at(f.heating, k) : 0.0);
    if (f.chem_heating.ptr) tt = tt + c.at(f.chem_heating, k);
    if (f.user_heating.ptr) tt = tt + c.at(f.user_heating, k);
  }
  *c.ptr(io.temperature, k) = tt + cond;
  *c.ptr(io.velocity, k, 0) = qin[1] + out[1];
  *c.ptr(io.velocity, k, 1) = qin[2] + out[2];
  {
    double *vu = c.ptr(io.velocity, k, 2);
    *vu = *vu + vup;
  }
#pragma unroll
  for (int s = 0; s < gv::NSP; s++) *c.ptr(io.vertical_velocity, k, s) = qin[3 + s] + out[3 + s];
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
  if (lds > (size_t)160 * 1024) return GITM_VERT_EUNSUPPORTED;  // 129 KB at nAlts 252
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

// Host arrays (GITM's own): every field staged on the device over the levels
// the step touches (the altitude on -1..nAlts+2, the heating terms on
// 1..nAlts, the rest on 0..nAlts+1), the written ones copied back.
extern "C" int gitm_vert_sources_host(gitm_vert_handle *h, const gitm_vert_sources_params *sp,
                                      const gitm_vert_columns *c,
                                      const gitm_vert_sources_fields *sf) {
  if (!h || !sp || !c || !sf) return GITM_VERT_EINVAL;
  const int n = c->n_lons, m = c->n_lats, nA = c->n_alts, nb = c->n_blocks;
  if (n < 1 || m < 1 || nA < 1 || nb < 1) return GITM_VERT_EINVAL;
  gitm_vert_columns d = *c;
  gitm_vert_sources_fields g = *sf;
  const bool apply = sp->apply != 0;
  std::vector<HostSpan> spans;
  auto add = [&](gitm_vert_field *f, int nc, int lo, int hi, bool written) {
    spans.push_back({f, 1, n, 1, m, lo, hi, nc, written});
  };
  add(&d.altitude, 1, -1, nA + 2, false);
  add(&d.temperature, 1, 0, nA + 1, apply);
  add(&d.velocity, 3, 0, nA + 1, apply);
  add(&d.vertical_velocity, gv::NSP, 0, nA + 1, apply);
  add(&d.temp_unit, 1, 0, nA + 1, false);
  add(&d.rho, 1, 0, nA + 1, false);
  add(&d.eddy, 1, 0, nA + 1, false);
  add(&d.ndensity_s, gv::NSP, 0, nA + 1, false);
  add(&g.kappa_temp, 1, 0, nA + 1, false);
  add(&g.visc_coef, 1, 0, nA + 1, false);
  add(&g.visc_coef_s, gv::NSP, 0, nA + 1, false);
  add(&g.cp, 1, 0, nA + 1, false);
  add(&g.radial_distance, 1, 0, nA + 1, false);
  add(&g.heating, 1, 1, nA, false);
  add(&g.chem_heating, 1, 1, nA, false);
  add(&g.user_heating, 1, 1, nA, false);
  add(&g.conduction, 1, 0, nA + 1, true);
  add(&g.viscosity, 3, 0, nA + 1, true);
  add(&g.vertical_viscosity, gv::NSP, 0, nA + 1, true);
  GV_HIP(hipSetDevice(h->device));
  HostStaging st;
  int rc = stage_in(h, spans, nb, 96, st);  // slots after the other host entries'
  if (rc) return rc;
  rc = gitm_vert_sources(h, sp, &d, &g);
  if (rc) return rc;
  rc = stage_out(h, spans, 96, st);
  if (rc) return rc;
  GV_HIP(hipStreamSynchronize(h->stream));
  return GITM_VERT_OK;
}
