// vert_ad.h -- reverse-mode (adjoint) build of the column step.
//
// The adjoint of one vertical step of a column is taken by running the
// column code of vert_column.h once more over the step's input state with a
// taped number type (gv::adm::ad): every arithmetic operation on state-
// dependent values appends a node (its two parents and the partial
// derivatives with respect to them) to the tape of the lane that performs
// it, then the tape is swept backwards, accumulating adjoints from the
// step's outputs to its inputs.  Values a lane reads from another lane's LDS
// record carry the producing lane's node number, so the sweep propagates
// across levels exactly as the stencils and the hydrostatic recurrence
// couple them.
//
// Ordering of the sweep.  A lane can read another lane's value only after a
// workgroup barrier.  Every barrier of the forward pass closes a phase (the
// lane's node count is recorded there), and the backward sweep runs the
// phases last to first with a barrier between them: when a lane sweeps a
// phase, every node of that phase has received all its adjoint
// contributions, from its own later nodes and from other lanes' later
// phases.  Contributions are atomic adds (several lanes may feed one node).
//
// Node numbers interleave the lanes of a launch (node n of lane g is
// n * stride + g), so consecutive lanes write consecutive tape entries.
// The forward values are computed by the same operations in the same order
// as the plain build's reference forms (device_math() is false here), so the
// recomputed step reproduces the oracle's arithmetic.
#pragma once

#include "vert_column.h"

#ifdef GV_HOST_EMULATION
#include <atomic>
#define GV_AD inline
#else
#define GV_AD __device__ __forceinline__
#endif

namespace gv {

// One launch's tape: node arrays indexed by node number, plus the phase
// table (node count of every lane at every barrier).
// One node's parents (-1: none) and the partial derivatives with respect to
// them, each pair stored and loaded as one access.
struct alignas(8) AdPar {
  int a, b;
};
struct alignas(16) AdDer {
  double a, b;
};

struct AdTape {
  AdDer *d;         // partial derivatives w.r.t. the parents
  AdPar *p;         // parent nodes
  double *adj;      // adjoints, all zero before the forward part (cleared by the caller)
  int *phase;       // [phase * stride + lane] node count at the barrier
  int *nphase;      // [lane] barriers of the lane's forward pass
  int stride;       // lanes in the launch
  int nmax;         // nodes per lane
  int maxph;        // barriers per lane
  int *overflow;    // set to 1 when a lane runs out of nodes or phases
};

#ifdef GV_HOST_EMULATION
// host build: each workgroup lane is a CPU thread
struct AdLane {
  const AdTape *t = nullptr;
  int gid = 0;  // lane number in the launch
  int n = 0;    // nodes made
};
inline thread_local AdLane ad_lane;
GV_AD const AdTape &ad_tape() { return *ad_lane.t; }
GV_AD int ad_gid() { return ad_lane.gid; }
GV_AD int &ad_count() { return ad_lane.n; }
#else
// device build: the tape of the workgroup and each lane's node count in LDS
__shared__ AdTape ad_tape_sh;
__shared__ int ad_count_sh[256];
GV_AD const AdTape &ad_tape() { return ad_tape_sh; }
GV_AD int ad_gid() { return (int)(blockIdx.x * blockDim.x + threadIdx.x); }
GV_AD int &ad_count() { return ad_count_sh[threadIdx.x]; }
#endif

// Append a node; returns its number, or -1 when both parents are constants
// (or the tape is full, which sets the overflow flag).
GV_AD int ad_record(int a, double da, int b, double db, bool leaf = false) {
  if (!leaf && a < 0 && b < 0) return -1;
  const AdTape &t = ad_tape();
  int &n = ad_count();
  if (n >= t.nmax) {
    *t.overflow = 1;
    return -1;
  }
  const long long id = (long long)n * t.stride + ad_gid();
  n = n + 1;
  t.p[id] = AdPar{a, b};
  t.d[id] = AdDer{da, db};
  return (int)id;
}

namespace adm {

struct ad {
  double v;
  int i;
  GV_AD ad() : v(0.0), i(-1) {}
  GV_AD ad(double x) : v(x), i(-1) {}  // a constant
};

GV_AD ad mk(double v, int a, double da, int b = -1, double db = 0.0) {
  ad r;
  r.v = v;
  r.i = ad_record(a, da, b, db);
  return r;
}

GV_AD ad operator+(ad a, ad b) { return mk(a.v + b.v, a.i, 1.0, b.i, 1.0); }
GV_AD ad operator+(ad a, double b) { return mk(a.v + b, a.i, 1.0); }
GV_AD ad operator+(double a, ad b) { return mk(a + b.v, b.i, 1.0); }
GV_AD ad operator-(ad a, ad b) { return mk(a.v - b.v, a.i, 1.0, b.i, -1.0); }
GV_AD ad operator-(ad a, double b) { return mk(a.v - b, a.i, 1.0); }
GV_AD ad operator-(double a, ad b) { return mk(a - b.v, b.i, -1.0); }
GV_AD ad operator-(ad a) { return mk(-a.v, a.i, -1.0); }
GV_AD ad operator*(ad a, ad b) { return mk(a.v * b.v, a.i, b.v, b.i, a.v); }
GV_AD ad operator*(ad a, double b) { return mk(a.v * b, a.i, b); }
GV_AD ad operator*(double a, ad b) { return mk(a * b.v, b.i, a); }
GV_AD ad operator/(ad a, ad b) {
  const double q = a.v / b.v;
  return mk(q, a.i, 1.0 / b.v, b.i, -q / b.v);
}
GV_AD ad operator/(ad a, double b) { return mk(a.v / b, a.i, 1.0 / b); }
GV_AD ad operator/(double a, ad b) {
  const double q = a / b.v;
  return mk(q, b.i, -q / b.v);
}

#define GV_AD_CMP(OP)                                                  \
  GV_AD bool operator OP(ad a, ad b) { return a.v OP b.v; }            \
  GV_AD bool operator OP(ad a, double b) { return a.v OP b; }          \
  GV_AD bool operator OP(double a, ad b) { return a OP b.v; }
GV_AD_CMP(<)
GV_AD_CMP(>)
GV_AD_CMP(<=)
GV_AD_CMP(>=)
GV_AD_CMP(==)
GV_AD_CMP(!=)
#undef GV_AD_CMP

GV_AD ad exp(ad a) {
  const double e = ::exp(a.v);
  return mk(e, a.i, e);
}
GV_AD ad log(ad a) { return mk(gv::ln(a.v), a.i, 1.0 / a.v); }
GV_AD ad sqrt(ad a) {
  const double s = ::sqrt(a.v);
  return mk(s, a.i, 0.5 / s);
}
GV_AD ad tanh(ad a) {
  const double t = ::tanh(a.v);
  return mk(t, a.i, 1.0 - t * t);
}
GV_AD ad pow(ad a, double p) {
  const double r = ::pow(a.v, p);
  return mk(r, a.i, p * ::pow(a.v, p - 1.0));
}
GV_AD ad fabs(ad a) { return mk(::fabs(a.v), a.i, a.v < 0.0 ? -1.0 : 1.0); }

}  // namespace adm

GV_FN double val(double x) { return x; }
GV_FN double val(adm::ad x) { return x.v; }

// Adjoint of one AUSM step of column `col` of io (no ions), forward part:
// the step is recomputed from io's state with taping and the adjoints
// `gout` of the step's outputs (NDensityS, VerticalVelocity, Velocity,
// Temperature, Rho after the step) are seeded.  The 17 state inputs of the
// lane's level are its first nodes (node q = input q of adjoint_store).
// The prep inputs (TempUnit, Gamma, MeanMajorMass, cMax, eddy coefficient,
// geometry) are held fixed.  Ctx: lane(), sync() (which closes a phase),
// sync_or(), report(), device_math() == false, diff_powers(), end_forward()
// (closes the last phase) and add(node, value) (an atomic adjoint add).
template <int NT, class Ctx>
GV_FN void adjoint_forward(Ctx &cx, const KParams &kp, const gitm_vert_columns &io,
                           const AdjFields &gout, int col, adm::ad *lds, double *scratch,
                           double dt, int err_col) {
  using adm::ad;
  Column<NT, Ctx, false> c(cx, kp, lds, err_col, dt);
  c.scr = reinterpret_cast<ad *>(scratch) + (long long)col * kp.scr_fields * (kp.n_alts + 4);
  const int nlb = io.n_lons * io.n_lats;
  const int ci = col % io.n_lons, cj = (col / io.n_lons) % io.n_lats, cb = col / nlb;
  const int k = c.k, nA = c.nA;
  // leaves: the 17 state inputs of the level are the lane's nodes 0..16
  ad in[17];
  double alt, grav, invr, dalt, mmm, gam, eddy = 0.0, cmax = 0.0, tu, coslat;
  {
    const gitm_vert_field &f = io.cos_latitude;
    coslat = f.ptr[col_base(f, ci, cj, cb)];
  }
#define GV_LD(F) (*fld(io.F, col_base(io.F, ci, cj, cb), k, 0))
  alt = GV_LD(altitude);
  grav = GV_LD(gravity);
  invr = GV_LD(inv_radial_distance);
  dalt = GV_LD(dalt);
  mmm = GV_LD(mean_major_mass);
  gam = GV_LD(gamma);
  tu = GV_LD(temp_unit);
  if (k >= 0 && k <= nA + 1) cmax = GV_LD(cmax);
  if (k >= 1 && k <= nA) eddy = GV_LD(eddy);
  {
    const long long bn = col_base(io.ndensity_s, ci, cj, cb);
    const long long bw = col_base(io.vertical_velocity, ci, cj, cb);
    const long long bv = col_base(io.velocity, ci, cj, cb);
    double v[17];
    for (int s = 0; s < NSP; s++) {
      v[s] = *fld(io.ndensity_s, bn, k, s);
      v[6 + s] = *fld(io.vertical_velocity, bw, k, s);
    }
    for (int d = 0; d < 3; d++) v[12 + d] = *fld(io.velocity, bv, k, d);
    v[15] = GV_LD(temperature);
    v[16] = GV_LD(rho);
    for (int q = 0; q < 17; q++) {
      in[q].v = v[q];
      in[q].i = ad_record(-1, 0.0, -1, 0.0, true);
    }
  }
#undef GV_LD
  c.setup(alt, grav, invr, dalt, mmm, gam, eddy, cmax, tu, coslat);
  bool bad = false;
  for (int s = 0; s < NSP; s++) bad = bad || !(in[s] > 0.0);
  if (cx.sync_or(bad)) {
    if (bad) c.fail(GITM_VERT_ENONPOSITIVE, k);
    return;
  }
  for (int s = 0; s < NSP; s++) {
    c.U[F_LN + s] = ln(in[s]);
    c.U[F_W + s] = in[6 + s];
  }
  c.U[F_VE] = in[12];
  c.U[F_VN] = in[13];
  c.U[F_VU] = in[14];
  c.U[F_T] = in[15] * tu;
  c.U[F_LR] = ln(in[16]);
  c.ausm_step();
  // the step's outputs as advance_vertical scatters them (:163-185)
  ad out[17];
  for (int s = 0; s < NSP; s++) {
    out[s] = exp(c.U[F_LN + s]);
    out[6 + s] = c.U[F_W + s];
  }
  out[12] = c.U[F_VE];
  out[13] = c.U[F_VN];
  out[14] = c.U[F_VU];
  out[15] = c.U[F_T] / tu;
  out[16] = exp(c.U[F_LR]);
  cx.end_forward();
  {
    const long long bn = col_base(gout.nds, ci, cj, cb);
    const long long bw = col_base(gout.vvel, ci, cj, cb);
    const long long bv = col_base(gout.vel, ci, cj, cb);
    double g[17];
    for (int s = 0; s < NSP; s++) {
      g[s] = *fld(gout.nds, bn, k, s);
      g[6 + s] = *fld(gout.vvel, bw, k, s);
    }
    for (int d = 0; d < 3; d++) g[12 + d] = *fld(gout.vel, bv, k, d);
    g[15] = *fld(gout.temp, col_base(gout.temp, ci, cj, cb), k, 0);
    g[16] = *fld(gout.rho, col_base(gout.rho, ci, cj, cb), k, 0);
    for (int q = 0; q < 17; q++)
      if (out[q].i >= 0 && g[q] != 0.0) cx.add(out[q].i, g[q]);
  }
}

// After the backward sweep: the adjoints of the lane's 17 inputs (its nodes
// 0..16; zero when the forward pass stopped at a non-positive density) to
// `gin`.  `adjoint(id)` reads a swept adjoint.
template <class Load>
GV_FN void adjoint_store(const AdTape &t, int gid, const gitm_vert_columns &io,
                         const AdjFields &gin, int col, int k, Load adjoint) {
  const int nlb = io.n_lons * io.n_lats;
  const int ci = col % io.n_lons, cj = (col / io.n_lons) % io.n_lats, cb = col / nlb;
  const long long bn = col_base(gin.nds, ci, cj, cb);
  const long long bw = col_base(gin.vvel, ci, cj, cb);
  const long long bv = col_base(gin.vel, ci, cj, cb);
  double g[17];
  for (int q = 0; q < 17; q++) g[q] = adjoint((long long)q * t.stride + gid);
  for (int s = 0; s < NSP; s++) {
    *fld(gin.nds, bn, k, s) = g[s];
    *fld(gin.vvel, bw, k, s) = g[6 + s];
  }
  for (int d = 0; d < 3; d++) *fld(gin.vel, bv, k, d) = g[12 + d];
  *fld(gin.temp, col_base(gin.temp, ci, cj, cb), k, 0) = g[15];
  *fld(gin.rho, col_base(gin.rho, ci, cj, cb), k, 0) = g[16];
}

// Forward pass, sweep and store in one (the host build): Ctx also provides
// sweep() and adjoint(id).
template <int NT, class Ctx>
GV_FN void adjoint_column(Ctx &cx, const KParams &kp, const gitm_vert_columns &io,
                          const AdjFields &gout, const AdjFields &gin, int col,
                          adm::ad *lds, double *scratch, double dt, int err_col) {
  adjoint_forward<NT>(cx, kp, io, gout, col, lds, scratch, dt, err_col);
  cx.sweep();
  adjoint_store(ad_tape(), ad_gid(), io, gin, col, cx.lane() - 1,
                [&](long long id) { return cx.adjoint((int)id); });
}

}  // namespace gv
