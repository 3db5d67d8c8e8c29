// emu_vert.cpp -- TEST INFRASTRUCTURE ONLY.
//
// Host build of the column kernels in gitm_amd/csrc/vert_column.h: each
// workgroup thread runs as a CPU thread, __syncthreads() is a std::barrier
// and the LDS block is an ordinary array.  With glibc's exp/log on both
// sides this checks the kernel logic (indexing, exchanges, operation order)
// against the CPU oracle bit for bit, without a GPU.  Nothing in the
// product library loads this code.
#define GV_HOST_EMULATION 1
#include <algorithm>
#include <atomic>
#include <barrier>
#include <memory>
#include <thread>
#include <vector>

// GV_EMU_FORWARD_ONLY leaves out the adjoint builds and the operation count
// (tests/emu/sanitize_main.cpp compiles only the forward step).
#ifndef GV_EMU_FORWARD_ONLY
#include "count.h"
#include "vert_ad.h"
#include "vert_adj.h"
#endif
#include "vert_column.h"
#include "vert_params.h"

namespace {

struct Shared {
  std::barrier<> bar;
  std::atomic<int> orflag{0};
  std::atomic<unsigned long long> err{~0ULL};
  bool device_math;  // the GPU's forms (exp(e ln T), log-space sums) instead of the reference's
  Shared(int n, bool dm) : bar(n), device_math(dm) {}
};

struct EmuCtx {
  using real = double;
  int lane_;
  Shared *sh;
  int lane() const { return lane_; }
  void sync() { sh->bar.arrive_and_wait(); }
  bool sync_or(bool p) {
    if (p) sh->orflag.store(1);
    sync();
    bool r = sh->orflag.load() != 0;
    sync();
    if (lane_ == 0) sh->orflag.store(0);
    sync();
    return r;
  }
  void report(unsigned long long key) {
    unsigned long long cur = sh->err.load();
    while (key < cur && !sh->err.compare_exchange_weak(cur, key)) {
    }
  }
  bool device_math() const { return sh->device_math; }
  void diff_powers(double temp, double sign, double (&out)[6]) const {
    if (sh->device_math) {
      const double lt = log(temp);
      for (int u = 0; u < 6; u++) out[u] = exp((sign * gv::DEXU[u]) * lt);
    } else {
      for (int u = 0; u < 6; u++) out[u] = pow(temp, sign * gv::DEXU[u]);
    }
  }
};

template <int NT, bool IONS>
void run_one(EmuCtx &cx, const gv::KParams &kp, const gitm_vert_columns &io, int c,
             double *lds) {
  if (kp.solver == GITM_VERT_AUSM)
    gv::run_column<NT, GITM_VERT_AUSM, IONS>(cx, kp, io, c, lds);
  else
    gv::run_column<NT, GITM_VERT_RUSANOV, IONS>(cx, kp, io, c, lds);
}

template <int NT>
unsigned long long run_all(const gv::KParams &kp, const gitm_vert_columns &io, int c0,
                           int c1, bool device_math) {
  const int nthreads = kp.n_alts + 4;  // one thread per level
  Shared sh(nthreads, device_math);
  // exactly the dynamic LDS a launch gets, so out-of-range indexing shows up
  // under the host sanitizers
  std::vector<double> lds(gv::lds_doubles(kp.n_alts, kp.ions != 0, gv::carry_in_lds(NT / 64)),
                          0.0);
  std::vector<std::thread> th;
  for (int t = 0; t < nthreads; t++)
    th.emplace_back([&, t] {
      EmuCtx cx{t, &sh};
      for (int c = c0; c < c1; c++) {
        if (kp.ions)
          run_one<NT, true>(cx, kp, io, c, lds.data());
        else
          run_one<NT, false>(cx, kp, io, c, lds.data());
        cx.sync();
      }
    });
  for (auto &x : th) x.join();
  return sh.err.load();
}

#ifndef GV_EMU_FORWARD_ONLY
// The adjoint build: taped arithmetic (vert_ad.h), lanes as CPU threads.
struct AdShared {
  std::barrier<> bar;
  std::atomic<int> orflag{0};
  std::atomic<unsigned long long> err{~0ULL};
  gv::AdTape tape;
  explicit AdShared(int n) : bar(n) {}
};

struct EmuAdCtx {
  using real = gv::adm::ad;
  int lane_;
  AdShared *sh;
  int nph = 0;
  int lane() const { return lane_; }
  void boundary() {
    const gv::AdTape &t = sh->tape;
    if (nph >= t.maxph)
      *t.overflow = 1;
    else
      t.phase[(long long)nph * t.stride + lane_] = gv::ad_count();
    nph++;
  }
  void sync() {
    boundary();
    sh->bar.arrive_and_wait();
  }
  bool sync_or(bool p) {
    boundary();
    if (p) sh->orflag.store(1);
    sh->bar.arrive_and_wait();
    bool r = sh->orflag.load() != 0;
    sh->bar.arrive_and_wait();
    if (lane_ == 0) sh->orflag.store(0);
    sh->bar.arrive_and_wait();
    return r;
  }
  void report(unsigned long long key) {
    unsigned long long cur = sh->err.load();
    while (key < cur && !sh->err.compare_exchange_weak(cur, key)) {
    }
  }
  bool device_math() const { return false; }
  void diff_powers(real temp, double sign, real (&out)[6]) const {
    for (int u = 0; u < 6; u++) out[u] = pow(temp, sign * gv::DEXU[u]);
  }
  void end_forward() { boundary(); }
  void add(int id, double v) { std::atomic_ref<double>(sh->tape.adj[id]).fetch_add(v); }
  double adjoint(int id) {
    return id >= 0 ? std::atomic_ref<double>(sh->tape.adj[id]).load() : 0.0;
  }
  void sweep() {
    const gv::AdTape &t = sh->tape;
    sh->bar.arrive_and_wait();  // every lane's seeds are in
    const int np = nph < t.maxph ? nph : t.maxph;
    for (int ph = np - 1; ph >= 0; ph--) {
      const int hi = t.phase[(long long)ph * t.stride + lane_];
      const int lo = ph ? t.phase[(long long)(ph - 1) * t.stride + lane_] : 0;
      for (int n = hi - 1; n >= lo; n--) {
        const long long id = (long long)n * t.stride + lane_;
        const double a = std::atomic_ref<double>(t.adj[id]).load();
        if (a == 0.0) continue;
        const gv::AdPar pp = t.p[id];
        const gv::AdDer dd = t.d[id];
        if (pp.a >= 0) add(pp.a, a * dd.a);
        if (pp.b >= 0) add(pp.b, a * dd.b);
      }
      sh->bar.arrive_and_wait();
    }
  }
};
#endif  // GV_EMU_FORWARD_ONLY

}  // namespace

#ifndef GV_EMU_FORWARD_ONLY
// Adjoint of one AUSM step (no ions) of columns c0..c1-1 of `c`: gout holds
// the adjoints of the step's outputs, gin receives those of its inputs.
// nmax/maxph size the tape of one column; the numbers of nodes and phases
// a column used are returned through used_nodes/used_phases.
extern "C" int gv_emu_adjoint(const gitm_vert_params *p, const gitm_vert_columns *c,
                              const gv::AdjFields *gout, const gv::AdjFields *gin, int c0,
                              int c1, int nmax, int maxph, int *used_nodes, int *used_phases,
                              gitm_vert_error *e) {
  int rc = gv::check_args(p, c, GITM_VERT_AUSM, 1);
  if (rc) return rc;
  if (p->use_ion_advection) return GITM_VERT_EUNSUPPORTED;
  gv::KParams kp = gv::make_kparams(p, c, GITM_VERT_AUSM, 1, nullptr);
  const int waves = gv::auto_waves(c->n_alts);
  kp.scr_fields = gv::scratch_fields(gv::carry_in_lds(waves), false);
  const int L = c->n_alts + 4;
  std::vector<gv::adm::ad> scratch((size_t)(c1 > 0 ? c1 : 1) * kp.scr_fields * L + 1);
  AdShared sh(L);
  std::vector<gv::AdDer> der((size_t)nmax * L);
  std::vector<gv::AdPar> par((size_t)nmax * L);
  std::vector<double> adj((size_t)nmax * L);
  std::vector<int> phase((size_t)maxph * L), nphase(L);
  int overflow = 0;
  sh.tape = gv::AdTape{der.data(), par.data(), adj.data(), phase.data(),
                       nphase.data(), L, nmax, maxph, &overflow};
  std::vector<gv::adm::ad> lds(gv::lds_doubles(c->n_alts, false, gv::carry_in_lds(waves)));
  std::atomic<int> max_nodes{0}, max_phases{0};
  std::vector<std::thread> th;
  for (int t = 0; t < L; t++)
    th.emplace_back([&, t] {
      gv::ad_lane.t = &sh.tape;
      gv::ad_lane.gid = t;
      for (int col = c0; col < c1; col++) {
        EmuAdCtx cx{t, &sh};
        gv::ad_lane.n = 0;
        // the lane's adjoints start at zero (before its first barrier; the
        // previous column's adds all ended at the barrier closing it)
        for (int n = 0; n < nmax; n++) adj[(size_t)n * L + t] = 0.0;
        auto run = [&](auto nt) {
          gv::adjoint_column<decltype(nt)::value>(cx, kp, *c, *gout, *gin, col, lds.data(),
                                                   reinterpret_cast<double *>(scratch.data()),
                                                   kp.dt, col);
        };
        switch (waves) {
          case 1: run(std::integral_constant<int, 64>()); break;
          case 2: run(std::integral_constant<int, 128>()); break;
          case 3: run(std::integral_constant<int, 192>()); break;
          default: run(std::integral_constant<int, 256>()); break;
        }
        int v = max_nodes.load();
        while (gv::ad_lane.n > v && !max_nodes.compare_exchange_weak(v, gv::ad_lane.n)) {
        }
        v = max_phases.load();
        while (cx.nph > v && !max_phases.compare_exchange_weak(v, cx.nph)) {
        }
        sh.bar.arrive_and_wait();
      }
    });
  for (auto &x : th) x.join();
  if (used_nodes) *used_nodes = max_nodes.load();
  if (used_phases) *used_phases = max_phases.load();
  gitm_vert_error tmp;
  if (!e) e = &tmp;
  gv::decode_error(sh.err.load(), e);
  if (overflow && e->code == GITM_VERT_OK) e->code = GITM_VERT_ENOMEM;
  return e->code;
}

// The hand-written adjoint (vert_adj.h) of one AUSM step of columns c0..c1-1,
// lanes as CPU threads: gout holds the adjoints of the step's outputs, gin
// receives those of its inputs, gfix (may be NULL) those of the column-prep
// inputs.  device_math as for gv_emu_advance.
extern "C" int gv_emu_adjoint_hand(const gitm_vert_params *p, const gitm_vert_columns *c,
                                   const gv::AdjFields *gout, const gv::AdjFields *gin,
                                   const gv::AdjFixFields *gfix, int c0, int c1,
                                   int device_math, gitm_vert_error *e) {
  int rc = gv::check_args(p, c, GITM_VERT_AUSM, 1);
  if (rc) return rc;
  if (p->use_ion_advection || p->use_msis) return GITM_VERT_EUNSUPPORTED;
  gv::KParams kp = gv::make_kparams(p, c, GITM_VERT_AUSM, 1, nullptr);
  const int waves = gv::auto_waves(c->n_alts);
  if (waves > 3) return GITM_VERT_EUNSUPPORTED;
  kp.scr_fields = 0;
  const int L = c->n_alts + 4;
  Shared sh(L, device_math != 0);
  std::vector<double> lds(gv::adj_lds_bytes(c->n_alts, true) / sizeof(double), 0.0);
  std::vector<double> rec((size_t)gv::adj_column_doubles(64 * waves), 0.0);
  std::vector<std::thread> th;
  for (int t = 0; t < L; t++)
    th.emplace_back([&, t] {
      for (int col = c0; col < c1; col++) {
        gv::TapCtx<EmuCtx> cx;
        cx.lane_ = t;
        cx.sh = &sh;
        switch (waves) {
          case 1:
            gv::adj_column<64>(cx, kp, *c, *gout, *gin, gfix, col, lds.data(), rec.data(), col);
            break;
          case 2:
            gv::adj_column<128>(cx, kp, *c, *gout, *gin, gfix, col, lds.data(), rec.data(), col);
            break;
          default:
            gv::adj_column<192>(cx, kp, *c, *gout, *gin, gfix, col, lds.data(), rec.data(), col);
            break;
        }
        sh.bar.arrive_and_wait();
      }
    });
  for (auto &x : th) x.join();
  gitm_vert_error tmp;
  if (!e) e = &tmp;
  gv::decode_error(sh.err.load(), e);
  return e->code;
}

// The operation count of the column step: the column code run over the
// counting number type (count.h), lanes as CPU threads.
struct EmuCntCtx {
  using real = cnt::fd;
  int lane_;
  Shared *sh;
  int lane() const { return lane_; }
  void sync() { sh->bar.arrive_and_wait(); }
  bool sync_or(bool p) {
    if (p) sh->orflag.store(1);
    sync();
    bool r = sh->orflag.load() != 0;
    sync();
    if (lane_ == 0) sh->orflag.store(0);
    sync();
    return r;
  }
  void report(unsigned long long) {}
  bool device_math() const { return sh->device_math; }
  void diff_powers(real temp, double sign, real (&out)[6]) const {
    if (sh->device_math) {
      const real lt = log(temp);
      for (int u = 0; u < 6; u++) out[u] = exp((sign * gv::DEXU[u]) * lt);
    } else {
      for (int u = 0; u < 6; u++) out[u] = pow(temp, sign * gv::DEXU[u]);
    }
  }
};

// One column's AUSM (solver 1) or Rusanov (0) step with gather and scatter
// as run_column does them; out[6] = the summed counts (add, mul, div, sqrt,
// trans, cmp) of all lanes, lanes_active[0] = the number of lanes.
extern "C" int gv_emu_count_ops(const gitm_vert_params *p, const gitm_vert_columns *c,
                                int solver, int col, int device_math, long long *out) {
  int rc = gv::check_args(p, c, solver, 1);
  if (rc) return rc;
  if (p->use_ion_advection || p->use_msis) return GITM_VERT_EUNSUPPORTED;
  gv::KParams kp = gv::make_kparams(p, c, solver, 1, nullptr);
  const int waves = gv::auto_waves(c->n_alts);
  if (waves > 3) return GITM_VERT_EUNSUPPORTED;
  kp.scr_fields = 0;
  const int L = c->n_alts + 4;
  Shared sh(L, device_math != 0);
  std::vector<cnt::fd> lds(gv::lds_doubles(c->n_alts, false, true));
  std::vector<cnt::Counts> per(L);
  std::vector<std::thread> th;
  for (int t = 0; t < L; t++)
    th.emplace_back([&, t] {
      cnt::tally = cnt::Counts();
      EmuCntCtx cx{t, &sh};
      auto run = [&](auto nt) {
        constexpr int NT = decltype(nt)::value;
        using cnt::fd;
        gv::Column<NT, EmuCntCtx, false> cl(cx, kp, lds.data(), col, fd(kp.dt));
        cl.scr = nullptr;
        const int nlb = c->n_lons * c->n_lats;
        const int ci = col % c->n_lons, cj = (col / c->n_lons) % c->n_lats, cb = col / nlb;
        const int k = cl.k, nA = cl.nA;
        auto ld = [&](const gitm_vert_field &f, int comp) {
          return *gv::fld(f, gv::col_base(f, ci, cj, cb), k, comp);
        };
        const double coslat = c->cos_latitude.ptr[gv::col_base(c->cos_latitude, ci, cj, cb)];
        const fd tu = ld(c->temp_unit, 0);
        cl.setup(ld(c->altitude, 0), ld(c->gravity, 0), ld(c->inv_radial_distance, 0),
                 ld(c->dalt, 0), ld(c->mean_major_mass, 0), ld(c->gamma, 0),
                 (k >= 1 && k <= nA) ? ld(c->eddy, 0) : 0.0,
                 (k >= 0 && k <= nA + 1) ? ld(c->cmax, 0) : 0.0, tu, coslat);
        // gather (advance_vertical.f90:105-118)
        for (int s = 0; s < gv::NSP; s++) {
          cl.U[gv::F_LN + s] = log(fd(ld(c->ndensity_s, s)));
          cl.U[gv::F_W + s] = ld(c->vertical_velocity, s);
        }
        for (int d = 0; d < 3; d++) cl.U[gv::F_VE + d] = ld(c->velocity, d);
        cl.U[gv::F_T] = fd(ld(c->temperature, 0)) * tu;
        cl.U[gv::F_LR] = log(fd(ld(c->rho, 0)));
        if (solver == GITM_VERT_AUSM)
          cl.ausm_step();
        else
          cl.rusanov_step();
        // scatter (:163-185)
        fd nsum = 0.0;
        for (int s = 0; s < gv::NSP; s++) nsum = nsum + exp(cl.U[gv::F_LN + s]);
        fd r = exp(cl.U[gv::F_LR]);
        fd t2 = cl.U[gv::F_T] / tu;
        (void)r;
        (void)t2;
      };
      switch (waves) {
        case 1: run(std::integral_constant<int, 64>()); break;
        case 2: run(std::integral_constant<int, 128>()); break;
        default: run(std::integral_constant<int, 192>()); break;
      }
      per[t] = cnt::tally;
    });
  for (auto &x : th) x.join();
  long long sum[6] = {0, 0, 0, 0, 0, 0};
  for (const cnt::Counts &q : per) {
    sum[0] += q.add;
    sum[1] += q.mul;
    sum[2] += q.div;
    sum[3] += q.sqrt;
    sum[4] += q.trans;
    sum[5] += q.cmp;
  }
  for (int i = 0; i < 6; i++) out[i] = sum[i];
  return GITM_VERT_OK;
}

#endif  // GV_EMU_FORWARD_ONLY

// device_math != 0 evaluates as the GPU does (friction powers as exp(e ln T),
// the log-space forms of Column); 0 uses the reference's forms, for bitwise
// checks against the oracle.
extern "C" int gv_emu_advance(const gitm_vert_params *p, const gitm_vert_columns *c,
                              int solver, int n_steps, int waves, int c0, int c1,
                              int device_math, gitm_vert_error *e) {
  int rc = gv::check_args(p, c, solver, n_steps);
  if (rc) return rc;
  gv::KParams kp = gv::make_kparams(p, c, solver, n_steps, nullptr);
  if (waves <= 0) waves = gv::auto_waves(c->n_alts);
  if (64 * waves < c->n_alts + 4) return GITM_VERT_EINVAL;
  kp.scr_fields = gv::scratch_fields(gv::carry_in_lds(waves), kp.ions != 0);
  std::vector<double> scratch((size_t)c->n_lons * c->n_lats * c->n_blocks * kp.scr_fields *
                              (c->n_alts + 4));
  kp.scratch = scratch.data();
  unsigned long long key;
  switch (waves) {
    case 1: key = run_all<64>(kp, *c, c0, c1, device_math != 0); break;
    case 2: key = run_all<128>(kp, *c, c0, c1, device_math != 0); break;
    case 3: key = run_all<192>(kp, *c, c0, c1, device_math != 0); break;
    case 4: key = run_all<256>(kp, *c, c0, c1, device_math != 0); break;
    default: return GITM_VERT_EINVAL;
  }
  gitm_vert_error tmp;
  if (!e) e = &tmp;
  gv::decode_error(key, e);
  return e->code;
}

// The short-column path of the ensemble step (gitm_vert_ensemble.hip): the
// columns [c0, c1), per_wg at a time, share one LDS block, each in a slot of
// its own (slot_doubles: the column's LDS rounded up to even), and each
// synchronises only its own lanes, as a wavefront's WaveCtx does on the GPU.
// carry_lds: the RK4 carries in the slot (1) or in scratch (0).
template <int SOLVER, bool IONS, bool CARRY>
unsigned long long run_slots(const gv::KParams &kp, const gitm_vert_columns &io, int c0,
                             int c1, int per_wg, bool device_math, double *scratch) {
  const int L = kp.n_alts + 4;
  int sd = (int)gv::lds_doubles(kp.n_alts, IONS, CARRY);
  sd += sd & 1;
  std::vector<double> lds((size_t)per_wg * sd, 0.0);
  unsigned long long key = ~0ULL;
  for (int g = c0; g < c1; g += per_wg) {
    const int n = std::min(per_wg, c1 - g);
    std::vector<std::unique_ptr<Shared>> sh;
    for (int w = 0; w < n; w++) sh.push_back(std::make_unique<Shared>(L, device_math));
    std::vector<std::thread> th;
    for (int w = 0; w < n; w++)
      for (int t = 0; t < L; t++)
        th.emplace_back([&, w, t] {
          EmuCtx cx{t, sh[w].get()};
          gv::run_column<64, SOLVER, IONS, CARRY>(cx, kp, io, g + w, lds.data() + (size_t)w * sd,
                                                  scratch, kp.dt, g + w, kp.n_alts,
                                                  gv::scratch_fields(CARRY, IONS));
        });
    for (auto &x : th) x.join();
    for (auto &s : sh) key = std::min(key, s->err.load());
  }
  return key;
}

extern "C" int gv_emu_advance_slots(const gitm_vert_params *p, const gitm_vert_columns *c,
                                    int solver, int n_steps, int per_wg, int carry_lds,
                                    int c0, int c1, int device_math, gitm_vert_error *e) {
  int rc = gv::check_args(p, c, solver, n_steps);
  if (rc) return rc;
  if (c->n_alts + 4 > 64 || per_wg < 1 || per_wg > 4) return GITM_VERT_EINVAL;
  gv::KParams kp = gv::make_kparams(p, c, solver, n_steps, nullptr);
  const bool ions = kp.ions != 0, carry = carry_lds != 0;
  std::vector<double> scratch((size_t)c->n_lons * c->n_lats * c->n_blocks *
                              gv::scratch_fields(carry, ions) * (c->n_alts + 4));
  const bool dm = device_math != 0;
  double *sc = scratch.data();
  unsigned long long key;
#define GV_SLOTS(S, I, C) key = run_slots<S, I, C>(kp, *c, c0, c1, per_wg, dm, sc)
  if (solver == GITM_VERT_AUSM) {
    if (ions) { if (carry) GV_SLOTS(GITM_VERT_AUSM, true, true); else GV_SLOTS(GITM_VERT_AUSM, true, false); }
    else { if (carry) GV_SLOTS(GITM_VERT_AUSM, false, true); else GV_SLOTS(GITM_VERT_AUSM, false, false); }
  } else {
    if (ions) { if (carry) GV_SLOTS(GITM_VERT_RUSANOV, true, true); else GV_SLOTS(GITM_VERT_RUSANOV, true, false); }
    else { if (carry) GV_SLOTS(GITM_VERT_RUSANOV, false, true); else GV_SLOTS(GITM_VERT_RUSANOV, false, false); }
  }
#undef GV_SLOTS
  gitm_vert_error tmp;
  if (!e) e = &tmp;
  gv::decode_error(key, e);
  return e->code;
}

extern "C" void gv_emu_default_params(gitm_vert_params *p) { gv::default_params(p); }
