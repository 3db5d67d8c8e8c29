// emu_vert.cpp -- TEST INFRASTRUCTURE ONLY.
//
// Host build of the column kernels in gitm_amd/csrc/vert_column.h: each
// workgroup thread runs as a CPU thread, __syncthreads() is a std::barrier
// and the LDS block is an ordinary array.  With glibc's exp/log on both
// sides this checks the kernel logic (indexing, exchanges, operation order)
// against the CPU oracle bit for bit, without a GPU.  Nothing in the
// product library loads this code.
#define GV_HOST_EMULATION 1
#include <atomic>
#include <barrier>
#include <thread>
#include <vector>

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

}  // namespace

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

extern "C" void gv_emu_default_params(gitm_vert_params *p) { gv::default_params(p); }
