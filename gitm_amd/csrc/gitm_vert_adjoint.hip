// gitm_vert_adjoint.hip -- the adjoint of the AUSM step: the hand-written
// adjoint (gitm_vert_adjoint, gitm_vert_adjoint_all; vert_adj.h), the
// adjoint of the column prep (gitm_vert_prepare_adjoint) and the taped
// adjoint (gitm_vert_adjoint_taped; vert_ad.h), which checks the former.
#include "vert_ad.h"
#include "vert_adj.h"
#include "vert_gpu.h"

namespace {

// The adjoint build (vert_ad.h): the column code over the taped number type.
// Every barrier closes a phase of the tape; sweep() runs the phases
// backwards with a barrier between them.
struct GpuAdCtx {
  using real = gv::adm::ad;
  unsigned long long *err;
  int nph = 0;
  __device__ int lane() const { return (int)threadIdx.x; }
  __device__ void boundary() {
    const gv::AdTape &t = gv::ad_tape();
    if (nph >= t.maxph)
      *t.overflow = 1;
    else
      t.phase[(long long)nph * t.stride + gv::ad_gid()] = gv::ad_count();
    nph++;
  }
  __device__ void sync() {
    boundary();
    __syncthreads();
  }
  __device__ bool sync_or(bool p) {
    boundary();
    return __syncthreads_or(p ? 1 : 0) != 0;
  }
  __device__ void report(unsigned long long key) const { atomicMin(err, key); }
  __device__ static constexpr bool device_math() { return false; }
  __device__ void diff_powers(real temp, double sign, real (&out)[6]) const {
#pragma unroll
    for (int u = 0; u < 6; u++) out[u] = pow(temp, sign * gv::DEXU[u]);
  }
  __device__ void end_forward() { boundary(); }
  __device__ void add(int id, double v) { unsafeAtomicAdd(gv::ad_tape().adj + id, v); }
  // adjoints are summed in L2 by atomics: read them past the CU's L1
  __device__ double adjoint(int id) const {
    return id >= 0 ? __hip_atomic_load(gv::ad_tape().adj + id, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT)
                   : 0.0;
  }
  __device__ void sweep() {
    const gv::AdTape &t = gv::ad_tape();
    const int g = gv::ad_gid();
    __threadfence();
    __syncthreads();  // every lane's seeds are in
    const int np = nph < t.maxph ? nph : t.maxph;
    for (int ph = np - 1; ph >= 0; ph--) {
      const int hi = t.phase[(long long)ph * t.stride + g];
      const int lo = ph ? t.phase[(long long)(ph - 1) * t.stride + g] : 0;
      for (int n = hi - 1; n >= lo; n--) {
        const long long id = (long long)n * t.stride + g;
        const double a = adjoint((int)id);
        if (a == 0.0) continue;
        const gv::AdPar pp = t.p[id];
        const gv::AdDer dd = t.d[id];
        if (pp.a >= 0) add(pp.a, a * dd.a);
        if (pp.b >= 0) add(pp.b, a * dd.b);
      }
      __threadfence();
      __syncthreads();
    }
  }
};

// Adjoint of one AUSM step of columns col0 .. col0+gridDim.x-1, forward
// part: one workgroup per column, one lane per level, as the forward kernel.
template <int NT>
__global__ void __launch_bounds__(NT)
    vert_adjoint_fwd_kernel(gv::KParams kp, gitm_vert_columns io, gv::AdjFields gout,
                            gv::AdTape tape, int col0) {
  extern __shared__ double lds[];
  if (threadIdx.x == 0) gv::ad_tape_sh = tape;
  gv::ad_count_sh[threadIdx.x] = 0;
  __syncthreads();
  GpuAdCtx cx{kp.err};
  const int col = col0 + (int)blockIdx.x;
  gv::adjoint_forward<NT>(cx, kp, io, gout, col, reinterpret_cast<gv::adm::ad *>(lds),
                          kp.scratch, kp.dt, col);
  tape.nphase[gv::ad_gid()] = cx.nph < tape.maxph ? cx.nph : tape.maxph;
}

// Backward sweep of the tape the forward part wrote, then the input
// adjoints.  It needs no LDS, so many more columns are in flight than in
// the forward part: the sweep is a chain of dependent memory accesses per
// lane, and its speed is the number of lanes that wait at once.
template <int NT>
__global__ void __launch_bounds__(NT)
    vert_adjoint_sweep_kernel(gitm_vert_columns io, gv::AdjFields gin, gv::AdTape t,
                              int col0) {
  const int g = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  const int nph = t.nphase[g];  // the same on every lane of the column
  double *adj = t.adj;
  // Nodes go in blocks of SB, last to first: the block's tape entries and
  // adjoints are loaded at once, a contribution to a parent inside the
  // block (the lane's own node of this phase, which no other lane touches
  // until the barrier) is added in registers, the others atomically.  A
  // node's adjoint is complete when it is reached: its children come after
  // it, inside the block (already in registers) or above it (their
  // atomics were issued before this block's loads).
  constexpr int SB = 16;
  for (int ph = nph - 1; ph >= 0; ph--) {
    const int hi = t.phase[(long long)ph * t.stride + g];
    const int lo = ph ? t.phase[(long long)(ph - 1) * t.stride + g] : 0;
    for (int top = hi - 1; top >= lo; top -= SB) {
      const int cnt = min(SB, top - lo + 1);  // nodes top, top-1, .., top-cnt+1
      double a[SB], d1[SB], d2[SB];
      int p1[SB], p2[SB];
#pragma unroll
      for (int j = 0; j < SB; j++) {
        a[j] = 0.0;
        p1[j] = p2[j] = -1;
        d1[j] = d2[j] = 0.0;
        if (j < cnt) {
          const long long id = (long long)(top - j) * t.stride + g;
          const gv::AdPar pp = t.p[id];
          const gv::AdDer dd = t.d[id];
          p1[j] = pp.a;
          p2[j] = pp.b;
          d1[j] = dd.a;
          d2[j] = dd.b;
          a[j] = __hip_atomic_load(adj + id, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
      }
      const long long base = (long long)(top - cnt + 1) * t.stride + g;  // lowest node's id
      const long long last = (long long)top * t.stride + g;
#pragma unroll
      for (int j = 0; j < SB; j++) {
        const double aj = a[j];
        if (aj == 0.0) continue;
        auto push = [&](int p, double d) {
          if (p < 0) return;
          const double v = aj * d;
          // inside the block: node top - k for id = last - k * stride
          if (p >= base && p < last && (p - base) % t.stride == 0) {
            const int k = (int)((last - p) / t.stride);
#pragma unroll
            for (int q = 0; q < SB; q++)
              if (q == k) a[q] = a[q] + v;
          } else {
            unsafeAtomicAdd(adj + p, v);
          }
        };
        push(p1[j], d1[j]);
        push(p2[j], d2[j]);
      }
    }
    __threadfence();
    __syncthreads();
  }
  gv::adjoint_store(t, g, io, gin, col0 + (int)blockIdx.x, (int)threadIdx.x - 1,
                    [&](long long id) {
                      return __hip_atomic_load(adj + id, __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT);
                    });
}

}  // namespace

// ---------------------------------------------------------------------------
// Adjoint of one AUSM step (vert_ad.h).  The columns run in chunks whose
// tapes fit the handle's adjoint memory budget.
// ---------------------------------------------------------------------------
namespace {

// Default tape budget: 40% of the device's memory, at most 96 GiB (the
// 5-degree grid's 2592 columns at nAlts = 50 take 58 GB, one chunk; the sweep
// needs many columns in flight to hide its memory latency).
size_t default_adjoint_budget(int device) {
  hipDeviceProp_t prop;
  size_t total = (size_t)16 << 30;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) total = prop.totalGlobalMem;
  return std::min((size_t)(0.4 * (double)total), (size_t)96 << 30);
}

// Tape entries per lane: about 10k nodes per level-lane and step, plus the
// serial loops of the boundary lanes, which grow with the column.
int adjoint_nmax(int n_alts) { return 12288 + 8 * n_alts; }
constexpr int kAdjMaxPhases = 128;
constexpr size_t kNodeBytes = sizeof(gv::AdDer) + sizeof(double) + sizeof(gv::AdPar);

// The tape's adjoints start at zero: one coalesced fill per chunk, rather
// than a store per node while taping.
__global__ void __launch_bounds__(256) clear_adjoints(double *adj, long long n) {
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (long long)gridDim.x * 256)
    adj[i] = 0.0;
}

template <int NT>
void launch_adjoint(int ncols, int col0, hipStream_t stream, const gv::KParams &kp,
                    const gitm_vert_columns &c, const gv::AdjFields &go,
                    const gv::AdjFields &gi, const gv::AdTape &tape) {
  const size_t lds =
      gv::lds_doubles(kp.n_alts, false, gv::carry_in_lds(NT / 64)) * sizeof(gv::adm::ad);
  hipLaunchKernelGGL(vert_adjoint_fwd_kernel<NT>, dim3(ncols), dim3(kp.n_alts + 4), lds,
                     stream, kp, c, go, tape, col0);
  hipLaunchKernelGGL(vert_adjoint_sweep_kernel<NT>, dim3(ncols), dim3(kp.n_alts + 4), 0,
                     stream, c, gi, tape, col0);
}

}  // namespace

extern "C" {

int gitm_vert_set_adjoint_memory(gitm_vert_handle *h, size_t bytes) {
  if (!h || (bytes != 0 && bytes < ((size_t)64 << 20))) return GITM_VERT_EINVAL;
  h->adj_budget = bytes;
  // a buffer larger than the new budget is given back at once
  const size_t cap = bytes ? bytes : (size_t)-1;
  if ((h->tape && h->tape_bytes > cap) || (h->rec && h->rec_bytes > cap)) {
    GV_HIP(hipSetDevice(h->device));
    GV_HIP(hipStreamSynchronize(h->stream));
    if (h->tape && h->tape_bytes > cap) {
      (void)hipFree(h->tape);
      h->tape = nullptr;
      h->tape_bytes = 0;
    }
    if (h->rec && h->rec_bytes > cap) {
      (void)hipFree(h->rec);
      h->rec = nullptr;
      h->rec_bytes = 0;
    }
  }
  return GITM_VERT_OK;
}

int gitm_vert_release_adjoint_memory(gitm_vert_handle *h) {
  if (!h) return GITM_VERT_EINVAL;
  GV_HIP(hipSetDevice(h->device));
  GV_HIP(hipStreamSynchronize(h->stream));
  if (h->tape) (void)hipFree(h->tape);
  if (h->rec) (void)hipFree(h->rec);
  h->tape = nullptr;
  h->rec = nullptr;
  h->tape_bytes = h->rec_bytes = 0;
  return GITM_VERT_OK;
}

int gitm_vert_adjoint_taped(gitm_vert_handle *h, const gitm_vert_params *p,
                            const gitm_vert_columns *c, const gitm_vert_adjoint_fields *gout,
                            const gitm_vert_adjoint_fields *gin) {
  if (!h || !gout || !gin) return GITM_VERT_EINVAL;
  int rc = gv::check_args(p, c, GITM_VERT_AUSM, 1);
  if (rc) return rc;
  if (p->use_ion_advection || p->use_msis || p->use_gswm_tides || p->use_waccm_tides)
    return GITM_VERT_EUNSUPPORTED;
  const gitm_vert_field *req[] = {&gout->ndensity_s, &gout->vertical_velocity, &gout->velocity,
                                  &gout->temperature, &gout->rho, &gin->ndensity_s,
                                  &gin->vertical_velocity, &gin->velocity, &gin->temperature,
                                  &gin->rho};
  for (const gitm_vert_field *f : req)
    if (!f->ptr) return GITM_VERT_EINVAL;
  const int waves = gv::auto_waves(c->n_alts);
  if (waves > 2) return GITM_VERT_EUNSUPPORTED;  // the taped LDS holds up to 124 levels
  GV_HIP(hipSetDevice(h->device));
  const int L = c->n_alts + 4;
  const long long ncols = (long long)c->n_lons * c->n_lats * c->n_blocks;
  const int nmax = adjoint_nmax(c->n_alts);
  const size_t lane_bytes =
      (size_t)nmax * kNodeBytes + (size_t)(kAdjMaxPhases + 1) * sizeof(int);
  const size_t budget = h->adj_budget ? h->adj_budget : default_adjoint_budget(h->device);
  long long chunk = (long long)(budget / (lane_bytes * (size_t)L));
  chunk = std::min(chunk, (long long)(0x7fffffffLL / ((long long)nmax * L)));
  chunk = std::min(chunk, ncols);
  if (chunk < 1) return GITM_VERT_ENOMEM;
  const size_t stride = (size_t)chunk * L;
  const size_t need = lane_bytes * stride + 256;
  if (h->tape_bytes < need) {
    GV_HIP(hipStreamSynchronize(h->stream));
    if (h->tape) (void)hipFree(h->tape);
    h->tape = nullptr;
    h->tape_bytes = 0;
    if (hipMalloc(&h->tape, need) != hipSuccess) return GITM_VERT_ENOMEM;
    h->tape_bytes = need;
  }
  if (!h->d_overflow) {
    if (hipMalloc(&h->d_overflow, sizeof(int)) != hipSuccess) return GITM_VERT_ENOMEM;
  }
  hipLaunchKernelGGL(set_int, dim3(1), dim3(1), 0, h->stream, h->d_overflow, 0);
  char *base = static_cast<char *>(h->tape);
  const size_t nn = (size_t)nmax * stride;
  gv::AdTape tape;
  tape.d = reinterpret_cast<gv::AdDer *>(base);
  tape.adj = reinterpret_cast<double *>(tape.d + nn);
  tape.p = reinterpret_cast<gv::AdPar *>(tape.adj + nn);
  tape.phase = reinterpret_cast<int *>(tape.p + nn);
  tape.nphase = tape.phase + (size_t)kAdjMaxPhases * stride;
  tape.stride = (int)stride;
  tape.nmax = nmax;
  tape.maxph = kAdjMaxPhases;
  tape.overflow = h->d_overflow;
  gv::KParams kp = gv::make_kparams(p, c, GITM_VERT_AUSM, 1, h->d_err);
  kp.scr_fields = 0;  // the carries live in LDS (waves <= 2)
  kp.scratch = nullptr;
  gv::AdjFields go, gi;
  go.nds = gout->ndensity_s;
  go.vvel = gout->vertical_velocity;
  go.vel = gout->velocity;
  go.temp = gout->temperature;
  go.rho = gout->rho;
  gi.nds = gin->ndensity_s;
  gi.vvel = gin->vertical_velocity;
  gi.vel = gin->velocity;
  gi.temp = gin->temperature;
  gi.rho = gin->rho;
  for (long long c0 = 0; c0 < ncols; c0 += chunk) {
    const int n = (int)std::min(chunk, ncols - c0);
    const long long nclear = (long long)nn;  // node n of lane g is n * stride + g
    hipLaunchKernelGGL(clear_adjoints, dim3((unsigned)std::min<long long>(
                                           (nclear + 255) / 256, 8192)),
                       dim3(256), 0, h->stream, tape.adj, nclear);
    if (waves == 1)
      launch_adjoint<64>(n, (int)c0, h->stream, kp, *c, go, gi, tape);
    else
      launch_adjoint<128>(n, (int)c0, h->stream, kp, *c, go, gi, tape);
    GV_HIP(hipGetLastError());
  }
  int overflow = 0;
  GV_HIP(hipMemcpyAsync(&overflow, h->d_overflow, sizeof(int), hipMemcpyDeviceToHost,
                        h->stream));
  GV_HIP(hipStreamSynchronize(h->stream));
  if (overflow) return GITM_VERT_ETAPE;
  return gitm_vert_get_error(h, nullptr);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// The hand-written adjoint (vert_adj.h): one workgroup per column, one lane
// per level; the record of the forward recompute in device memory.
// ---------------------------------------------------------------------------
namespace {

template <int NT>
__global__ void __launch_bounds__(NT)
    vert_adj_kernel(gv::KParams kp, gitm_vert_columns io, gv::AdjFields gout,
                    gv::AdjFields gin, gv::AdjFixFields gfix, int with_fix, double *rec,
                    int col0) {
  extern __shared__ double lds[];
  gv::TapCtx<GpuCtx> cx;
  cx.err = kp.err;
  const int col = col0 + (int)blockIdx.x;
  double *r = rec + (size_t)blockIdx.x * (size_t)gv::adj_column_doubles(NT);
  gv::adj_column<NT>(cx, kp, io, gout, gin, with_fix ? &gfix : nullptr, col, lds, r, col);
}

// The record and work area of a column: ADJ_STAGES x NREC + NWORK doubles
// per lane.
size_t adj_record_bytes(int nt) { return (size_t)gv::adj_column_doubles(nt) * sizeof(double); }

// Default record budget: 10% of the device's memory, at most 4 GiB (the
// 1-degree grid at nAlts = 100 then runs in chunks of about 7,900 columns).
size_t default_record_budget(int device) {
  hipDeviceProp_t prop;
  size_t total = (size_t)16 << 30;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) total = prop.totalGlobalMem;
  return std::min((size_t)(0.1 * (double)total), (size_t)4 << 30);
}

template <int NT>
void launch_adj(int n, int col0, hipStream_t stream, const gv::KParams &kp,
                const gitm_vert_columns &c, const gv::AdjFields &go, const gv::AdjFields &gi,
                const gv::AdjFixFields &gf, int with_fix, double *rec) {
  const size_t lds = gv::adj_lds_bytes(kp.n_alts, true);
  hipLaunchKernelGGL(vert_adj_kernel<NT>, dim3(n), dim3(kp.n_alts + 4), lds, stream, kp, c,
                     go, gi, gf, with_fix, rec, col0);
}

gv::AdjFields adj_fields(const gitm_vert_adjoint_fields *g) {
  gv::AdjFields a;
  a.nds = g->ndensity_s;
  a.vvel = g->vertical_velocity;
  a.vel = g->velocity;
  a.temp = g->temperature;
  a.rho = g->rho;
  return a;
}

int adjoint_hand(gitm_vert_handle *h, const gitm_vert_params *p, const gitm_vert_columns *c,
                 const gitm_vert_adjoint_fields *gout, const gitm_vert_adjoint_fields *gin,
                 const gitm_vert_adjoint_prep_fields *gfix) {
  if (!h || !gout || !gin) return GITM_VERT_EINVAL;
  int rc = gv::check_args(p, c, GITM_VERT_AUSM, 1);
  if (rc) return rc;
  if (p->use_ion_advection || p->use_msis || p->use_gswm_tides || p->use_waccm_tides)
    return GITM_VERT_EUNSUPPORTED;
  const gitm_vert_field *req[] = {&gout->ndensity_s, &gout->vertical_velocity, &gout->velocity,
                                  &gout->temperature, &gout->rho, &gin->ndensity_s,
                                  &gin->vertical_velocity, &gin->velocity, &gin->temperature,
                                  &gin->rho};
  for (const gitm_vert_field *f : req)
    if (!f->ptr) return GITM_VERT_EINVAL;
  gv::AdjFixFields gf = {};
  if (gfix) {
    const gitm_vert_field *fr[] = {&gfix->temp_unit, &gfix->gamma, &gfix->mean_major_mass,
                                   &gfix->cmax, &gfix->eddy};
    for (const gitm_vert_field *f : fr)
      if (!f->ptr) return GITM_VERT_EINVAL;
    if (gfix->cmax.first_level < -1 || gfix->cmax.first_level > 0) return GITM_VERT_EINVAL;
    gf.temp_unit = gfix->temp_unit;
    gf.gamma = gfix->gamma;
    gf.mean_major_mass = gfix->mean_major_mass;
    gf.cmax = gfix->cmax;
    gf.eddy = gfix->eddy;
  }
  const int waves = gv::auto_waves(c->n_alts);
  if (!gv::carry_in_lds(waves)) return GITM_VERT_EUNSUPPORTED;  // nAlts <= 188
  GV_HIP(hipSetDevice(h->device));
  const long long ncols = (long long)c->n_lons * c->n_lats * c->n_blocks;
  const int nt = 64 * waves;
  const size_t per_col = adj_record_bytes(nt);
  const size_t budget = h->adj_budget ? h->adj_budget : default_record_budget(h->device);
  long long chunk = std::min((long long)(budget / per_col), ncols);
  if (chunk < 1) return GITM_VERT_ENOMEM;
  const size_t need = (size_t)chunk * per_col;
  if (h->rec_bytes < need) {
    GV_HIP(hipStreamSynchronize(h->stream));
    if (h->rec) (void)hipFree(h->rec);
    h->rec = nullptr;
    h->rec_bytes = 0;
    if (hipMalloc(&h->rec, need) != hipSuccess) return GITM_VERT_ENOMEM;
    h->rec_bytes = need;
  }
  gv::KParams kp = gv::make_kparams(p, c, GITM_VERT_AUSM, 1, h->d_err);
  kp.scr_fields = 0;
  kp.scratch = nullptr;
  const gv::AdjFields go = adj_fields(gout), gi = adj_fields(gin);
  double *rec = static_cast<double *>(h->rec);
  for (long long c0 = 0; c0 < ncols; c0 += chunk) {
    const int n = (int)std::min(chunk, ncols - c0);
    switch (waves) {
      case 1: launch_adj<64>(n, (int)c0, h->stream, kp, *c, go, gi, gf, gfix != nullptr, rec); break;
      case 2: launch_adj<128>(n, (int)c0, h->stream, kp, *c, go, gi, gf, gfix != nullptr, rec); break;
      default: launch_adj<192>(n, (int)c0, h->stream, kp, *c, go, gi, gf, gfix != nullptr, rec); break;
    }
    GV_HIP(hipGetLastError());
  }
  return GITM_VERT_OK;
}

// The column prep's adjoint, one thread per (column, level).
__global__ void __launch_bounds__(256)
    prep_adjoint_kernel(gv::PrepParams pp, gitm_vert_columns io,
                        gitm_vert_adjoint_prep_fields gf, gitm_vert_field gcp,
                        gitm_vert_adjoint_fields gin, long long n) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int L = pp.n_alts + 4;
  const int col = (int)(i / L), k = (int)(i % L) - 1;
  const int nlb = io.n_lons * io.n_lats;
  const int ci = col % io.n_lons, cj = (col / io.n_lons) % io.n_lats, cb = col / nlb;
  gv::prep_level_adjoint(pp, io, gf, gcp, gin, ci, cj, cb, k);
}

// The heating source's adjoint, one thread per (column, level) of L levels:
// a = lam Dt / (cp TempUnit) on the interior levels, g_heat += a, and with
// g_cp, g_tu -= a Q / TempUnit, g_cp = -(a Q / cp).
__global__ void __launch_bounds__(256)
    source_adjoint_kernel(long long n, int L, double dt, const double *lam, const double *cp,
                          const double *tu, const double *heat, double *g_heat, double *g_tu,
                          double *g_cp) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int k = (int)(i % L);
  const double a = (k >= 2 && k < L - 2) ? lam[i] * (dt / (cp[i] * tu[i])) : 0.0;
  g_heat[i] += a;
  if (g_cp) {
    const double b = a * heat[i];
    g_tu[i] -= b / tu[i];
    g_cp[i] = -(b / cp[i]);
  }
}

}  // namespace

extern "C" {

int gitm_vert_adjoint(gitm_vert_handle *h, const gitm_vert_params *p,
                      const gitm_vert_columns *c, const gitm_vert_adjoint_fields *gout,
                      const gitm_vert_adjoint_fields *gin) {
  return adjoint_hand(h, p, c, gout, gin, nullptr);
}

int gitm_vert_adjoint_all(gitm_vert_handle *h, const gitm_vert_params *p,
                          const gitm_vert_columns *c, const gitm_vert_adjoint_fields *gout,
                          const gitm_vert_adjoint_fields *gin,
                          const gitm_vert_adjoint_prep_fields *gprep) {
  if (!gprep) return GITM_VERT_EINVAL;
  return adjoint_hand(h, p, c, gout, gin, gprep);
}

int gitm_vert_prepare_adjoint(gitm_vert_handle *h, const gitm_vert_prep_params *p,
                              const gitm_vert_columns *c,
                              const gitm_vert_adjoint_prep_fields *gprep,
                              const gitm_vert_field *gcp, const gitm_vert_adjoint_fields *gin) {
  if (!h || !gprep || !gin) return GITM_VERT_EINVAL;
  int rc = check_prep(p, c);
  if (rc) return rc;
  const gitm_vert_field *req[] = {&gprep->temp_unit, &gprep->gamma, &gprep->mean_major_mass,
                                  &gprep->cmax, &gprep->eddy, &gin->ndensity_s,
                                  &gin->vertical_velocity, &gin->temperature, &gin->rho};
  for (const gitm_vert_field *f : req)
    if (!f->ptr) return GITM_VERT_EINVAL;
  if (gprep->cmax.first_level < -1 || gprep->cmax.first_level > 0) return GITM_VERT_EINVAL;
  GV_HIP(hipSetDevice(h->device));
  const gv::PrepParams pp = make_prep(p, c->n_alts);
  const gitm_vert_field none = {};
  const long long n = (long long)c->n_lons * c->n_lats * c->n_blocks * (c->n_alts + 4);
  hipLaunchKernelGGL(prep_adjoint_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     h->stream, pp, *c, *gprep, gcp ? *gcp : none, *gin, n);
  GV_HIP(hipGetLastError());
  return GITM_VERT_OK;
}

int gitm_vert_source_adjoint(gitm_vert_handle *h, long long n_cols, int n_alts, double dt,
                             const double *g_temp, const double *cp, const double *temp_unit,
                             const double *heat, double *g_heat, double *g_temp_unit,
                             double *g_cp) {
  if (!h || n_cols < 0 || n_alts < 1 || !g_temp || !cp || !temp_unit || !g_heat)
    return GITM_VERT_EINVAL;
  if ((g_temp_unit == nullptr) != (g_cp == nullptr) || (g_cp && !heat)) return GITM_VERT_EINVAL;
  const long long n = n_cols * (long long)(n_alts + 4);
  if (n_cols > ((long long)1 << 38) / (n_alts + 4)) return GITM_VERT_EINVAL;
  if (n == 0) return GITM_VERT_OK;
  GV_HIP(hipSetDevice(h->device));
  hipLaunchKernelGGL(source_adjoint_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     h->stream, n, n_alts + 4, dt, g_temp, cp, temp_unit, heat, g_heat,
                     g_temp_unit, g_cp);
  GV_HIP(hipGetLastError());
  return GITM_VERT_OK;
}

}  // extern "C"
