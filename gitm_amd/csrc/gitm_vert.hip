// gitm_vert.hip -- HIP kernels and C ABI of the MI355X vertical step.
//
// Entry points are declared in include/gitm_vert.h.  One workgroup per
// column (gitm_vert_columns column index = blockIdx.x), NT = 64*W threads,
// one level per thread; see vert_column.h for the column algorithm.
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "vert_ad.h"
#include "vert_params.h"
#include "vert_prep.h"

namespace {

struct GpuCtx {
  using real = double;
  unsigned long long *err;
  __device__ int lane() const { return (int)threadIdx.x; }
  __device__ void sync() const { __syncthreads(); }
  __device__ bool sync_or(bool p) const { return __syncthreads_or(p ? 1 : 0) != 0; }
  __device__ void report(unsigned long long key) const { atomicMin(err, key); }
  // Temp**DiffExp as exp(DiffExp*ln Temp): one log and six exp in place of
  // fifteen pow (a pow costs five times an exp here); within a few ulp of pow.
  __device__ static constexpr bool device_math() { return true; }
  __device__ void diff_powers(double temp, double sign, double (&out)[6]) const {
    const double lt = log(temp);
#pragma unroll
    for (int u = 0; u < 6; u++) out[u] = exp((sign * gv::DEXU[u]) * lt);
  }
};

template <int NT, int SOLVER, bool IONS>
__global__ void __launch_bounds__(NT) vert_column_kernel(gv::KParams kp, gitm_vert_columns io) {
  extern __shared__ double lds[];
  GpuCtx cx{kp.err};
  gv::run_column<NT, SOLVER, IONS>(cx, kp, io, (int)blockIdx.x, lds);
}

// One member of a batch as the kernels see it.  A group of members with the
// same nAlts runs as one launch; workgroup b of the launch belongs to the
// member with the largest first_block <= b.
struct BatchMember {
  gitm_vert_columns io;
  double *scratch;  // RK4 carries of the member's columns
  int first_block;  // the member's first workgroup in its group's launch
  int member;       // index in the caller's member list (dt_device)
  int col_offset;   // first column of the member in the batch-wide numbering
  int reserved;
};

__device__ __forceinline__ int find_member(const BatchMember *__restrict__ m, int n,
                                           int block) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (m[mid].first_block <= block)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

template <int NT, int SOLVER, bool IONS>
__global__ void __launch_bounds__(NT)
    vert_batch_kernel(gv::KParams kp, const BatchMember *__restrict__ mem, int n_mem,
                      const double *__restrict__ dt_dev) {
  extern __shared__ double lds[];
  const int b = (int)blockIdx.x;
  const BatchMember &e = mem[find_member(mem, n_mem, b)];
  const double dt = dt_dev ? dt_dev[e.member] : kp.dt;
  GpuCtx cx{kp.err};
  const int col = b - e.first_block;
  gv::run_column<NT, SOLVER, IONS>(cx, kp, e.io, col, lds, e.scratch, dt, e.col_offset + col);
}

template <int NT, bool IONS>
void launch_batch_ions(int solver, dim3 grid, hipStream_t stream, const gv::KParams &kp,
                       const BatchMember *mem, int n_mem, const double *dt_dev) {
  const dim3 block(kp.n_alts + 4);
  const size_t lds = gv::lds_bytes(kp.n_alts, IONS, gv::carry_in_lds(NT / 64));
  if (solver == GITM_VERT_AUSM)
    hipLaunchKernelGGL((vert_batch_kernel<NT, GITM_VERT_AUSM, IONS>), grid, block, lds,
                       stream, kp, mem, n_mem, dt_dev);
  else
    hipLaunchKernelGGL((vert_batch_kernel<NT, GITM_VERT_RUSANOV, IONS>), grid, block, lds,
                       stream, kp, mem, n_mem, dt_dev);
}

template <int NT>
void launch_batch(int solver, dim3 grid, hipStream_t stream, const gv::KParams &kp,
                  const BatchMember *mem, int n_mem, const double *dt_dev) {
  if (kp.ions)
    launch_batch_ions<NT, true>(solver, grid, stream, kp, mem, n_mem, dt_dev);
  else
    launch_batch_ions<NT, false>(solver, grid, stream, kp, mem, n_mem, dt_dev);
}

// One thread per level: blockDim = nAlts+4 (<= NT, the compile-time bound).
template <int NT, bool IONS>
void launch_ions(int solver, dim3 grid, hipStream_t stream, const gv::KParams &kp,
                 const gitm_vert_columns &c) {
  const dim3 block(kp.n_alts + 4);
  const size_t lds = gv::lds_bytes(kp.n_alts, IONS, gv::carry_in_lds(NT / 64));
  if (solver == GITM_VERT_AUSM)
    hipLaunchKernelGGL((vert_column_kernel<NT, GITM_VERT_AUSM, IONS>), grid, block, lds,
                       stream, kp, c);
  else
    hipLaunchKernelGGL((vert_column_kernel<NT, GITM_VERT_RUSANOV, IONS>), grid, block, lds,
                       stream, kp, c);
}

template <int NT>
void launch(int solver, dim3 grid, hipStream_t stream, const gv::KParams &kp,
            const gitm_vert_columns &c) {
  if (kp.ions)
    launch_ions<NT, true>(solver, grid, stream, kp, c);
  else
    launch_ions<NT, false>(solver, grid, stream, kp, c);
}

__global__ void reset_error_word(unsigned long long *e) { *e = ~0ULL; }

// Column prep, one workgroup per column and one lane per level; lane 0 folds
// the column's CFL limit into *dt_bits (non-negative doubles order as their
// bit patterns, so an integer atomic min is a double min).
template <int NT>
__device__ __forceinline__ void prep_column(const gv::PrepParams &pp,
                                            const gitm_vert_columns &io,
                                            const gitm_vert_field &pf,
                                            const gitm_vert_field &cpf, int col,
                                            unsigned long long *dt_bits) {
  __shared__ double red[2][NT / 64];
  const int j = (int)threadIdx.x;
  const int nlb = io.n_lons * io.n_lats;
  const int ci = col % io.n_lons, cj = (col / io.n_lons) % io.n_lats, cb = col / nlb;
  double rate = 0.0, irate = 0.0;
  if (j < pp.n_alts + 4) gv::prep_level(pp, io, pf, cpf, ci, cj, cb, j - 1, rate, irate);
  // maxval over the column's levels 1..nAlts (calc_timestep.f90:122,132)
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    rate = fmax(rate, __shfl_xor(rate, o));
    irate = fmax(irate, __shfl_xor(irate, o));
  }
  if ((j & 63) == 0) {
    red[0][j >> 6] = rate;
    red[1][j >> 6] = irate;
  }
  __syncthreads();
  if (j == 0) {
    double r = red[0][0], ri = red[1][0];
    for (int w = 1; w < (int)((blockDim.x + 63) / 64); w++) {
      r = fmax(r, red[0][w]);
      ri = fmax(ri, red[1][w]);
    }
    double dt = pp.cfl / r;
    if (pp.ions) dt = fmin(dt, pp.cfl / ri);
    atomicMin(dt_bits, (unsigned long long)__double_as_longlong(dt));
  }
}

template <int NT>
__global__ void __launch_bounds__(NT) prep_kernel(gv::PrepParams pp, gitm_vert_columns io,
                                                  gitm_vert_field pf, gitm_vert_field cpf,
                                                  unsigned long long *dt_bits) {
  prep_column<NT>(pp, io, pf, cpf, (int)blockIdx.x, dt_bits);
}

// Batch prep: member e's step goes to dt_bits[e.member].
template <int NT>
__global__ void __launch_bounds__(NT) prep_batch_kernel(gv::PrepParams pp,
                                                        const BatchMember *__restrict__ mem,
                                                        int n_mem, unsigned long long *dt_bits) {
  const int b = (int)blockIdx.x;
  const BatchMember &e = mem[find_member(mem, n_mem, b)];
  const gitm_vert_field none = {};
  prep_column<NT>(pp, e.io, none, none, b - e.first_block, dt_bits + e.member);
}

__global__ void fill_words(unsigned long long *w, int n, unsigned long long v) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i < n) w[i] = v;
}

__global__ void set_word(unsigned long long *w, unsigned long long v) { *w = v; }
__global__ void set_int(int *w, int v) { *w = v; }

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
        const int p1 = t.p1[id], p2 = t.p2[id];
        if (p1 >= 0) add(p1, a * t.d1[id]);
        if (p2 >= 0) add(p2, a * t.d2[id]);
      }
      __threadfence();
      __syncthreads();
    }
  }
};

// Adjoint of one AUSM step of columns col0 .. col0+gridDim.x-1: one
// workgroup per column, one lane per level, as the forward kernel.
template <int NT>
__global__ void __launch_bounds__(NT)
    vert_adjoint_kernel(gv::KParams kp, gitm_vert_columns io, gv::AdjFields gout,
                        gv::AdjFields gin, gv::AdTape tape, int col0) {
  extern __shared__ double lds[];
  if (threadIdx.x == 0) gv::ad_tape_sh = tape;
  gv::ad_count_sh[threadIdx.x] = 0;
  __syncthreads();
  GpuAdCtx cx{kp.err};
  const int col = col0 + (int)blockIdx.x;
  gv::adjoint_column<NT>(cx, kp, io, gout, gin, col, reinterpret_cast<gv::adm::ad *>(lds),
                         kp.scratch, kp.dt, col);
}

}  // namespace

// Arguments of the column prep (gitm_vert_prepare and the batch form).
static int check_prep(const gitm_vert_prep_params *p, const gitm_vert_columns *c) {
  if (!p || !c) return GITM_VERT_EINVAL;
  if (c->n_alts < 10 || c->n_alts > GITM_VERT_MAX_NALTS) return GITM_VERT_EINVAL;
  if (c->n_lons < 1 || c->n_lats < 1 || c->n_blocks < 1) return GITM_VERT_EINVAL;
  if ((long long)c->n_lons * c->n_lats * c->n_blocks > 0x7fffffffLL) return GITM_VERT_EINVAL;
  const gitm_vert_field *req[] = {&c->ndensity_s, &c->temperature, &c->rho,
                                  &c->vertical_velocity, &c->dalt, &c->gamma,
                                  &c->mean_major_mass, &c->temp_unit, &c->eddy, &c->cmax};
  for (const gitm_vert_field *f : req)
    if (!f->ptr) return GITM_VERT_EINVAL;
  if (p->use_ion_advection && !c->ivelocity.ptr) return GITM_VERT_EINVAL;
  if (!(p->cfl > 0.0) || !(p->fixed_dt > 0.0)) return GITM_VERT_EINVAL;
  if (c->cmax.first_level < -1 || c->cmax.first_level > 0) return GITM_VERT_EINVAL;
  for (int s = 0; s < GITM_VERT_NSPECIES; s++)
    if (!(p->mass[s] > 0.0) || p->vibration[s] == 2.0) return GITM_VERT_EINVAL;
  return GITM_VERT_OK;
}

static gv::PrepParams make_prep(const gitm_vert_prep_params *p, int n_alts) {
  gv::PrepParams pp;
  for (int s = 0; s < gv::NSP; s++) {
    pp.mass[s] = p->mass[s];
    pp.vib2[s] = p->vibration[s] - 2;
    pp.kb_over_m[s] = gv::KB / p->mass[s];
  }
  pp.cfl = p->cfl;
  pp.eddy_coef = p->eddy_coef;
  pp.eddy_p0 = p->eddy_p0;
  pp.eddy_p1 = p->eddy_p1;
  pp.n_alts = n_alts;
  pp.ions = p->use_ion_advection != 0;
  return pp;
}

// min(FixedDt, DtEnd, 1e32) before the columns' limits (calc_timestep.f90
// :105,142,147), as the bits the device MIN starts from.
static unsigned long long prep_dt0_bits(const gitm_vert_prep_params *p) {
  double d0 = 1.0e32;
  if (p->dt_end < d0) d0 = p->dt_end;
  if (p->fixed_dt < d0) d0 = p->fixed_dt;
  unsigned long long bits;
  memcpy(&bits, &d0, sizeof(bits));
  return bits;
}

struct gitm_vert_handle {
  int device = 0;
  hipStream_t stream = nullptr;
  unsigned long long *d_err = nullptr;
  int waves = 0;
  // RK4 carries of the columns in flight (grown on demand, reused)
  double *scratch = nullptr;
  size_t scratch_bytes = 0;
  // device staging of the host-array entry point
  std::vector<void *> stage;
  std::vector<size_t> stage_bytes;
  // adjoint tape (grown on demand up to adj_budget bytes, reused)
  void *tape = nullptr;
  size_t tape_bytes = 0;
  size_t adj_budget = (size_t)16 << 30;
  int *d_overflow = nullptr;
};

#define GV_HIP(call)                         \
  do {                                       \
    hipError_t e_ = (call);                  \
    if (e_ != hipSuccess) return GITM_VERT_EHIP; \
  } while (0)

extern "C" {

void gitm_vert_default_params(gitm_vert_params *p) {
  if (p) gv::default_params(p);
}

const char *gitm_vert_strerror(int code) {
  switch (code) {
    case GITM_VERT_OK: return "ok";
    case GITM_VERT_ENONPOSITIVE: return "non-positive neutral density entering a column";
    case GITM_VERT_ESINGULAR: return "singular neutral-friction matrix";
    case GITM_VERT_ENEGION: return "negative ion density too close to the upper boundary";
    case GITM_VERT_EINVAL: return "invalid arguments";
    case GITM_VERT_EUNSUPPORTED:
      return "unsupported option (UseMSIS, tides, BC perturbation, UseDamping)";
    case GITM_VERT_EHIP: return "HIP runtime error";
    case GITM_VERT_ENOMEM: return "out of device memory";
    case GITM_VERT_ETAPE: return "adjoint tape too small for the column";
    default: return "unknown error";
  }
}

int gitm_vert_create(gitm_vert_handle **out, int device) {
  if (!out) return GITM_VERT_EINVAL;
  *out = nullptr;
  int dev = device;
  if (dev < 0) GV_HIP(hipGetDevice(&dev));
  GV_HIP(hipSetDevice(dev));
  gitm_vert_handle *h = new gitm_vert_handle();
  h->device = dev;
  if (hipMalloc(&h->d_err, sizeof(unsigned long long)) != hipSuccess) {
    delete h;
    return GITM_VERT_ENOMEM;
  }
  unsigned long long init = ~0ULL;
  if (hipMemcpy(h->d_err, &init, sizeof(init), hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(h->d_err);
    delete h;
    return GITM_VERT_EHIP;
  }
  *out = h;
  return GITM_VERT_OK;
}

int gitm_vert_destroy(gitm_vert_handle *h) {
  if (!h) return GITM_VERT_OK;
  (void)hipSetDevice(h->device);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  else (void)hipDeviceSynchronize();
  for (void *p : h->stage) (void)hipFree(p);
  if (h->scratch) (void)hipFree(h->scratch);
  if (h->tape) (void)hipFree(h->tape);
  if (h->d_overflow) (void)hipFree(h->d_overflow);
  (void)hipFree(h->d_err);
  delete h;
  return GITM_VERT_OK;
}

int gitm_vert_set_stream(gitm_vert_handle *h, void *hip_stream) {
  if (!h) return GITM_VERT_EINVAL;
  h->stream = (hipStream_t)hip_stream;
  return GITM_VERT_OK;
}

int gitm_vert_set_tile(gitm_vert_handle *h, int waves_per_column) {
  if (!h || (waves_per_column != 0 && !gv::valid_waves(waves_per_column)))
    return GITM_VERT_EINVAL;
  h->waves = waves_per_column;
  return GITM_VERT_OK;
}

int gitm_vert_advance(gitm_vert_handle *h, const gitm_vert_params *p,
                      const gitm_vert_columns *c, int solver, int n_steps) {
  if (!h) return GITM_VERT_EINVAL;
  int rc = gv::check_args(p, c, solver, n_steps);
  if (rc) return rc;
  const int ncols = c->n_lons * c->n_lats * c->n_blocks;
  if (n_steps == 0) return GITM_VERT_OK;
  int waves = h->waves ? h->waves : gv::auto_waves(c->n_alts);
  if (!gv::valid_waves(waves) || !gv::tile_fits(waves, c->n_alts)) return GITM_VERT_EINVAL;
  GV_HIP(hipSetDevice(h->device));
  gv::KParams kp = gv::make_kparams(p, c, solver, n_steps, h->d_err);
  kp.scr_fields = gv::scratch_fields(gv::carry_in_lds(waves), kp.ions != 0);
  const size_t need = (size_t)ncols * kp.scr_fields * (size_t)(c->n_alts + 4) * sizeof(double);
  if (h->scratch_bytes < need) {
    if (h->scratch) {
      // the previous launches on this handle's stream may still use it
      GV_HIP(hipStreamSynchronize(h->stream));
      (void)hipFree(h->scratch);
      h->scratch = nullptr;
      h->scratch_bytes = 0;
    }
    if (hipMalloc(&h->scratch, need) != hipSuccess) return GITM_VERT_ENOMEM;
    h->scratch_bytes = need;
  }
  kp.scratch = h->scratch;
  const dim3 grid(ncols);
  switch (waves) {
    case 1: launch<64>(solver, grid, h->stream, kp, *c); break;
    case 2: launch<128>(solver, grid, h->stream, kp, *c); break;
    case 3: launch<192>(solver, grid, h->stream, kp, *c); break;
    case 4: launch<256>(solver, grid, h->stream, kp, *c); break;
    default: return GITM_VERT_EINVAL;
  }
  GV_HIP(hipGetLastError());
  return GITM_VERT_OK;
}

void gitm_vert_default_prep_params(gitm_vert_prep_params *p) {
  if (!p) return;
  memset(p, 0, sizeof(*p));
  gitm_vert_params sp;
  gv::default_params(&sp);
  const double vib[GITM_VERT_NSPECIES] = {5.0, 7.0, 7.0, 5.0, 7.0, 5.0};  // ModEarth.f90:307-312
  for (int s = 0; s < GITM_VERT_NSPECIES; s++) {
    p->mass[s] = sp.mass[s];
    p->vibration[s] = vib[s];
  }
  p->cfl = 0.75;          // ModInputs.f90:123
  p->fixed_dt = 2.0;
  p->dt_end = 1.0e32;
  p->eddy_coef = 50.0;    // :302
  p->eddy_p0 = 0.01;      // :303
  p->eddy_p1 = 0.005;     // :304
}

int gitm_vert_prepare(gitm_vert_handle *h, const gitm_vert_prep_params *p,
                      const gitm_vert_columns *c, const gitm_vert_field *pressure,
                      const gitm_vert_field *cp, double *dt_device) {
  if (!h || !dt_device) return GITM_VERT_EINVAL;
  const int rc = check_prep(p, c);
  if (rc) return rc;
  GV_HIP(hipSetDevice(h->device));
  const gv::PrepParams pp = make_prep(p, c->n_alts);
  const gitm_vert_field none = {};
  const gitm_vert_field pf = pressure ? *pressure : none;
  const gitm_vert_field cpf = cp ? *cp : none;
  unsigned long long *bits = reinterpret_cast<unsigned long long *>(dt_device);
  hipLaunchKernelGGL(set_word, dim3(1), dim3(1), 0, h->stream, bits, prep_dt0_bits(p));
  const int ncols = c->n_lons * c->n_lats * c->n_blocks;
  const dim3 grid(ncols), block(c->n_alts + 4);
  switch (gv::auto_waves(c->n_alts)) {
    case 1: hipLaunchKernelGGL(prep_kernel<64>, grid, block, 0, h->stream, pp, *c, pf, cpf, bits); break;
    case 2: hipLaunchKernelGGL(prep_kernel<128>, grid, block, 0, h->stream, pp, *c, pf, cpf, bits); break;
    case 3: hipLaunchKernelGGL(prep_kernel<192>, grid, block, 0, h->stream, pp, *c, pf, cpf, bits); break;
    case 4: hipLaunchKernelGGL(prep_kernel<256>, grid, block, 0, h->stream, pp, *c, pf, cpf, bits); break;
    default: return GITM_VERT_EINVAL;
  }
  GV_HIP(hipGetLastError());
  return GITM_VERT_OK;
}

int gitm_vert_get_error(gitm_vert_handle *h, gitm_vert_error *err) {
  if (!h) return GITM_VERT_EINVAL;
  GV_HIP(hipSetDevice(h->device));
  unsigned long long key = ~0ULL;
  GV_HIP(hipMemcpyAsync(&key, h->d_err, sizeof(key), hipMemcpyDeviceToHost, h->stream));
  GV_HIP(hipStreamSynchronize(h->stream));
  if (key != ~0ULL) {
    hipLaunchKernelGGL(reset_error_word, dim3(1), dim3(1), 0, h->stream, h->d_err);
    GV_HIP(hipGetLastError());
    GV_HIP(hipStreamSynchronize(h->stream));
  }
  gitm_vert_error tmp;
  if (!err) err = &tmp;
  gv::decode_error(key, err);
  return err->code;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Host-array entry point (the Fortran drop-in): stage every field on the
// device, run, copy the written fields back.
// ---------------------------------------------------------------------------
namespace {

struct FieldSpec {
  gitm_vert_field gitm_vert_columns::*member;
  int ncomp;
  int last_level_offset;  // last level read = nAlts + this
  bool written;
  bool per_column;
  bool ions_only;  // staged only with UseIonAdvection
};

const FieldSpec kFields[] = {
    {&gitm_vert_columns::ndensity_s, 6, 2, true, false, false},
    {&gitm_vert_columns::vertical_velocity, 6, 2, true, false, false},
    {&gitm_vert_columns::velocity, 3, 2, true, false, false},
    {&gitm_vert_columns::temperature, 1, 2, true, false, false},
    {&gitm_vert_columns::rho, 1, 2, true, false, false},
    {&gitm_vert_columns::log_ns, 6, 2, true, false, false},
    {&gitm_vert_columns::ndensity, 1, 2, true, false, false},
    {&gitm_vert_columns::temp_unit, 1, 2, false, false, false},
    {&gitm_vert_columns::mean_major_mass, 1, 2, false, false, false},
    {&gitm_vert_columns::gamma, 1, 2, false, false, false},
    {&gitm_vert_columns::cmax, 1, 1, false, false, false},
    {&gitm_vert_columns::eddy, 1, 2, false, false, false},
    {&gitm_vert_columns::altitude, 1, 2, false, false, false},
    {&gitm_vert_columns::gravity, 1, 2, false, false, false},
    {&gitm_vert_columns::inv_radial_distance, 1, 2, false, false, false},
    {&gitm_vert_columns::dalt, 1, 2, false, false, false},
    {&gitm_vert_columns::cos_latitude, 1, 0, false, true, false},
    {&gitm_vert_columns::idensity_s, 10, 2, true, false, true},
    {&gitm_vert_columns::ivelocity, 3, 2, false, false, true},
    {&gitm_vert_columns::sza, 1, 0, false, true, true},
    {&gitm_vert_columns::mlat, 1, 0, false, true, true},
};

size_t field_span(const gitm_vert_columns &c, const gitm_vert_field &f, const FieldSpec &fs) {
  if (f.s_lon < 0 || f.s_lat < 0 || f.s_blk < 0 || f.s_lev < 0 || f.s_comp < 0) return 0;
  long long last = (long long)(c.n_lons - 1) * f.s_lon + (long long)(c.n_lats - 1) * f.s_lat +
                   (long long)(c.n_blocks - 1) * f.s_blk + (long long)(fs.ncomp - 1) * f.s_comp;
  if (!fs.per_column)
    last += (long long)(c.n_alts + fs.last_level_offset - f.first_level) * f.s_lev;
  return (size_t)(last + 1);
}

}  // namespace

extern "C" int gitm_vert_advance_host(gitm_vert_handle *h, const gitm_vert_params *p,
                                      const gitm_vert_columns *c, int solver, int n_steps) {
  if (!h) return GITM_VERT_EINVAL;
  int rc = gv::check_args(p, c, solver, n_steps);
  if (rc) return rc;
  GV_HIP(hipSetDevice(h->device));
  const size_t nf = sizeof(kFields) / sizeof(kFields[0]);
  if (h->stage.size() < nf) {
    h->stage.resize(nf, nullptr);
    h->stage_bytes.resize(nf, 0);
  }
  gitm_vert_columns d = *c;
  for (size_t i = 0; i < nf; i++) {
    const FieldSpec &fs = kFields[i];
    const gitm_vert_field &f = c->*fs.member;
    if (!f.ptr || (fs.ions_only && !p->use_ion_advection)) continue;
    size_t n = field_span(*c, f, fs);
    if (!n) return GITM_VERT_EINVAL;
    size_t bytes = n * sizeof(double);
    if (h->stage_bytes[i] < bytes) {
      if (h->stage[i]) (void)hipFree(h->stage[i]);
      h->stage[i] = nullptr;
      h->stage_bytes[i] = 0;
      if (hipMalloc(&h->stage[i], bytes) != hipSuccess) return GITM_VERT_ENOMEM;
      h->stage_bytes[i] = bytes;
    }
    GV_HIP(hipMemcpyAsync(h->stage[i], f.ptr, bytes, hipMemcpyHostToDevice, h->stream));
    (d.*fs.member).ptr = (double *)h->stage[i];
  }
  rc = gitm_vert_advance(h, p, &d, solver, n_steps);
  if (rc) return rc;
  for (size_t i = 0; i < nf; i++) {
    const FieldSpec &fs = kFields[i];
    const gitm_vert_field &f = c->*fs.member;
    if (!f.ptr || !fs.written || (fs.ions_only && !p->use_ion_advection)) continue;
    size_t bytes = field_span(*c, f, fs) * sizeof(double);
    GV_HIP(hipMemcpyAsync(f.ptr, h->stage[i], bytes, hipMemcpyDeviceToHost, h->stream));
  }
  return gitm_vert_get_error(h, nullptr);
}

// ---------------------------------------------------------------------------
// Batches of ensemble members with their own grids.  Members with the same
// nAlts form a group and run as one launch (one workgroup per column, one
// lane per level, so a launch needs a single nAlts); a step of the whole
// batch is one launch per distinct nAlts.  The member table and the scratch
// memory are made when the batch is created (or its tiles change), so the
// advance and prepare calls allocate and copy nothing and can be captured in
// a hipGraph.
// ---------------------------------------------------------------------------
struct gitm_vert_batch {
  gitm_vert_handle *h = nullptr;
  int ions = 0;
  std::vector<gitm_vert_columns> members;  // the caller's order
  std::vector<long long> col_offset;       // batch-wide column numbering, n+1 entries
  struct Group {
    int n_alts, waves, first, count, blocks;
  };
  std::vector<Group> groups;
  std::vector<BatchMember> table;  // group order
  BatchMember *d_table = nullptr;
  double *scratch = nullptr;
  size_t scratch_bytes = 0;
};

namespace {

long long member_columns(const gitm_vert_columns &c) {
  return (long long)c.n_lons * c.n_lats * c.n_blocks;
}

// Scratch offsets of the members for the current tiles; (re)allocates the
// scratch when it has to grow and uploads the member table.
int batch_layout(gitm_vert_batch *b) {
  size_t total = 0;  // doubles
  for (const gitm_vert_batch::Group &g : b->groups) {
    const int fields = gv::scratch_fields(gv::carry_in_lds(g.waves), b->ions != 0);
    for (int i = g.first; i < g.first + g.count; i++) {
      BatchMember &m = b->table[i];
      m.scratch = reinterpret_cast<double *>(total);  // offset for now
      total += (size_t)member_columns(m.io) * fields * (size_t)(g.n_alts + 4);
    }
  }
  gitm_vert_handle *h = b->h;
  GV_HIP(hipSetDevice(h->device));
  // the launches queued so far may still read the table and the scratch
  GV_HIP(hipStreamSynchronize(h->stream));
  const size_t need = total * sizeof(double);
  if (need > b->scratch_bytes) {
    if (b->scratch) (void)hipFree(b->scratch);
    b->scratch = nullptr;
    b->scratch_bytes = 0;
    if (hipMalloc(&b->scratch, need) != hipSuccess) return GITM_VERT_ENOMEM;
    b->scratch_bytes = need;
  }
  std::vector<BatchMember> dev = b->table;
  for (BatchMember &m : dev)
    m.scratch = b->scratch + reinterpret_cast<size_t>(m.scratch);
  GV_HIP(hipMemcpy(b->d_table, dev.data(), dev.size() * sizeof(BatchMember),
                   hipMemcpyHostToDevice));
  return GITM_VERT_OK;
}

}  // namespace

extern "C" {

int gitm_vert_batch_create(gitm_vert_handle *h, const gitm_vert_columns *members,
                           int n_members, int use_ion_advection, gitm_vert_batch **out) {
  if (!out) return GITM_VERT_EINVAL;
  *out = nullptr;
  if (!h || !members || n_members < 1) return GITM_VERT_EINVAL;
  long long total = 0;
  for (int m = 0; m < n_members; m++) {
    const gitm_vert_columns &c = members[m];
    if (c.n_alts < 10 || c.n_alts > GITM_VERT_MAX_NALTS) return GITM_VERT_EINVAL;
    if (c.n_lons < 1 || c.n_lats < 1 || c.n_blocks < 1) return GITM_VERT_EINVAL;
    total += member_columns(c);
  }
  if (total > 0x7fffffffLL) return GITM_VERT_EINVAL;
  gitm_vert_batch *b = new gitm_vert_batch();
  b->h = h;
  b->ions = use_ion_advection != 0;
  b->members.assign(members, members + n_members);
  b->col_offset.resize(n_members + 1, 0);
  for (int m = 0; m < n_members; m++)
    b->col_offset[m + 1] = b->col_offset[m] + member_columns(members[m]);
  // groups by nAlts, in increasing order; members keep the caller's order
  // within a group
  std::vector<int> order(n_members);
  for (int m = 0; m < n_members; m++) order[m] = m;
  std::stable_sort(order.begin(), order.end(), [&](int a, int c) {
    return members[a].n_alts < members[c].n_alts;
  });
  for (int i = 0; i < n_members; i++) {
    const int m = order[i];
    const gitm_vert_columns &c = members[m];
    if (b->groups.empty() || b->groups.back().n_alts != c.n_alts)
      b->groups.push_back({c.n_alts, gv::auto_waves(c.n_alts), i, 0, 0});
    gitm_vert_batch::Group &g = b->groups.back();
    BatchMember e;
    memset(&e, 0, sizeof(e));
    e.io = c;
    e.first_block = g.blocks;
    e.member = m;
    e.col_offset = (int)b->col_offset[m];
    b->table.push_back(e);
    g.count++;
    g.blocks += (int)member_columns(c);
  }
  int rc = GITM_VERT_OK;
  if (hipSetDevice(h->device) != hipSuccess) {
    rc = GITM_VERT_EHIP;
  } else if (hipMalloc(&b->d_table, b->table.size() * sizeof(BatchMember)) != hipSuccess) {
    rc = GITM_VERT_ENOMEM;
  } else {
    rc = batch_layout(b);
  }
  if (rc) {
    gitm_vert_batch_destroy(b);
    return rc;
  }
  *out = b;
  return GITM_VERT_OK;
}

int gitm_vert_batch_destroy(gitm_vert_batch *b) {
  if (!b) return GITM_VERT_OK;
  (void)hipSetDevice(b->h->device);
  (void)hipStreamSynchronize(b->h->stream);
  if (b->d_table) (void)hipFree(b->d_table);
  if (b->scratch) (void)hipFree(b->scratch);
  delete b;
  return GITM_VERT_OK;
}

int gitm_vert_batch_set_tile(gitm_vert_batch *b, int n_alts, int waves_per_column) {
  if (!b || (waves_per_column != 0 && !gv::valid_waves(waves_per_column)))
    return GITM_VERT_EINVAL;
  bool found = false;
  for (gitm_vert_batch::Group &g : b->groups) {
    if (n_alts != 0 && g.n_alts != n_alts) continue;
    const int w = waves_per_column ? waves_per_column : gv::auto_waves(g.n_alts);
    if (!gv::tile_fits(w, g.n_alts)) return GITM_VERT_EINVAL;
    found = true;
  }
  if (!found) return GITM_VERT_EINVAL;
  for (gitm_vert_batch::Group &g : b->groups)
    if (n_alts == 0 || g.n_alts == n_alts)
      g.waves = waves_per_column ? waves_per_column : gv::auto_waves(g.n_alts);
  return batch_layout(b);
}

int gitm_vert_batch_advance(gitm_vert_batch *b, const gitm_vert_params *p, int solver,
                            int n_steps, const double *dt_device) {
  if (!b || !p) return GITM_VERT_EINVAL;
  for (const gitm_vert_columns &c : b->members) {
    const int rc = gv::check_args(p, &c, solver, n_steps);
    if (rc) return rc;
  }
  if (p->use_ion_advection && !b->ions) return GITM_VERT_EINVAL;
  if (n_steps == 0) return GITM_VERT_OK;
  gitm_vert_handle *h = b->h;
  GV_HIP(hipSetDevice(h->device));
  for (const gitm_vert_batch::Group &g : b->groups) {
    gv::KParams kp = gv::make_kparams(p, &b->members[b->table[g.first].member], solver,
                                      n_steps, h->d_err);
    kp.scr_fields = gv::scratch_fields(gv::carry_in_lds(g.waves), kp.ions != 0);
    const BatchMember *mem = b->d_table + g.first;
    const dim3 grid(g.blocks);
    switch (g.waves) {
      case 1: launch_batch<64>(solver, grid, h->stream, kp, mem, g.count, dt_device); break;
      case 2: launch_batch<128>(solver, grid, h->stream, kp, mem, g.count, dt_device); break;
      case 3: launch_batch<192>(solver, grid, h->stream, kp, mem, g.count, dt_device); break;
      case 4: launch_batch<256>(solver, grid, h->stream, kp, mem, g.count, dt_device); break;
      default: return GITM_VERT_EINVAL;
    }
    GV_HIP(hipGetLastError());
  }
  return GITM_VERT_OK;
}

int gitm_vert_batch_prepare(gitm_vert_batch *b, const gitm_vert_prep_params *p,
                            double *dt_device) {
  if (!b || !p || !dt_device) return GITM_VERT_EINVAL;
  for (const gitm_vert_columns &c : b->members) {
    const int rc = check_prep(p, &c);
    if (rc) return rc;
  }
  gitm_vert_handle *h = b->h;
  GV_HIP(hipSetDevice(h->device));
  unsigned long long *bits = reinterpret_cast<unsigned long long *>(dt_device);
  const int n = (int)b->members.size();
  hipLaunchKernelGGL(fill_words, dim3((n + 255) / 256), dim3(256), 0, h->stream, bits, n,
                     prep_dt0_bits(p));
  for (const gitm_vert_batch::Group &g : b->groups) {
    const gv::PrepParams pp = make_prep(p, g.n_alts);
    const BatchMember *mem = b->d_table + g.first;
    const dim3 grid(g.blocks), block(g.n_alts + 4);
    switch (gv::auto_waves(g.n_alts)) {
      case 1: hipLaunchKernelGGL(prep_batch_kernel<64>, grid, block, 0, h->stream, pp, mem, g.count, bits); break;
      case 2: hipLaunchKernelGGL(prep_batch_kernel<128>, grid, block, 0, h->stream, pp, mem, g.count, bits); break;
      case 3: hipLaunchKernelGGL(prep_batch_kernel<192>, grid, block, 0, h->stream, pp, mem, g.count, bits); break;
      case 4: hipLaunchKernelGGL(prep_batch_kernel<256>, grid, block, 0, h->stream, pp, mem, g.count, bits); break;
      default: return GITM_VERT_EINVAL;
    }
    GV_HIP(hipGetLastError());
  }
  return GITM_VERT_OK;
}

int gitm_vert_batch_locate(const gitm_vert_batch *b, int column, int *member,
                           int *member_column) {
  if (!b || column < 0 || column >= b->col_offset.back()) return GITM_VERT_EINVAL;
  const int m = (int)(std::upper_bound(b->col_offset.begin(), b->col_offset.end(),
                                       (long long)column) -
                      b->col_offset.begin()) - 1;
  if (member) *member = m;
  if (member_column) *member_column = (int)(column - b->col_offset[m]);
  return GITM_VERT_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Adjoint of one AUSM step (vert_ad.h).  The columns run in chunks whose
// tapes fit the handle's adjoint memory budget.
// ---------------------------------------------------------------------------
namespace {

// Tape entries per lane: about 10k nodes per level-lane and step, plus the
// serial loops of the boundary lanes, which grow with the column.
int adjoint_nmax(int n_alts) { return 12288 + 8 * n_alts; }
constexpr int kAdjMaxPhases = 128;
constexpr size_t kNodeBytes = 2 * sizeof(double) + sizeof(double) + 2 * sizeof(int);

template <int NT>
void launch_adjoint(int ncols, int col0, hipStream_t stream, const gv::KParams &kp,
                    const gitm_vert_columns &c, const gv::AdjFields &go,
                    const gv::AdjFields &gi, const gv::AdTape &tape) {
  const size_t lds =
      gv::lds_doubles(kp.n_alts, false, gv::carry_in_lds(NT / 64)) * sizeof(gv::adm::ad);
  hipLaunchKernelGGL(vert_adjoint_kernel<NT>, dim3(ncols), dim3(kp.n_alts + 4), lds, stream,
                     kp, c, go, gi, tape, col0);
}

}  // namespace

extern "C" {

int gitm_vert_set_adjoint_memory(gitm_vert_handle *h, size_t bytes) {
  if (!h || bytes < ((size_t)64 << 20)) return GITM_VERT_EINVAL;
  h->adj_budget = bytes;
  return GITM_VERT_OK;
}

int gitm_vert_adjoint(gitm_vert_handle *h, const gitm_vert_params *p,
                      const gitm_vert_columns *c, const gitm_vert_adjoint_fields *gout,
                      const gitm_vert_adjoint_fields *gin) {
  if (!h || !gout || !gin) return GITM_VERT_EINVAL;
  int rc = gv::check_args(p, c, GITM_VERT_AUSM, 1);
  if (rc) return rc;
  if (p->use_ion_advection) return GITM_VERT_EUNSUPPORTED;
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
  const size_t lane_bytes = (size_t)nmax * kNodeBytes + (size_t)kAdjMaxPhases * sizeof(int);
  long long chunk = (long long)(h->adj_budget / (lane_bytes * (size_t)L));
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
  tape.d1 = reinterpret_cast<double *>(base);
  tape.d2 = tape.d1 + nn;
  tape.adj = tape.d2 + nn;
  tape.p1 = reinterpret_cast<int *>(tape.adj + nn);
  tape.p2 = tape.p1 + nn;
  tape.phase = tape.p2 + nn;
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
