// vert_gpu.h -- pieces shared by the library's HIP translation units
// (gitm_vert.hip: single-grid step, prep and C ABI; gitm_vert_batch.hip:
// ensemble batches; gitm_vert_adjoint.hip: the adjoint): the kernel context
// of the plain build, the batch member table, the column prep, small
// kernels, the handle and argument helpers.  Internal; not installed.
#pragma once

#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

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
    const double lt = gv::ln(temp);
#pragma unroll
    for (int u = 0; u < 6; u++) out[u] = exp((sign * gv::DEXU[u]) * lt);
  }
};

// One column per wavefront of a larger workgroup (the short columns of a
// batch step): lanes are numbered from the wavefront's first thread, and a
// synchronisation is the wavefront's own.  A wavefront's LDS operations are
// carried out in the order it issues them, so the fences (which wait for
// them and keep the compiler from moving memory operations across) order
// its lanes' LDS writes before the reads that follow, and no s_barrier,
// which would stop the workgroup's other columns, is needed.
struct WaveCtx : GpuCtx {
  int base;  // first thread of the wavefront
  __device__ int lane() const { return (int)threadIdx.x - base; }
  __device__ void sync() const {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
  }
  __device__ bool sync_or(bool p) const {
    sync();
    return __builtin_amdgcn_ballot_w64(p) != 0;
  }
};

}  // namespace

// The batch's member table and step layout are shared by the batch's host
// code and the ensemble kernel's objects, so they have external linkage.
namespace gv {

// One member of a batch as the kernels see it.  The batch's columns are
// numbered per class: the columns of more than 64 levels (tall, one
// workgroup each) and those of at most 64 (short, one wavefront each); the
// member table holds the tall members, then the short ones, and column c of
// a class belongs to the member of that class with the largest
// first_block <= c.
struct BatchMember {
  gitm_vert_columns io;
  double *scratch;  // RK4 carries of the member's columns
  int first_block;  // the member's first column in its class
  int member;       // index in the caller's member list (dt_device)
  int col_offset;   // first column of the member in the batch-wide numbering
  int carry_lds;    // the RK4 carries of its columns in LDS (1) or in scratch (0)
};

// Work of one batch step: tasks 0..n_tall-1 are the tall columns, one per
// workgroup; the tasks after them hold per_wg short columns each, one per
// wavefront, in LDS slots of slot_doubles doubles.
struct EnsLayout {
  int n_tall_members, n_short_members;  // table entries of each class
  int n_tall, n_short;                  // columns of each class
  int per_wg;                           // short columns per workgroup
  int n_tasks;                          // n_tall + ceil(n_short / per_wg)
  int slot_doubles;
  int reserved;
};

}  // namespace gv

namespace {

using gv::BatchMember;
using gv::EnsLayout;

// Column of workgroup b in a launch of G workgroups, one column each.  The
// hardware hands consecutive workgroups to the eight XCDs in turn, and each
// XCD has an L2 of its own; giving each XCD a contiguous range of columns
// keeps neighbouring columns, which share cache lines in GITM's lon-fastest
// block arrays, on one L2 while they run.  A bijection of [0, G) for any G.
__device__ __forceinline__ int xcd_column(int b, int G) {
  constexpr int NXCD = 8;
  const int full = G / NXCD;
  if (b >= NXCD * full) return b;
  return (b % NXCD) * full + b / NXCD;
}

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
    // The word only decreases: a column whose step cannot lower it skips
    // the atomic, so the 10^4-10^5 columns do not queue on one address.
    const unsigned long long bits = (unsigned long long)__double_as_longlong(dt);
    if (bits < __hip_atomic_load(dt_bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      atomicMin(dt_bits, bits);
  }
}

__global__ void reset_error_word(unsigned long long *e) { *e = ~0ULL; }
__global__ void set_word(unsigned long long *w, unsigned long long v) { *w = v; }
__global__ void set_int(int *w, int v) { *w = v; }
__global__ void fill_words(unsigned long long *w, int n, unsigned long long v) {
  const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (i < n) w[i] = v;
}

}  // namespace

// Arguments of the column prep (gitm_vert_prepare and the batch form).
static inline int check_prep(const gitm_vert_prep_params *p, const gitm_vert_columns *c) {
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

static inline gv::PrepParams make_prep(const gitm_vert_prep_params *p, int n_alts) {
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
static inline unsigned long long prep_dt0_bits(const gitm_vert_prep_params *p) {
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
  // the hand-written adjoint's record (likewise)
  void *rec = nullptr;
  size_t rec_bytes = 0;
  size_t adj_budget = 0;  // 0: default_adjoint_budget()
  int *d_overflow = nullptr;
};

#define GV_HIP(call)                         \
  do {                                       \
    hipError_t e_ = (call);                  \
    if (e_ != hipSuccess) return GITM_VERT_EHIP; \
  } while (0)


// Host-array entry points: a field staged on the device over the index
// ranges a step touches (iLon, iLat in the descriptor's 1-based numbering,
// levels as first_level does), then redirected to the device copy.
struct HostSpan {
  gitm_vert_field *f;
  int lon_lo, lon_hi, lat_lo, lat_hi, lev_lo, lev_hi, ncomp;
  bool written;
};

struct HostStaging {
  std::vector<double *> host;
  std::vector<long long> lo, count;
};

// Copies every span in (into the handle's staging slots slot0..) and points
// its descriptor at the device copy; NULL fields are skipped.
static inline int stage_in(gitm_vert_handle *h, std::vector<HostSpan> &sp, int nblocks,
                           size_t slot0, HostStaging &st) {
  if (h->stage.size() < slot0 + sp.size()) {
    h->stage.resize(slot0 + sp.size(), nullptr);
    h->stage_bytes.resize(slot0 + sp.size(), 0);
  }
  st.host.assign(sp.size(), nullptr);
  st.lo.assign(sp.size(), 0);
  st.count.assign(sp.size(), 0);
  for (size_t i = 0; i < sp.size(); i++) {
    gitm_vert_field &f = *sp[i].f;
    st.host[i] = f.ptr;
    if (!f.ptr) continue;
    if (f.s_lon < 0 || f.s_lat < 0 || f.s_lev < 0 || f.s_comp < 0 || f.s_blk < 0)
      return GITM_VERT_EINVAL;
    const long long lo = (long long)(sp[i].lon_lo - 1) * f.s_lon +
                         (long long)(sp[i].lat_lo - 1) * f.s_lat +
                         (long long)(sp[i].lev_lo - f.first_level) * f.s_lev;
    const long long hi = (long long)(sp[i].lon_hi - 1) * f.s_lon +
                         (long long)(sp[i].lat_hi - 1) * f.s_lat +
                         (long long)(sp[i].lev_hi - f.first_level) * f.s_lev +
                         (long long)(sp[i].ncomp - 1) * f.s_comp +
                         (long long)(nblocks - 1) * f.s_blk;
    if (hi < lo) return GITM_VERT_EINVAL;
    const size_t bytes = (size_t)(hi - lo + 1) * sizeof(double), k = slot0 + i;
    if (h->stage_bytes[k] < bytes) {
      if (h->stage[k]) (void)hipFree(h->stage[k]);
      h->stage[k] = nullptr;
      h->stage_bytes[k] = 0;
      if (hipMalloc(&h->stage[k], bytes) != hipSuccess) return GITM_VERT_ENOMEM;
      h->stage_bytes[k] = bytes;
    }
    if (hipMemcpyAsync(h->stage[k], f.ptr + lo, bytes, hipMemcpyHostToDevice, h->stream) !=
        hipSuccess)
      return GITM_VERT_EHIP;
    f.ptr = (double *)h->stage[k] - lo;
    st.lo[i] = lo;
    st.count[i] = hi - lo + 1;
  }
  return GITM_VERT_OK;
}

// Copies the written spans back to the host arrays.
static inline int stage_out(gitm_vert_handle *h, const std::vector<HostSpan> &sp, size_t slot0,
                            const HostStaging &st) {
  for (size_t i = 0; i < sp.size(); i++) {
    if (!st.host[i] || !sp[i].written) continue;
    if (hipMemcpyAsync(st.host[i] + st.lo[i], h->stage[slot0 + i],
                       (size_t)st.count[i] * sizeof(double), hipMemcpyDeviceToHost,
                       h->stream) != hipSuccess)
      return GITM_VERT_EHIP;
  }
  return GITM_VERT_OK;
}

