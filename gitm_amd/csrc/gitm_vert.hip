// gitm_vert.hip -- HIP kernels and C ABI of the MI355X vertical step:
// the single-grid step, the column prep, errors and the host-array entry.
//
// Entry points are declared in include/gitm_vert.h.  One workgroup per
// column (the column of workgroup blockIdx.x is xcd_column's), NT = 64*W threads,
// one level per thread; see vert_column.h for the column algorithm.
#include "vert_gpu.h"

namespace {

template <int NT, int SOLVER, bool IONS>
__global__ void __launch_bounds__(NT) vert_column_kernel(gv::KParams kp, gitm_vert_columns io) {
  extern __shared__ double lds[];
  GpuCtx cx{kp.err};
  gv::run_column<NT, SOLVER, IONS>(cx, kp, io, xcd_column((int)blockIdx.x, (int)gridDim.x),
                                   lds);
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

template <int NT>
__global__ void __launch_bounds__(NT) prep_kernel(gv::PrepParams pp, gitm_vert_columns io,
                                                  gitm_vert_field pf, gitm_vert_field cpf,
                                                  unsigned long long *dt_bits) {
  prep_column<NT>(pp, io, pf, cpf, xcd_column((int)blockIdx.x, (int)gridDim.x), dt_bits);
}

}  // namespace

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
      return "unsupported option (BC perturbation, UseDamping with the AUSM solver)";
    case GITM_VERT_EHIP: return "HIP runtime error";
    case GITM_VERT_ENOMEM: return "out of device memory";
    case GITM_VERT_ETAPE: return "adjoint tape too small for the column";
    case GITM_VERT_ELOWTEMP: return "temperature below 100 after the message exchange";
    case GITM_VERT_ESCRATCH:
      return "columns too large for the fixed scratch of gitm_vert_advance_noalloc";
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
  if (h->rec) (void)hipFree(h->rec);
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
  bool bottom = false;  // levels first_level..0 only (the tides)
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
    {&gitm_vert_columns::msis_bc, 9, 0, false, true, false},
    {&gitm_vert_columns::tides_east, 1, 0, false, false, false, true},
    {&gitm_vert_columns::tides_north, 1, 0, false, false, false, true},
    {&gitm_vert_columns::tides_temp, 1, 0, false, false, false, true},
};

size_t field_span(const gitm_vert_columns &c, const gitm_vert_field &f, const FieldSpec &fs) {
  if (f.s_lon < 0 || f.s_lat < 0 || f.s_blk < 0 || f.s_lev < 0 || f.s_comp < 0) return 0;
  long long last = (long long)(c.n_lons - 1) * f.s_lon + (long long)(c.n_lats - 1) * f.s_lat +
                   (long long)(c.n_blocks - 1) * f.s_blk + (long long)(fs.ncomp - 1) * f.s_comp;
  if (fs.bottom)
    last += (long long)(0 - f.first_level) * f.s_lev;
  else if (!fs.per_column)
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
// gitm_vert_advance_noalloc: gitm_vert_advance without a handle and without
// any allocation, for callers that cannot allocate device memory (code under
// hipGraph capture, serving loops).  Its scratch memory is the file-scope
// array below, which every device the library runs on holds until the
// process ends: the first 256 bytes hold the error word, the rest the RK4
// carries that gitm_vert_advance keeps in its handle's scratch (ion carries
// and the carries of columns taller than 188 levels).
// ---------------------------------------------------------------------------
#include <mutex>

namespace {

constexpr size_t kNoallocScratchBytes = (size_t)64 << 20;
constexpr size_t kNoallocCarryOffset = 256 / sizeof(double);  // doubles before the carries
static __device__ __align__(256) double scratch[kNoallocScratchBytes / sizeof(double)];

// The array's address on each device, from hipGetSymbolAddress with that
// device current; the name `scratch` is a host-side stand-in in host code and
// is never passed to a kernel.
constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
double *g_noalloc_ptr[kMaxDevices] = {};

int noalloc_scratch(int dev, double **out) {
  if (dev < 0 || dev >= kMaxDevices) return GITM_VERT_EINVAL;
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  if (!g_noalloc_ptr[dev]) {
    void *ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr)
      return GITM_VERT_EHIP;
    g_noalloc_ptr[dev] = static_cast<double *>(ptr);
  }
  *out = g_noalloc_ptr[dev];
  return GITM_VERT_OK;
}

__global__ void copy_word(unsigned long long *dst, const unsigned long long *src) { *dst = *src; }

}  // namespace

extern "C" int gitm_vert_advance_noalloc(const gitm_vert_params *p, const gitm_vert_columns *c,
                                         int solver, int n_steps, void *hip_stream,
                                         unsigned long long *err_word) {
  int rc = gv::check_args(p, c, solver, n_steps);
  if (rc) return rc;
  const int waves = gv::auto_waves(c->n_alts);
  if (!gv::valid_waves(waves)) return GITM_VERT_EINVAL;
  const size_t ncols = (size_t)c->n_lons * c->n_lats * c->n_blocks;
  const int fields = gv::scratch_fields(gv::carry_in_lds(waves), p->use_ion_advection != 0);
  const size_t need = ncols * (size_t)fields * (size_t)(c->n_alts + 4);
  if (need > kNoallocScratchBytes / sizeof(double) - kNoallocCarryOffset)
    return GITM_VERT_ESCRATCH;
  if (n_steps == 0) return GITM_VERT_OK;
  int dev = 0;
  GV_HIP(hipGetDevice(&dev));
  double *base = nullptr;
  rc = noalloc_scratch(dev, &base);
  if (rc) return rc;
  hipStream_t stream = (hipStream_t)hip_stream;
  unsigned long long *err = reinterpret_cast<unsigned long long *>(base);
  // every call starts with a clear error word (a fill kernel, so that the
  // call can be captured in a hipGraph)
  hipLaunchKernelGGL(set_word, dim3(1), dim3(1), 0, stream, err, ~0ULL);
  gv::KParams kp = gv::make_kparams(p, c, solver, n_steps, err);
  kp.scr_fields = fields;
  kp.scratch = base + kNoallocCarryOffset;
  const dim3 grid((unsigned)ncols);
  switch (waves) {
    case 1: launch<64>(solver, grid, stream, kp, *c); break;
    case 2: launch<128>(solver, grid, stream, kp, *c); break;
    case 3: launch<192>(solver, grid, stream, kp, *c); break;
    case 4: launch<256>(solver, grid, stream, kp, *c); break;
    default: return GITM_VERT_EINVAL;
  }
  if (err_word) hipLaunchKernelGGL(copy_word, dim3(1), dim3(1), 0, stream, err_word, err);
  GV_HIP(hipGetLastError());
  return GITM_VERT_OK;
}
