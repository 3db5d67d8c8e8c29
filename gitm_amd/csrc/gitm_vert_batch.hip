// gitm_vert_batch.hip -- ensemble batches of the vertical step (the
// gitm_vert_batch_* entry points of include/gitm_vert.h).
#include "vert_gpu.h"

namespace {

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

}  // namespace

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

