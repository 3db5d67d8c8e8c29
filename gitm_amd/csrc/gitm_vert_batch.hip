// gitm_vert_batch.hip -- ensemble batches of the vertical step (the
// gitm_vert_batch_* entry points of include/gitm_vert.h).
//
// A batch step is one launch of vert_ensemble_kernel (gitm_vert_ensemble.hip,
// one object per workgroup size) over a task list: the columns of more than
// 64 levels one workgroup each, then the shorter columns one wavefront each,
// several per workgroup.  The batch prep is one launch as well.
#include "vert_gpu.h"

#define GV_ENS_DECL(W)                                                                     \
  int ensemble_occupancy_w##W(int solver, bool ions, size_t lds_bytes);                    \
  hipError_t launch_ensemble_w##W(int solver, bool ions, int grid, size_t lds_bytes,       \
                                  hipStream_t stream, const gv::KParams &kp,               \
                                  const BatchMember *mem, const EnsLayout &lay,            \
                                  const double *dt_dev);
GV_ENS_DECL(1)
GV_ENS_DECL(2)
GV_ENS_DECL(3)
GV_ENS_DECL(4)
#undef GV_ENS_DECL

namespace {

int ensemble_occupancy(int waves, int solver, bool ions, size_t lds) {
  switch (waves) {
    case 1: return ensemble_occupancy_w1(solver, ions, lds);
    case 2: return ensemble_occupancy_w2(solver, ions, lds);
    case 3: return ensemble_occupancy_w3(solver, ions, lds);
    case 4: return ensemble_occupancy_w4(solver, ions, lds);
    default: return 0;
  }
}

hipError_t launch_ensemble(int waves, int solver, bool ions, int grid, size_t lds,
                           hipStream_t stream, const gv::KParams &kp, const BatchMember *mem,
                           const EnsLayout &lay, const double *dt_dev) {
  switch (waves) {
    case 1: return launch_ensemble_w1(solver, ions, grid, lds, stream, kp, mem, lay, dt_dev);
    case 2: return launch_ensemble_w2(solver, ions, grid, lds, stream, kp, mem, lay, dt_dev);
    case 3: return launch_ensemble_w3(solver, ions, grid, lds, stream, kp, mem, lay, dt_dev);
    case 4: return launch_ensemble_w4(solver, ions, grid, lds, stream, kp, mem, lay, dt_dev);
    default: return hipErrorInvalidValue;
  }
}

// Batch prep: one workgroup per column of any member (tall columns first,
// as the step numbers them), the lanes above the column's top idle; member
// e's step goes to dt_bits[e.member].
template <int NT>
__global__ void __launch_bounds__(NT) prep_batch_kernel(gv::PrepParams pp,
                                                        const BatchMember *__restrict__ mem,
                                                        EnsLayout lay,
                                                        unsigned long long *dt_bits) {
  const int b = (int)blockIdx.x;
  const bool tall = b < lay.n_tall;
  const int c = tall ? b : b - lay.n_tall;
  const int m = tall ? find_member(mem, lay.n_tall_members, c)
                     : lay.n_tall_members +
                           find_member(mem + lay.n_tall_members, lay.n_short_members, c);
  const BatchMember &e = mem[m];
  gv::PrepParams q = pp;
  q.n_alts = e.io.n_alts;
  const gitm_vert_field none = {};
  prep_column<NT>(q, e.io, none, none, c - e.first_block, dt_bits + e.member);
}

}  // namespace

struct gitm_vert_batch {
  gitm_vert_handle *h = nullptr;
  int ions = 0;
  std::vector<gitm_vert_columns> members;  // the caller's order
  std::vector<long long> col_offset;       // batch-wide column numbering, n+1 entries
  std::vector<BatchMember> table;          // tall members, then short ones
  EnsLayout lay{};                         // counts; per_wg, n_tasks with the tile
  int max_alts = 0, max_alts_tall = 0, max_alts_short = 0;
  bool uniform_tall = true;  // every tall column needs the same wavefront count
  int tall_waves = 0;        // ... which is this one
  gitm_vert_batch_tile tile{};
  int occupancy[2][2] = {};  // workgroups per CU of the tile's kernel [ausm][ions]
  int slot[2] = {};          // short columns' LDS slot in doubles [ions]
  size_t lds[2] = {};        // LDS of a workgroup [ions]
  int n_cu = 0;
  size_t lds_max = 0;
  BatchMember *d_table = nullptr;
  double *scratch = nullptr;
  size_t scratch_bytes = 0;
};

namespace {

long long member_columns(const gitm_vert_columns &c) {
  return (long long)c.n_lons * c.n_lats * c.n_blocks;
}

int waves_of(int n_alts) { return (n_alts + 4 + 63) / 64; }

// LDS doubles of a short column of n_alts levels, rounded up to even so
// that every slot starts 16-byte aligned
int short_doubles(int n_alts, bool ions, bool carry) {
  const int d = (int)gv::lds_doubles(n_alts, ions, carry);
  return d + (d & 1);
}

// Where the tile puts each member's RK4 carries (1: LDS, 0: scratch), in
// table order: carries_in_lds 0 and 1 put all of them in scratch or in LDS;
// 2 puts a member's in LDS when its column's LDS (a tall column's, or a
// short column's times the short columns per workgroup) then stays within
// the share of a CU's LDS that each of the workgroups its registers allow
// gets.
std::vector<int> place_carries(const gitm_vert_batch *b, const gitm_vert_batch_tile &t) {
  std::vector<int> flag(b->table.size(), t.carries_in_lds ? 1 : 0);
  if (t.carries_in_lds != 2) return flag;
  const bool ions = b->ions != 0;
  const int occ = std::max(1, ensemble_occupancy(t.waves_per_workgroup, GITM_VERT_AUSM, ions, 0));
  const size_t budget = b->lds_max / (size_t)occ;
  for (size_t i = 0; i < b->table.size(); i++) {
    const int na = b->table[i].io.n_alts;
    const size_t need = (int)i < b->lay.n_tall_members
                            ? (size_t)gv::lds_bytes(na, ions, true)
                            : (size_t)t.columns_per_workgroup * short_doubles(na, ions, true) *
                                  sizeof(double);
    flag[i] = need <= budget;
  }
  return flag;
}

// LDS doubles of one short column's slot: the largest short column with its
// carries where `flag` puts them
int slot_doubles(const gitm_vert_batch *b, bool ions, const std::vector<int> &flag) {
  int d = 0;
  for (int i = b->lay.n_tall_members; i < (int)b->table.size(); i++)
    d = std::max(d, short_doubles(b->table[i].io.n_alts, ions, flag[i] != 0));
  return d;
}

// LDS of one workgroup of the step under tile t
size_t tile_lds(const gitm_vert_batch *b, const gitm_vert_batch_tile &t, bool ions) {
  const std::vector<int> flag = place_carries(b, t);
  size_t need = 0;
  for (int i = 0; i < b->lay.n_tall_members; i++)
    need = std::max(need, (size_t)gv::lds_bytes(b->table[i].io.n_alts, ions, flag[i] != 0));
  return std::max(need, (size_t)t.columns_per_workgroup * slot_doubles(b, ions, flag) *
                            sizeof(double));
}

int check_tile(const gitm_vert_batch *b, const gitm_vert_batch_tile &t) {
  const int w = t.waves_per_workgroup;
  if (w < 1 || w > 4 || 64 * w < b->max_alts + 4) return GITM_VERT_EINVAL;
  if (t.columns_per_workgroup < 1 || t.columns_per_workgroup > w) return GITM_VERT_EINVAL;
  if (t.carries_in_lds < 0 || t.carries_in_lds > 2 || (t.persistent != 0 && t.persistent != 1))
    return GITM_VERT_EINVAL;
  if (tile_lds(b, t, b->ions != 0) > b->lds_max) return GITM_VERT_EINVAL;
  // a looping grid needs every wavefront of a tall task to hold levels
  if (t.persistent && b->lay.n_tall && !(b->uniform_tall && b->tall_waves == w))
    return GITM_VERT_EINVAL;
  return GITM_VERT_OK;
}

// The smallest workgroup that holds the tallest column, one wavefront per
// short column, the carries in LDS where they fit.
gitm_vert_batch_tile auto_tile(const gitm_vert_batch *b) {
  gitm_vert_batch_tile t;
  t.waves_per_workgroup = waves_of(b->max_alts);
  t.columns_per_workgroup = t.waves_per_workgroup;
  t.carries_in_lds = 2;
  t.persistent = 0;
  return t;
}

// Scratch offsets of the members for the tile; (re)allocates the scratch
// when it has to grow, uploads the member table and records the persistent
// grid's occupancy.
int batch_layout(gitm_vert_batch *b) {
  const std::vector<int> flag = place_carries(b, b->tile);
  size_t total = 0;  // doubles
  for (size_t i = 0; i < b->table.size(); i++) {
    BatchMember &m = b->table[i];
    m.carry_lds = flag[i];
    const int fields = gv::scratch_fields(flag[i] != 0, b->ions != 0);
    m.scratch = reinterpret_cast<double *>(total);  // offset for now
    total += (size_t)member_columns(m.io) * fields * (size_t)(m.io.n_alts + 4);
  }
  b->lay.per_wg = b->tile.columns_per_workgroup;
  b->lay.n_tasks = b->lay.n_tall + (b->lay.n_short + b->lay.per_wg - 1) / b->lay.per_wg;
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
  for (int i = 0; i < 2; i++) {
    b->slot[i] = slot_doubles(b, i != 0, flag);
    b->lds[i] = tile_lds(b, b->tile, i != 0);
    for (int s = 0; s < 2; s++)
      b->occupancy[s][i] =
          (i && !b->ions) ? 0
                          : ensemble_occupancy(b->tile.waves_per_workgroup,
                                               s ? GITM_VERT_AUSM : GITM_VERT_RUSANOV, i != 0,
                                               b->lds[i]);
  }
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
  // the table: tall members in the caller's order, then the short ones
  for (int pass = 0; pass < 2; pass++) {
    long long first = 0;
    for (int m = 0; m < n_members; m++) {
      const gitm_vert_columns &c = members[m];
      const bool tall = waves_of(c.n_alts) > 1;
      if (tall != (pass == 0)) continue;
      BatchMember e;
      memset(&e, 0, sizeof(e));
      e.io = c;
      e.first_block = (int)first;
      e.member = m;
      e.col_offset = (int)b->col_offset[m];
      b->table.push_back(e);
      first += member_columns(c);
      b->max_alts = std::max(b->max_alts, c.n_alts);
      if (tall) {
        b->max_alts_tall = std::max(b->max_alts_tall, c.n_alts);
        if (b->tall_waves && b->tall_waves != waves_of(c.n_alts)) b->uniform_tall = false;
        b->tall_waves = waves_of(c.n_alts);
      } else {
        b->max_alts_short = std::max(b->max_alts_short, c.n_alts);
      }
    }
    if (pass == 0) {
      b->lay.n_tall_members = (int)b->table.size();
      b->lay.n_tall = (int)first;
    } else {
      b->lay.n_short_members = (int)b->table.size() - b->lay.n_tall_members;
      b->lay.n_short = (int)first;
    }
  }
  int rc = GITM_VERT_OK;
  int lds_max = 0;
  if (hipSetDevice(h->device) != hipSuccess ||
      hipDeviceGetAttribute(&b->n_cu, hipDeviceAttributeMultiprocessorCount, h->device) !=
          hipSuccess ||
      hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, h->device) !=
          hipSuccess) {
    rc = GITM_VERT_EHIP;
  } else if (hipMalloc(&b->d_table, b->table.size() * sizeof(BatchMember)) != hipSuccess) {
    rc = GITM_VERT_ENOMEM;
  } else {
    b->lds_max = (size_t)lds_max;
    b->tile = auto_tile(b);
    rc = check_tile(b, b->tile);
    if (!rc) rc = batch_layout(b);
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

int gitm_vert_batch_set_tile(gitm_vert_batch *b, const gitm_vert_batch_tile *t) {
  if (!b) return GITM_VERT_EINVAL;
  GV_HIP(hipSetDevice(b->h->device));
  const gitm_vert_batch_tile nt = t ? *t : auto_tile(b);
  const int rc = check_tile(b, nt);
  if (rc) return rc;
  b->tile = nt;
  return batch_layout(b);
}

int gitm_vert_batch_get_tile(const gitm_vert_batch *b, gitm_vert_batch_tile *t) {
  if (!b || !t) return GITM_VERT_EINVAL;
  *t = b->tile;
  return GITM_VERT_OK;
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
  // nAlts and the carries' stride come from each member in the kernel
  gv::KParams kp = gv::make_kparams(p, &b->members[0], solver, n_steps, h->d_err);
  const bool ions = kp.ions != 0;
  EnsLayout lay = b->lay;
  lay.slot_doubles = b->slot[ions];
  const size_t lds = b->lds[ions];
  int grid = lay.n_tasks;
  if (b->tile.persistent) {
    const int occ = b->occupancy[solver == GITM_VERT_AUSM][ions];
    if (occ < 1) return GITM_VERT_EHIP;
    grid = std::min(grid, occ * b->n_cu);
  }
  if (grid < 1) return GITM_VERT_OK;
  GV_HIP(launch_ensemble(b->tile.waves_per_workgroup, solver, ions, grid, lds, h->stream, kp,
                         b->d_table, lay, dt_device));
  GV_HIP(hipGetLastError());
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
  const gv::PrepParams pp = make_prep(p, b->max_alts);
  const dim3 grid(b->lay.n_tall + b->lay.n_short);
  switch (waves_of(b->max_alts)) {
    case 1: hipLaunchKernelGGL(prep_batch_kernel<64>, grid, dim3(64), 0, h->stream, pp, b->d_table, b->lay, bits); break;
    case 2: hipLaunchKernelGGL(prep_batch_kernel<128>, grid, dim3(128), 0, h->stream, pp, b->d_table, b->lay, bits); break;
    case 3: hipLaunchKernelGGL(prep_batch_kernel<192>, grid, dim3(192), 0, h->stream, pp, b->d_table, b->lay, bits); break;
    case 4: hipLaunchKernelGGL(prep_batch_kernel<256>, grid, dim3(256), 0, h->stream, pp, b->d_table, b->lay, bits); break;
    default: return GITM_VERT_EINVAL;
  }
  GV_HIP(hipGetLastError());
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
