// gitm_vert_ensemble.hip -- the batch step kernel for one workgroup size.
//
// Compiled once per size: GV_ENS_WAVES = 1..4 wavefronts per workgroup
// (gitm_amd/build.py), each object defining launch_ensemble_w<GV_ENS_WAVES>
// for gitm_vert_batch.hip.  One launch steps every member of a batch.
#include "vert_gpu.h"

#ifndef GV_ENS_WAVES
#error "GV_ENS_WAVES (wavefronts per workgroup, 1..4) must be defined"
#endif

namespace {

constexpr int NT = 64 * GV_ENS_WAVES;

// Tasks t = blockIdx.x, blockIdx.x + gridDim.x, ... of the step (EnsLayout):
//  * a tall column takes the whole workgroup, one lane per level, the lanes
//    above its top masked (as in a single-grid launch of blockDim nAlts+4);
//  * a task of short columns gives each of its first per_wg wavefronts one
//    column and an LDS slot of its own; those columns synchronise within
//    their wavefront only (WaveCtx).
// The barrier between two tasks of a workgroup lets the next task reuse the
// LDS; with one task per workgroup (gridDim.x = n_tasks) there is none, and
// a wavefront whose lanes all lie above a tall column's top ends at once,
// which removes it from the column's barriers.  A looping (persistent) grid
// therefore needs every tall column to fill all NT/64 wavefronts; the host
// checks that before it launches one.  Each member's RK4 carries live where
// its table entry says (carry_lds), so that the members whose columns fit
// the LDS budget of the tile keep them there and the others in scratch.
template <int W, int SOLVER, bool IONS, class Ctx>
__device__ __forceinline__ void step_column(Ctx &cx, const gv::KParams &kp,
                                            const BatchMember &e, int col, double *lds,
                                            double dt) {
  if (e.carry_lds)
    gv::run_column<64 * W, SOLVER, IONS, true>(cx, kp, e.io, col, lds, e.scratch, dt,
                                               e.col_offset + col, e.io.n_alts,
                                               gv::scratch_fields(true, IONS));
  else
    gv::run_column<64 * W, SOLVER, IONS, false>(cx, kp, e.io, col, lds, e.scratch, dt,
                                                e.col_offset + col, e.io.n_alts,
                                                gv::scratch_fields(false, IONS));
}

template <int SOLVER, bool IONS>
__global__ void __launch_bounds__(NT)
    vert_ensemble_kernel(gv::KParams kp, const BatchMember *__restrict__ mem, EnsLayout lay,
                         const double *__restrict__ dt_dev) {
  extern __shared__ double lds[];
  // The member of the current task: found by bisection for the first task
  // of each class, then followed forward (tasks only grow) past the first
  // column of the next member, which is kept, so that a task normally costs
  // no load of the table.
  int mt = -1, ms = -1, mt_end = 0, ms_end = 0;
  const BatchMember *shrt = mem + lay.n_tall_members;
  const auto next_first = [](const BatchMember *m, int n, int i) {
    return i + 1 < n ? m[i + 1].first_block : 0x7fffffff;
  };
  for (int t = (int)blockIdx.x; t < lay.n_tasks; t += (int)gridDim.x) {
    bool tall = false;
    if constexpr (NT > 64) {
      if (t < lay.n_tall) {
        tall = true;
        if (mt < 0) {
          mt = find_member(mem, lay.n_tall_members, t);
          mt_end = next_first(mem, lay.n_tall_members, mt);
        }
        while (t >= mt_end) mt_end = next_first(mem, lay.n_tall_members, ++mt);
        const BatchMember &e = mem[mt];
        const int col = t - e.first_block;
        const double dt = dt_dev ? dt_dev[e.member] : kp.dt;
        GpuCtx cx{kp.err};
        if ((int)threadIdx.x < e.io.n_alts + 4)
          step_column<NT / 64, SOLVER, IONS>(cx, kp, e, col, lds, dt);
      }
    }
    if (!tall) {
      const int w = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
      const long long s = (long long)(t - lay.n_tall) * lay.per_wg + w;
      if (w < lay.per_wg && s < lay.n_short) {
        if (ms < 0) {
          ms = find_member(shrt, lay.n_short_members, (int)s);
          ms_end = next_first(shrt, lay.n_short_members, ms);
        }
        while ((int)s >= ms_end) ms_end = next_first(shrt, lay.n_short_members, ++ms);
        const BatchMember &e = shrt[ms];
        const int col = (int)s - e.first_block;
        const double dt = dt_dev ? dt_dev[e.member] : kp.dt;
        WaveCtx cx{{kp.err}, w << 6};
        if (cx.lane() < e.io.n_alts + 4)
          step_column<1, SOLVER, IONS>(cx, kp, e, col, lds + (long long)w * lay.slot_doubles,
                                       dt);
      }
    }
    if (t + (int)gridDim.x < lay.n_tasks) __syncthreads();
  }
}

template <int SOLVER, bool IONS>
const void *kernel_of() {
  return reinterpret_cast<const void *>(&vert_ensemble_kernel<SOLVER, IONS>);
}

const void *select_kernel(int solver, bool ions) {
  if (solver == GITM_VERT_AUSM)
    return ions ? kernel_of<GITM_VERT_AUSM, true>() : kernel_of<GITM_VERT_AUSM, false>();
  return ions ? kernel_of<GITM_VERT_RUSANOV, true>() : kernel_of<GITM_VERT_RUSANOV, false>();
}

}  // namespace

#define GV_ENS_CAT2(a, b) a##b
#define GV_ENS_CAT(a, b) GV_ENS_CAT2(a, b)

// Workgroups of this size that one CU holds at once with `lds_bytes` of LDS
// each (the persistent grid is that times the CU count); 0 on error.
int GV_ENS_CAT(ensemble_occupancy_w, GV_ENS_WAVES)(int solver, bool ions, size_t lds_bytes) {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, select_kernel(solver, ions), NT,
                                                   lds_bytes) != hipSuccess)
    return 0;
  return n;
}

// Queues one batch step: `grid` workgroups of NT threads, `lds_bytes` each.
hipError_t GV_ENS_CAT(launch_ensemble_w, GV_ENS_WAVES)(int solver, bool ions, int grid,
                                                       size_t lds_bytes,
                                                       hipStream_t stream,
                                                       const gv::KParams &kp,
                                                       const BatchMember *mem,
                                                       const EnsLayout &lay,
                                                       const double *dt_dev) {
  void *args[] = {const_cast<gv::KParams *>(&kp), &mem, const_cast<EnsLayout *>(&lay),
                  &dt_dev};
  return hipLaunchKernel(select_kernel(solver, ions), dim3(grid), dim3(NT), args,
                         lds_bytes, stream);
}
