// Task-loop shape: one workgroup owns one independent component (task) of
// one batch element and runs every iteration of it in a single launch;
// variable and factor phases are separated by workgroup barriers only, so
// the component's working set stays in that CU's caches. The kernels are
// instantiated per temperature class in pgx_task_t0.hip (max-product) and
// pgx_task_t1.hip (T > 0), which compile in parallel.
#ifndef PGX_TASK_H_
#define PGX_TASK_H_

#include "pgx_launch.h"

namespace pgx {
namespace {

// One workgroup = one (task, batch element); all iterations in one launch.
// kFused: low-degree pairwise graph, the factor side recomputes variable
// sums and no variable phase runs. With r.converge the workgroup stops after
// the first iteration whose max-abs message change is <= tol, leaving the
// messages of exactly that many iterations in x.
template <int BS, bool kT0, bool kFused, int KINDS>
__device__ __forceinline__ void task_loop_body(const DevGraph& g, const RunBufs& r, float t) {
  const int32_t ti = blockIdx.x % g.num_tasks;
  const int32_t b = blockIdx.x / g.num_tasks;
  const Task task = g.tasks[ti];
  const float* ev = r.ev + b * r.ev_stride;
  float* vsum = r.vsum + b * g.num_vs;
  float* xb = r.x + b * g.num_es;
  float* yb = r.y + b * g.num_es;
  for (int32_t it = 0; it < r.n_it; ++it) {
    const float* mo = ((r.n_it - it) & 1) ? yb : xb;
    float* mn = ((r.n_it - 1 - it) & 1) ? yb : xb;
    float dl = 0.f;
    if (kFused) {
      FusedIO io;
      io.vsum = nullptr;
      io.mo = mo;
      io.mn = mn;
      io.lp = r.lp + b * r.lp_stride;
      io.d = r.d;
      io.omd = r.omd;
      io.edge_vs0 = g.edge_vs0;
      io.even2 = g.even2 != 0;
      io.var_eptr = g.var_eptr;
      io.var_adj_es = g.var_adj_es;
      io.edge_var = g.edge_var;
      io.edge_pos4 = g.edge_pos4;
      io.packed = g.packed != 0;
      io.ev = ev;
      for (int32_t bi = task.b0; bi < task.b1; ++bi) {
        const Block blk = g.blocks[bi];
        for (int32_t i = threadIdx.x; i < blk.count; i += BS)
          fused_factor_update<kT0, KINDS>(g, blk, i, io, t, dl);
      }
    } else {
      for (int32_t v = task.v0 + threadIdx.x; v < task.v1; v += BS) var_sums(g, v, mo, ev, vsum);
      if (task.h1 > task.h0) {
        float* part = r.partial + b * g.partial_size;
        const int32_t c1 = g.heavy_cptr[task.h1];
        for (int32_t hc = g.heavy_cptr[task.h0] + threadIdx.x; hc < c1; hc += BS)
          heavy_chunk(g, hc, mo, part);
        __syncthreads();
        for (int32_t h = task.h0 + threadIdx.x; h < task.h1; h += BS)
          heavy_combine(g, h, part, ev, vsum);
      }
      __syncthreads();
      EngineIO io{vsum, mo, mn, r.lp + b * r.lp_stride, r.d, r.omd, g.edge_vs0, g.even2 != 0,
                  g.packed != 0};
      for (int32_t bi = task.b0; bi < task.b1; ++bi) {
        const Block blk = g.blocks[bi];
        for (int32_t i = threadIdx.x; i < blk.count; i += BS)
          factor_update<EngineIO, kT0, KINDS>(g, blk, i, io, t, true, dl);
      }
    }
    const uint32_t dm = block_delta<BS>(dl, r.converge ? nullptr : r.deltas + b * r.n_it + it);
    if (r.converge) {
      const int64_t slot = int64_t(b) * g.num_tasks + ti;
      const float dv = __uint_as_float(dm);
      if (threadIdx.x == 0) r.task_deltas[slot * r.n_it + it] = dv;
      if (dv <= r.tol || it == r.n_it - 1) {
        if (mn != xb) {  // move this task's messages to the output buffer
          for (int32_t bi = task.b0; bi < task.b1; ++bi) {
            const Block blk = g.blocks[bi];
            const int32_t a = g.edge_es0[g.fac_e0[blk.f0]];
            const int32_t z = g.edge_es0[g.fac_e0[blk.f0 + blk.count]];
            for (int32_t k = a + threadIdx.x; k < z; k += BS) xb[k] = mn[k];
          }
        }
        if (threadIdx.x == 0) r.iters_done[slot] = it + 1;
        return;
      }
    }
  }
}

// At T = 1 (sum-product) the temperature scalings are identities: kernels
// for T > 0 pass a constant 1 down when they can, so the compiler drops them.
template <int BS, bool kT0, bool kFused, int KINDS>
__global__ void __launch_bounds__(BS) k_task_loop(DevGraph g, RunBufs r) {
  if (!kT0 && r.t == 1.f)
    task_loop_body<BS, kT0, kFused, KINDS>(g, r, 1.f);
  else
    task_loop_body<BS, kT0, kFused, KINDS>(g, r, r.t);
}

}  // namespace

template <bool kT0, bool kFused, int KINDS>
static void launch_task_loop_t(pgx_graph* g, const RunBufs& r, int32_t batch, int bs,
                               hipStream_t s) {
  const dim3 grid(g->dev.num_tasks * batch);
  if (bs <= 256)
    hipLaunchKernelGGL((k_task_loop<256, kT0, kFused, KINDS>), grid, dim3(256), 0, s, g->dev, r);
  else if (bs <= 512)
    hipLaunchKernelGGL((k_task_loop<512, kT0, kFused, KINDS>), grid, dim3(512), 0, s, g->dev, r);
  else
    hipLaunchKernelGGL((k_task_loop<1024, kT0, kFused, KINDS>), grid, dim3(1024), 0, s, g->dev,
                       r);
}

template <bool kT0>
int launch_task_loop(pgx_graph* g, const RunBufs& r, int32_t batch, int bs,
                            hipStream_t s) {
  const int m = g->kinds_mask;
  const bool bin2_only = m == kKindsBin2;
  const bool pairwise = (m & ~kKindsPairwise) == 0;
  if (g->dev.fusable) {
    if (bin2_only)
      launch_task_loop_t<kT0, true, kKindsBin2>(g, r, batch, bs, s);
    else
      launch_task_loop_t<kT0, true, kKindsPairwise>(g, r, batch, bs, s);
  } else if (bin2_only) {
    launch_task_loop_t<kT0, false, kKindsBin2>(g, r, batch, bs, s);
  } else if (pairwise) {
    launch_task_loop_t<kT0, false, kKindsPairwise>(g, r, batch, bs, s);
  } else {
    launch_task_loop_t<kT0, false, kKindsAll>(g, r, batch, bs, s);
  }
  PGX_HIP(hipGetLastError());
  return PGX_OK;
}

}  // namespace pgx

#endif  // PGX_TASK_H_
