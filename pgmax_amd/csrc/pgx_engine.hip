// Kernels and C ABI of the loopy-BP engine (include/pgx.h).
//
// Two execution shapes for pgx_run, chosen per call:
//   * task loop: one workgroup owns one independent component of one batch
//     element and runs every iteration of it in a single launch; variable
//     and factor phases are separated by workgroup barriers only, so the
//     graph's working set stays in that CU's caches.
//   * phased: per iteration one launch for the variable phase and one for
//     the factor phase over the whole (batched) graph; used when there are
//     too few independent components to fill the GPU.
// Both run the same per-factor device code (pgx_device.h).

#include <hip/hip_cooperative_groups.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <string>
#include <vector>

#include "../../include/pgx.h"
#include "pgx_device.h"
#include "pgx_internal.h"
#include "pgx_plan.h"

namespace pgx {

thread_local std::string g_last_error;

int set_error(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

namespace {

int64_t g_opt_mode = 0;   // 0 auto, 1 task loop, 2 phased
int64_t g_opt_block = 0;  // threads per workgroup for the task loop, 0 auto
int64_t g_opt_profile = 0;
int64_t g_opt_fuse = 1;   // recompute variable sums on the factor side when the graph allows
int64_t g_opt_vec = 3;    // bit 0: 8-byte binary message loads, bit 1: 16-byte 2x2 potentials,
                          // bit 2: non-temporal 2x2 potential loads
int64_t g_opt_merge = 4096;  // components with fewer edge+variable states merge into one task
int64_t g_opt_planar = 1;    // planar message layout for 2x2 Enum blocks (read at graph create)
int64_t g_opt_tile = 0;      // factors (threads) per workgroup of the per-phase kernels, 0 auto
int64_t g_opt_graph = 0;     // 1: capture pgx_run's launches as a HIP graph and replay them
int64_t g_opt_pack = 1;      // 1: one float per binary message where the graph allows
std::mutex g_opt_mu;
std::vector<std::pair<hipEvent_t, hipEvent_t>> g_prof_events;
// per-kernel records (option "profile" = 2): kernel id, start, end
struct KernelRecord {
  int32_t id;
  hipEvent_t a, b;
  int32_t launches;  // back-to-back launches between a and b
};
std::vector<KernelRecord> g_kernel_events;

// Records one launch between two events when per-kernel profiling is on.
struct KernelClock {
  bool on = false;
  hipStream_t s = nullptr;
  template <class F>
  void run(int32_t id, F&& launch) const {
    if (!on) {
      launch();
      return;
    }
    hipEvent_t a = nullptr, b = nullptr;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
      launch();
      return;
    }
    (void)hipEventRecord(a, s);
    launch();
    (void)hipEventRecord(b, s);
    std::lock_guard<std::mutex> lock(g_opt_mu);
    g_kernel_events.push_back({id, a, b, 1});
  }
  // n back-to-back launches of one kernel between one pair of events (the
  // per-launch event records would otherwise add to every launch's time)
  template <class F>
  void run_n(int32_t id, int32_t n, F&& launches) const {
    if (!on) {
      launches();
      return;
    }
    hipEvent_t a = nullptr, b = nullptr;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
      launches();
      return;
    }
    (void)hipEventRecord(a, s);
    launches();
    (void)hipEventRecord(b, s);
    std::lock_guard<std::mutex> lock(g_opt_mu);
    g_kernel_events.push_back({id, a, b, n});
  }
};

#define PGX_HIP(call)                                                              \
  do {                                                                             \
    hipError_t _e = (call);                                                        \
    if (_e != hipSuccess)                                                          \
      return set_error(PGX_EHIP, std::string(#call) + ": " + hipGetErrorString(_e)); \
  } while (0)

constexpr int kPhaseThreads = 256;
constexpr int32_t kMaxBatchPerLaunch = 65535;  // batch elements index grid dimension y
constexpr int kTileFactors = 256;
constexpr int32_t kBin2PerThread = 4;  // 2x2 factors per thread in k_fac_phase (msets tiles)
constexpr int64_t kFusedMaxStates = int64_t(1) << 24;  // untuned fused shape up to this many
                                                       // edge states per run
constexpr uint32_t kXcds = 8;  // accelerator complex dies of an MI355X

// Block coordinates renumbered for the XCDs. Workgroups reach the kXcds XCDs
// round robin by linear id (x fastest); renumbered, each XCD runs one
// contiguous range of the grid in x-fastest order, so data that neighbouring
// workgroups share is fetched into one XCD's L2 instead of every XCD's. Used
// where every range costs the same (k_bipartite: whole batch elements, whose
// chunk sums and evidence all their tiles read); the per-phase kernels keep
// the round-robin order, which measured faster for them (uneven tiles would
// leave some XCDs with more of the work).
__device__ __forceinline__ void xcd_block(uint32_t& bx, uint32_t& by) {
  const uint64_t nx = gridDim.x, n = nx * gridDim.y;
  const uint64_t p = uint64_t(blockIdx.y) * nx + blockIdx.x;
  const uint64_t x = p % kXcds, j = p / kXcds, base = n / kXcds, rem = n % kXcds;
  const uint64_t q = x * base + (x < rem ? x : rem) + j;
  bx = uint32_t(q % nx);
  by = uint32_t(q / nx);
}

// ---------------------------------------------------------------------------
// small kernels
// ---------------------------------------------------------------------------

__global__ void k_fill_u32(uint32_t* p, int64_t n, uint32_t v) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    p[i] = v;
}

// Normalise and clip incoming messages once at the start of a run
// (pgmax/infer/bp.py:92-96, :249-259).
__global__ void k_normalize_entry(DevGraph g, const float* in, int64_t in_stride, float* out) {
  const int32_t b = blockIdx.y;
  const float* src = in + b * in_stride;
  float* dst = out + b * g.num_es;
  for (int32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < g.num_edges;
       e += gridDim.x * blockDim.x) {
    const int32_t a = g.edge_es0[e], z = g.edge_es0[e + 1];
    float mx = kNegInf;
    for (int32_t k = a; k < z; ++k) mx = nmax(mx, src[k]);
    for (int32_t k = a; k < z; ++k) dst[g.perm ? g.perm[k] : k] = nmax(src[k] - mx, kMsgNegInf);
  }
}

// Engine layout -> reference layout (planar graphs, after the last iteration).
__global__ void k_unpermute(DevGraph g, const float* in, float* out) {
  const int32_t b = blockIdx.y;
  const float* src = in + b * g.num_es;
  float* dst = out + b * g.num_es;
  for (int64_t k = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; k < g.num_es;
       k += int64_t(gridDim.x) * blockDim.x)
    dst[k] = src[g.perm[k]];
}

// Workgroup max (blockDim.x <= BS, a multiple of 64) of a non-negative (or NaN) float. For such values the
// unsigned order of the bit patterns is the float order with NaN above +inf,
// so a max over the bits is a NaN-propagating max. Thread 0 folds the result
// into *dst with atomicMax (when dst is not null); every thread gets it back.
template <int BS>
__device__ __forceinline__ uint32_t block_delta(float dl, uint32_t* dst) {
  __shared__ uint32_t red[BS / 64];
  __shared__ uint32_t res;
  uint32_t u = __float_as_uint(fabsf(dl));
  for (int o = 32; o > 0; o >>= 1) u = max(u, __shfl_xor(u, o, 64));
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) red[w] = u;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t m = red[0];
    for (int k = 1; k < int(blockDim.x >> 6); ++k) m = max(m, red[k]);
    res = m;
    if (dst) atomicMax(dst, m);
  }
  __syncthreads();
  return res;
}

// block_delta without the result broadcast (one barrier): folds the
// workgroup max into *dst.
template <int BS>
__device__ __forceinline__ void block_delta_store(float dl, uint32_t* dst) {
  __shared__ uint32_t red[BS / 64];
  uint32_t u = __float_as_uint(fabsf(dl));
  for (int o = 32; o > 0; o >>= 1) u = max(u, __shfl_xor(u, o, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = u;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t m = red[0];
    for (int k = 1; k < int(blockDim.x >> 6); ++k) m = max(m, red[k]);
    atomicMax(dst, m);
  }
}

struct RunBufs;
__device__ __forceinline__ uint32_t* delta_slot(const RunBufs& r, int32_t b, int32_t it);

struct RunBufs {
  const float* lp;
  int64_t lp_stride;
  const float* ev;
  int64_t ev_stride;
  float* x;  // final output buffer (ftov_out)
  float* y;  // scratch message buffer
  float* vsum;
  float* partial;  // heavy-variable chunk sums, batch x partial_size
  uint32_t* deltas;
  // Per-launch kernels with many workgroups per (batch element, iteration)
  // fold their deltas into dslot_n slots per pair (a workgroup picks slot
  // blockIdx.x % dslot_n), reduced into deltas after the run: one address
  // taking every workgroup's atomic serialises them. dslot_n = 1: dslot is
  // deltas.
  uint32_t* dslot;
  int32_t dslot_n;
  int32_t n_it;
  float d, omd, t;
  // convergence runs (pgx_run_converge): per-task deltas and early exit
  int32_t converge;
  float tol;
  float* task_deltas;  // batch x num_tasks x n_it
  int32_t* iters_done;  // batch x num_tasks
};

__device__ __forceinline__ uint32_t* delta_slot(const RunBufs& r, int32_t b, int32_t it) {
  const int64_t pair = int64_t(b) * r.n_it + it;
  return r.dslot + pair * r.dslot_n + (r.dslot_n > 1 ? int32_t(blockIdx.x) % r.dslot_n : 0);
}

// deltas[i] = max over the dslot_n slots of pair i (unsigned bits: NaN-propagating).
__global__ void k_delta_reduce(const uint32_t* dslot, int32_t n_slot, uint32_t* deltas,
                               int64_t pairs) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < pairs;
       i += int64_t(gridDim.x) * blockDim.x) {
    uint32_t m = 0;
    for (int32_t k = 0; k < n_slot; ++k) m = max(m, dslot[i * n_slot + k]);
    deltas[i] = m;
  }
}

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

// Complete bipartite graphs, and in the per-phase shape all graphs of 2x2
// factors whose variables have at most 4 edges, iterate on one float per
// binary edge (pair_encode), stride floats per batch element. k_pair_pack
// turns the engine-layout messages (planar pairs) into that form;
// k_pair_unpack writes the caller's reference layout from it.
__global__ void k_pair_pack(DevGraph g, const float* full, float* packed, int64_t stride) {
  const int32_t b = blockIdx.y;
  const int64_t npair = g.num_es / 2;
  const float* src = full + b * g.num_es;
  float* dst = packed + b * stride;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < npair;
       i += int64_t(gridDim.x) * blockDim.x) {
    const float2 m = ld2(src + 2 * i);
    dst[i] = pair_encode(m.x, m.y);
  }
}

__global__ void k_pair_unpack(DevGraph g, const float* packed, int64_t stride, float* out) {
  const int32_t b = blockIdx.y;
  const float* src = packed + b * stride;
  float* dst = out + b * g.num_es;
  for (int64_t k = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; k < g.num_es;
       k += int64_t(gridDim.x) * blockDim.x) {
    const int32_t p = g.perm[k];
    const float2 m = pair_decode(src[p >> 1]);
    dst[k] = (p & 1) ? m.y : m.x;
  }
}

// Phased shape: variable phase over all variables of all batch elements.
// Blocks [0, light_blocks) take one light variable per thread, the rest one
// heavy-variable chunk per thread (k_var_combine then adds the chunks).
__global__ void __launch_bounds__(kPhaseThreads)
    k_var_phase(DevGraph g, const float* m, int64_t m_stride, const float* ev,
                int64_t ev_stride, float* out, float* partial, int32_t light_blocks,
                int32_t col_blocks) {
  const int32_t b = blockIdx.y;
  const int32_t blk_x = blockIdx.x;
  if (blk_x < light_blocks) {
    const int32_t v = blk_x * kPhaseThreads + threadIdx.x;
    if (v < g.num_vars)
      var_sums(g, v, m + b * m_stride, ev + b * ev_stride, out + b * g.num_vs);
    return;
  }
  if (blk_x < light_blocks + col_blocks) {
    const int32_t i = (blk_x - light_blocks) * kPhaseThreads + threadIdx.x;
    if (i < g.num_heavy_col)
      heavy_chunk_d(g, g.col_desc ? g.col_desc[i] : g.chunk_desc[i], m + b * m_stride,
                    partial + b * g.partial_size);
    return;
  }
  __shared__ float stage[(kPhaseThreads / kRowLanes) * kRowStride];
  const int32_t grp = int32_t(threadIdx.x) / kRowLanes;
  const int32_t i = (blk_x - light_blocks - col_blocks) * (kPhaseThreads / kRowLanes) + grp;
  heavy_chunk_row(g, i < g.num_heavy_row ? i : -1, m + b * m_stride,
                  partial + b * g.partial_size, int32_t(threadIdx.x) % kRowLanes,
                  stage + grp * kRowStride);
}

// Group sums of the variables with more than kVarChunk chunks.
__global__ void __launch_bounds__(kPhaseThreads) k_var_groups(DevGraph g, float* partial) {
  const int32_t b = blockIdx.y;
  const int32_t gi = blockIdx.x * kPhaseThreads + threadIdx.x;
  if (gi < g.num_vh_groups) heavy_group(g, gi, partial + b * g.partial_size);
}

// Heavy variables: chunk (or group) sums in order, then + evidence.
__global__ void __launch_bounds__(kPhaseThreads)
    k_var_combine(DevGraph g, const float* partial, const float* ev, int64_t ev_stride,
                  float* out) {
  const int32_t b = blockIdx.y;
  const int32_t h = blockIdx.x * kPhaseThreads + threadIdx.x;
  if (h < g.num_heavy)
    heavy_combine(g, h, partial + b * g.partial_size, ev + b * ev_stride, out + b * g.num_vs,
                  true);
}

static inline dim3 var_grid(const DevGraph& d, int32_t batch, int32_t* light_blocks,
                            int32_t* col_blocks) {
  *light_blocks = (d.num_vars + kPhaseThreads - 1) / kPhaseThreads;
  *col_blocks = (d.num_heavy_col + kPhaseThreads - 1) / kPhaseThreads;
  const int32_t per = kPhaseThreads / kRowLanes;
  const int32_t row_blocks = (d.num_heavy_row + per - 1) / per;
  return dim3(*light_blocks + *col_blocks + row_blocks, batch);
}

static inline void launch_var_phase(const DevGraph& d, const float* m, int64_t m_stride,
                                    const float* ev, int64_t ev_stride, float* out,
                                    float* partial, int32_t batch, hipStream_t s,
                                    const KernelClock* clk = nullptr) {
  int32_t light_blocks, col_blocks;
  const dim3 vgrid = var_grid(d, batch, &light_blocks, &col_blocks);
  const KernelClock off;
  const KernelClock& c = clk ? *clk : off;
  c.run(PGX_KERNEL_VAR_PHASE, [&] {
    hipLaunchKernelGGL(k_var_phase, vgrid, dim3(kPhaseThreads), 0, s, d, m, m_stride, ev,
                       ev_stride, out, partial, light_blocks, col_blocks);
  });
  if (d.num_vh_groups > 0)
    c.run(PGX_KERNEL_VAR_COMBINE, [&] {
      hipLaunchKernelGGL(k_var_groups,
                         dim3((d.num_vh_groups + kPhaseThreads - 1) / kPhaseThreads, batch),
                         dim3(kPhaseThreads), 0, s, d, partial);
    });
  if (d.num_heavy > 0)
    c.run(PGX_KERNEL_VAR_COMBINE, [&] {
      hipLaunchKernelGGL(k_var_combine,
                         dim3((d.num_heavy + kPhaseThreads - 1) / kPhaseThreads, batch),
                         dim3(kPhaseThreads), 0, s, d, partial, ev, ev_stride, out);
    });
}

struct Tile {
  int32_t block, i0, i1, pad;
};

// Phased shape: factor phase over tiles of <= kTileFactors factors.
template <bool kT0, int KIND>
__device__ __forceinline__ void fac_phase_body(const DevGraph& g, const Tile* tiles,
                                               const RunBufs& r, int32_t it, float t) {
  const int32_t b = blockIdx.y;
  const Tile tile = tiles[blockIdx.x];
  const Block blk = g.blocks[tile.block];
  const float* xb = r.x + b * g.num_es;
  const float* yb = r.y + b * g.num_es;
  const float* mo = ((r.n_it - it) & 1) ? yb : xb;
  float* mn = const_cast<float*>(((r.n_it - 1 - it) & 1) ? yb : xb);
  EngineIO io{r.vsum + b * g.num_vs, mo, mn, r.lp + b * r.lp_stride, r.d, r.omd, g.edge_vs0,
              g.even2 != 0, g.packed != 0};
  float dl = 0.f;
  if constexpr (KIND == kOr || KIND == kAnd) {
    if (tile.pad == kWideLanes) {  // one lane group per factor
      const int32_t i = tile.i0 + int32_t(threadIdx.x) / kWideLanes;
      if (i < tile.i1)
        logical_factor_wide<EngineIO>(g, blk.f0 + i, blk.off, io, t, true, dl,
                                      int32_t(threadIdx.x) % kWideLanes);
      block_delta_store<kPhaseThreads>(dl, delta_slot(r, b, it));
      return;
    }
  }
  if constexpr (KIND == kEnumGeneral) {
    if (tile.pad == 64) {  // one wavefront per factor
      __shared__ float wl[kPhaseThreads / 64][kWaveMaxConfigs + kWaveMaxStates];
      const int32_t w = int32_t(threadIdx.x) >> 6;
      const int32_t i = tile.i0 + w;
      enum_general_wave<EngineIO, kT0>(g, blk, i, i < tile.i1, io, t, dl,
                                       int32_t(threadIdx.x) & 63, wl[w], wl[w] + kWaveMaxConfigs);
      block_delta_store<kPhaseThreads>(dl, delta_slot(r, b, it));
      return;
    }
  }
  if constexpr (KIND == kEnumBin2) {
    if (tile.pad == -kBin2PerThread) {  // kBin2PerThread factors per thread, strided
      int32_t idx[kBin2PerThread];
#pragma unroll
      for (int k = 0; k < kBin2PerThread; ++k) {
        const int32_t i = tile.i0 + int32_t(threadIdx.x) + k * int32_t(blockDim.x);
        idx[k] = i < tile.i1 ? i : -1;
      }
      if (g.even2) {
        enum_bin2_n<kT0, kBin2PerThread>(g, blk, idx, io, t, dl);
      } else {
        for (int k = 0; k < kBin2PerThread; ++k)
          if (idx[k] >= 0) factor_update<EngineIO, kT0, (1 << KIND)>(g, blk, idx[k], io, t, true, dl);
      }
      block_delta_store<kPhaseThreads>(dl, delta_slot(r, b, it));
      return;
    }
  }
  const int32_t i = tile.i0 + threadIdx.x;
  if (i < tile.i1) factor_update<EngineIO, kT0, (1 << KIND)>(g, blk, i, io, t, true, dl);
  block_delta_store<kPhaseThreads>(dl, delta_slot(r, b, it));
}

template <bool kT0, int KIND>
__global__ void __launch_bounds__(kPhaseThreads)
    k_fac_phase(DevGraph g, const Tile* tiles, RunBufs r, int32_t it) {
  if (!kT0 && r.t == 1.f)
    fac_phase_body<kT0, KIND>(g, tiles, r, it, 1.f);
  else
    fac_phase_body<kT0, KIND>(g, tiles, r, it, r.t);
}

// Phased shape for fusable graphs: the whole iteration in one launch.
template <bool kT0, int KINDS>
__device__ __forceinline__ void fac_fused_body(const DevGraph& g, const Tile* tiles,
                                               const RunBufs& r, int32_t it, float t) {
  const int32_t b = blockIdx.y;
  const Tile tile = tiles[blockIdx.x];
  const Block blk = g.blocks[tile.block];
  const float* xb = r.x + b * g.num_es;
  const float* yb = r.y + b * g.num_es;
  FusedIO io;
  io.vsum = nullptr;
  io.mo = ((r.n_it - it) & 1) ? yb : xb;
  io.mn = const_cast<float*>(((r.n_it - 1 - it) & 1) ? yb : xb);
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
  io.ev = r.ev + b * r.ev_stride;
  float dl = 0.f;
  const int32_t i = tile.i0 + threadIdx.x;
  if (i < tile.i1) fused_factor_update<kT0, KINDS>(g, blk, i, io, t, dl);
  block_delta_store<kPhaseThreads>(dl, delta_slot(r, b, it));
}

template <bool kT0, int KINDS>
__global__ void __launch_bounds__(kPhaseThreads)
    k_fac_fused(DevGraph g, const Tile* tiles, RunBufs r, int32_t it) {
  if (!kT0 && r.t == 1.f)
    fac_fused_body<kT0, KINDS>(g, tiles, r, it, 1.f);
  else
    fac_fused_body<kT0, KINDS>(g, tiles, r, it, r.t);
}

// Cooperative shape for fusable graphs: every iteration in one launch of
// co-resident workgroups that stride over the (tile, batch element) units,
// with a grid-wide barrier between iterations.
template <bool kT0, int KINDS>
__global__ void __launch_bounds__(kPhaseThreads)
    k_coop_fused(DevGraph g, const Tile* tiles, int32_t num_tiles, RunBufs r, int32_t batch) {
  cooperative_groups::grid_group grid = cooperative_groups::this_grid();
  const int64_t units = int64_t(num_tiles) * batch;
  for (int32_t it = 0; it < r.n_it; ++it) {
    for (int64_t w = blockIdx.x; w < units; w += gridDim.x) {
      const int32_t b = int32_t(w / num_tiles);
      const Tile tile = tiles[w % num_tiles];
      const Block blk = g.blocks[tile.block];
      const float* xb = r.x + b * g.num_es;
      const float* yb = r.y + b * g.num_es;
      FusedIO io;
      io.vsum = nullptr;
      io.mo = ((r.n_it - it) & 1) ? yb : xb;
      io.mn = const_cast<float*>(((r.n_it - 1 - it) & 1) ? yb : xb);
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
      io.ev = r.ev + b * r.ev_stride;
      float dl = 0.f;
      const int32_t i = tile.i0 + threadIdx.x;
      if (i < tile.i1) fused_factor_update<kT0, KINDS>(g, blk, i, io, r.t, dl);
      block_delta<kPhaseThreads>(dl, r.deltas + b * r.n_it + it);
    }
    grid.sync();
  }
}

// Complete bipartite binary graphs (Plan::bip_*): the whole iteration in
// one launch with no separate variable phase. A workgroup owns kBipRows row
// variables x kCols column variables of one batch element, which are whole
// chunks of both sides' edge lists. It forms its variables' sums from the
// previous iteration's chunk sums (the order of k_var_combine), updates the
// tile's factors (enum_bin2's arithmetic), and sums the new messages of each
// chunk in edge order from 0 (the order of k_var_phase) into part_out.
// Reading and writing different chunk-sum buffers keeps the workgroups of
// one launch independent. Thread t takes column t % kCols of rows
// t / kCols, t / kCols + kThreads / kCols, ...
constexpr int kBipRows = kVarChunk;  // rows per tile: one column chunk

template <bool kT0, int kCols, int kThreads>
__global__ void __launch_bounds__(kThreads)
    k_bipartite(DevGraph g, RunBufs r, const float* msg_in, float* msg_out,
                const float* part_in, float* part_out, int32_t it, int32_t col_tiles,
                int32_t batch) {
  static_assert(kCols % kVarChunk == 0 && kThreads % kCols == 0, "whole row chunks per tile");
  static_assert(kThreads >= 2 * kBipRows + 2 * kCols, "two lanes per variable sum");
  constexpr int kStep = kThreads / kCols;  // rows one pass of the workgroup covers
  constexpr int kPer = kBipRows / kStep;   // rows per thread
  static_assert(kPer * kStep == kBipRows, "rows split evenly");
  __shared__ float2 srow[kBipRows][kCols + 1];  // new first-edge messages (padded)
  __shared__ float2 scol[kBipRows][kCols];      // new second-edge messages
  __shared__ float2 vrow[kBipRows];             // variable sums of the tile's rows
  __shared__ float2 vcol[kCols];                // and columns
  // Workgroups go to the 8 XCDs round robin by id; number them so that each
  // XCD runs whole batch elements, whose chunk sums and evidence then stay
  // in that XCD's L2.
  uint32_t bx, by;
  xcd_block(bx, by);
  const int32_t wg = int32_t(bx);
  const int32_t tiles = int32_t(gridDim.x / batch);
  const int32_t b = wg / tiles;
  const int32_t rt = (wg % tiles) / col_tiles, ct = (wg % tiles) % col_tiles;
  const Block blk = g.blocks[0];
  const int32_t na = g.bip_na, nb = g.bip_nb;
  const int32_t a0 = rt * kBipRows, c0 = ct * kCols;
  const int32_t nr = min(kBipRows, na - a0), ncl = min(kCols, nb - c0);
  const int32_t tid = threadIdx.x;
  const int32_t col = tid % kCols, row0 = tid / kCols;
  // messages as one float per edge (pair_encode): first edges at [0, n),
  // second edges at [n, 2n) of each batch element's num_es / 2 floats
  const int64_t npair = g.num_es / 2;
  const float* mo = msg_in + b * npair;
  float* mn = msg_out + b * npair;
  const float* lp = r.lp + b * r.lp_stride;
  const bool col_ok = col < ncl;
  // Variable sums: two lanes per variable, rows in wave 0 and columns in
  // the next 2 * kCols threads. The even lane adds chunks 0 .. kSumLoads-1
  // from 0 and the odd lane continues with the next kSumLoads -- the one
  // sequential sum chunk_total2 forms. These loads and the tile's message
  // and log-potential loads have clamped (always valid) indices and are all
  // issued before any is used, so the chunk sums wait only for their own
  // loads. Variables of more chunks use chunk_total2.
  constexpr int kSumLoads = 8;
  const float* pin = part_in + b * g.partial_size;
  const float* ev = r.ev + b * r.ev_stride;
  const int32_t nch_row = (nb + kVarChunk - 1) / kVarChunk;
  const int32_t nch_col = (na + kVarChunk - 1) / kVarChunk;
  const bool fast = max(nch_row, nch_col) <= 2 * kSumLoads;  // uniform
  const bool is_row = tid < 2 * kBipRows;
  const bool is_col = !is_row && tid < 2 * kBipRows + 2 * kCols;
  const int32_t vi = (is_row ? tid : tid - 2 * kBipRows) >> 1, half = tid & 1;
  const int32_t vg = is_row ? a0 + min(vi, nr - 1) : c0 + min(is_col ? vi : 0, ncl - 1);
  int2 vinfo;
  if (g.bip_affine) {  // no table load ahead of the chunk-sum loads
    vinfo = is_row ? make_int2(g.bip_aff[0] + vg * g.bip_aff[1], g.bip_aff[2] + vg * g.bip_aff[3])
                   : make_int2(g.bip_aff[4] + vg * g.bip_aff[5], g.bip_aff[6] + vg * g.bip_aff[7]);
  } else {
    vinfo = is_row ? g.bip_row[vg] : g.bip_col[vg];
  }
  const int32_t nch = is_row ? nch_row : nch_col;
  float2 cv[kSumLoads];
#pragma unroll
  for (int u = 0; u < kSumLoads; ++u)
    cv[u] = ld2(pin + vinfo.y + 2 * min(half * kSumLoads + u, nch - 1));
  const float2 evv = make_float2(ev[vinfo.x], ev[vinfo.x + 1]);
  // the tile's old messages and raw log-potentials
  float ua[kPer], ub[kPer];  // encoded old messages
  float4 lq[kPer];
  const int32_t cc = min(col, ncl - 1);
  if ((blk.pad[0] & 1) && g.lp_al4) {
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const int32_t f = (a0 + min(row0 + kStep * k, nr - 1)) * nb + c0 + cc;
      ua[k] = mo[f];
      ub[k] = mo[blk.count + f];
      lq[k] = ld4(lp + blk.c0 + 4 * f);
    }
  } else {
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const int32_t f = (a0 + min(row0 + kStep * k, nr - 1)) * nb + c0 + cc;
      const float* q = lp + blk.c0 + 4 * f;
      ua[k] = mo[f];
      ub[k] = mo[blk.count + f];
      lq[k] = make_float4(q[0], q[1], q[2], q[3]);
    }
  }
  if (fast) {
    float s0 = 0.f, s1 = 0.f;
    if (half == 0) {
#pragma unroll
      for (int u = 0; u < kSumLoads; ++u)
        if (u < nch) {
          s0 += cv[u].x;
          s1 += cv[u].y;
        }
    }
    const float p0 = __shfl_xor(s0, 1), p1 = __shfl_xor(s1, 1);
    if (half == 1) {
      s0 = p0;
      s1 = p1;
#pragma unroll
      for (int u = 0; u < kSumLoads; ++u)
        if (kSumLoads + u < nch) {
          s0 += cv[u].x;
          s1 += cv[u].y;
        }
      float t0 = 0.f, t1 = 0.f;  // the group total, as chunk_total2 adds it
      t0 += s0;
      t1 += s1;
      if (is_row && vi < nr) vrow[vi] = make_float2(t0 + evv.x, t1 + evv.y);
      if (is_col && vi < ncl) vcol[vi] = make_float2(t0 + evv.x, t1 + evv.y);
    }
  } else if (tid < nr) {
    const int2 ri = g.bip_row[a0 + tid];
    float t0, t1;
    chunk_total2(pin + ri.y, nch_row, t0, t1);
    vrow[tid] = make_float2(t0 + ev[ri.x], t1 + ev[ri.x + 1]);
  } else if (tid >= 64 && tid - 64 < ncl) {
    const int2 ci = g.bip_col[c0 + tid - 64];
    float t0, t1;
    chunk_total2(pin + ci.y, nch_col, t0, t1);
    vcol[tid - 64] = make_float2(t0 + ev[ci.x], t1 + ev[ci.x + 1]);
  }
  __syncthreads();
  float dl = 0.f;
  const float2 vc = col_ok ? vcol[col] : make_float2(0.f, 0.f);
  auto update = [&](float t) {
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const int32_t ra = row0 + kStep * k;
      if (col_ok && ra < nr) {
        const int32_t f = (a0 + ra) * nb + c0 + col;
        const float2 vr = vrow[ra];
        const float2 oa = pair_decode(ua[k]), ob = pair_decode(ub[k]);
        const float4 lc = make_float4(clip_lp(lq[k].x), clip_lp(lq[k].y), clip_lp(lq[k].z),
                                      clip_lp(lq[k].w));
        float fa0, fa1, fb0, fb1;
        bin2_ftov<kT0>(vr.x - oa.x, vr.y - oa.y, vc.x - ob.x, vc.y - ob.y, lc, t, fa0, fa1, fb0,
                       fb1);
        float na0, na1, nb0, nb1;
        damp_norm2(r.d, r.omd, oa.x, oa.y, fa0, fa1, na0, na1, dl);
        damp_norm2(r.d, r.omd, ob.x, ob.y, fb0, fb1, nb0, nb1, dl);
        mn[f] = pair_encode(na0, na1);
        mn[blk.count + f] = pair_encode(nb0, nb1);
        srow[ra][col] = make_float2(na0, na1);
        scol[ra][col] = make_float2(nb0, nb1);
      }
    }
  };
  // at T = 1 (sum-product) the temperature scalings are identities: a
  // constant argument lets the compiler drop them
  if (!kT0 && r.t == 1.f)
    update(1.f);
  else
    update(r.t);
  __syncthreads();
  float* pout = part_out + b * g.partial_size;
  if (tid < kCols) {  // row chunks: thread = (run, row)
    const int32_t ra = tid % kBipRows, cb = (tid / kBipRows) * kVarChunk;
    const int32_t len = min(kVarChunk, ncl - cb);
    if (ra < nr && len > 0) {
      float s0 = 0.f, s1 = 0.f;
      if (len == kVarChunk) {
        float2 m[kVarChunk];
#pragma unroll
        for (int u = 0; u < kVarChunk; ++u) m[u] = srow[ra][cb + u];
#pragma unroll
        for (int u = 0; u < kVarChunk; ++u) {
          s0 += m[u].x;
          s1 += m[u].y;
        }
      } else {
        for (int32_t u = 0; u < len; ++u) {
          s0 += srow[ra][cb + u].x;
          s1 += srow[ra][cb + u].y;
        }
      }
      st2(pout + g.bip_row[a0 + ra].y + 2 * ((c0 + cb) / kVarChunk), s0, s1);
    }
  } else if (tid < 2 * kCols && tid - kCols < ncl) {  // column chunks: thread = column
    const int32_t c = tid - kCols;
    float s0 = 0.f, s1 = 0.f;
    if (nr == kBipRows) {
      float2 m[kBipRows];
#pragma unroll
      for (int u = 0; u < kBipRows; ++u) m[u] = scol[u][c];
#pragma unroll
      for (int u = 0; u < kBipRows; ++u) {
        s0 += m[u].x;
        s1 += m[u].y;
      }
    } else {
      for (int32_t u = 0; u < nr; ++u) {
        s0 += scol[u][c].x;
        s1 += scol[u][c].y;
      }
    }
    st2(pout + g.bip_col[c0 + c].y + 2 * rt, s0, s1);
  }
  block_delta_store<kThreads>(dl, delta_slot(r, b, it));
}

template <bool kT0, int kCols, int kThreads>
static void launch_bipartite_t(const DevGraph& d, const RunBufs& r, const float* msg_in,
                               float* msg_out, const float* pin, float* pout, int32_t it,
                               int32_t batch, hipStream_t s) {
  const int32_t col_tiles = (d.bip_nb + kCols - 1) / kCols;
  const int32_t row_tiles = (d.bip_na + kBipRows - 1) / kBipRows;
  hipLaunchKernelGGL((k_bipartite<kT0, kCols, kThreads>),
                     dim3(unsigned(row_tiles * col_tiles * batch)), dim3(kThreads), 0, s, d, r,
                     msg_in, msg_out, pin, pout, it, col_tiles, batch);
}

// Tile shapes of k_bipartite, (columns, threads) from the "tile" and
// "block" options: (32, 128), (32, 256), (64, 256), (64, 512), (128, 512);
// tile 32 / 64 / 128 columns, block 128 / 256 / 512 threads (else 256), the
// nearest existing pair. Other tile values pick 32 columns (the smallest
// LDS footprint: most workgroups per CU).
template <bool kT0>
static void launch_bipartite(const DevGraph& d, const RunBufs& r, const float* msg_in, float* msg_out,
                             const float* pin, float* pout, int32_t it, int32_t batch,
                             hipStream_t s, int64_t cols, int64_t threads) {
  if (cols != 32 && cols != 64 && cols != 128) cols = 32;
  if (cols == 32 && threads == 128)
    launch_bipartite_t<kT0, 32, 128>(d, r, msg_in, msg_out, pin, pout, it, batch, s);
  else if (cols == 32)
    launch_bipartite_t<kT0, 32, 256>(d, r, msg_in, msg_out, pin, pout, it, batch, s);
  else if (cols == 128)
    launch_bipartite_t<kT0, 128, 512>(d, r, msg_in, msg_out, pin, pout, it, batch, s);
  else if (threads == 512)
    launch_bipartite_t<kT0, 64, 512>(d, r, msg_in, msg_out, pin, pout, it, batch, s);
  else
    launch_bipartite_t<kT0, 64, 256>(d, r, msg_in, msg_out, pin, pout, it, batch, s);
}

// One factor type's update on given vtof (plugin API).
template <bool kT0, int KIND>
__global__ void __launch_bounds__(kPhaseThreads)
    k_fac_op(DevGraph g, const Tile* tiles, const float* vtof, const float* lp, float* out,
             float t, int32_t normalize) {
  const Tile tile = tiles[blockIdx.x];
  const Block blk = g.blocks[tile.block];
  OpIO io{vtof, out, lp};
  float dl = 0.f;
  if constexpr (KIND == kEnumGeneral) {
    if (tile.pad == 64) {  // one wavefront per factor
      __shared__ float wl[kPhaseThreads / 64][kWaveMaxConfigs + kWaveMaxStates];
      const int32_t w = int32_t(threadIdx.x) >> 6;
      const int32_t i = tile.i0 + w;
      enum_general_wave<OpIO, kT0>(g, blk, i, i < tile.i1, io, t, dl, int32_t(threadIdx.x) & 63,
                                   wl[w], wl[w] + kWaveMaxConfigs);
      return;
    }
  }
  const int32_t i = tile.i0 + threadIdx.x;
  if (i < tile.i1) factor_update<OpIO, kT0, (1 << KIND)>(g, blk, i, io, t, normalize != 0, dl);
}

// First maximal state of every variable (pgmax/infer/inferer.py:251-264).
__global__ void k_decode(DevGraph g, const float* beliefs, int32_t* out) {
  const int32_t b = blockIdx.y;
  const int32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= g.num_vars) return;
  const float* x = beliefs + b * g.num_vs;
  const int32_t a = g.var_vs0[v], z = g.var_vs0[v + 1];
  int32_t best = 0;
  float bv = x[a];
  for (int32_t k = a + 1; k < z && bv == bv; ++k)
    if (x[k] > bv || x[k] != x[k]) {
      bv = x[k];
      best = k - a;
    }
  out[int64_t(b) * g.num_vars + v] = best;
}

// Row-level Enum op (pgx_enum_rows_fac_to_var): summaries per config ...
__global__ void k_enum_rows_summary(const int32_t* cfg_row_ptr, const int32_t* row_es,
                                    int64_t num_configs, const float* vtof, const float* lp,
                                    float* summary) {
  for (int64_t c = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; c < num_configs;
       c += int64_t(gridDim.x) * blockDim.x) {
    float s = 0.f;
    for (int32_t r = cfg_row_ptr[c]; r < cfg_row_ptr[c + 1]; ++r) s += vtof[row_es[r]];
    summary[c] = s + lp[c];
  }
}

// ... then max / T-logsumexp per edge state over its rows in row order.
template <bool kT0>
__global__ void k_enum_rows_out(const int32_t* es_row_ptr, const int32_t* es_row_cfg,
                                int64_t num_es, const float* vtof, const float* summary,
                                float* ftov, float t) {
  for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < num_es;
       e += int64_t(gridDim.x) * blockDim.x) {
    const int32_t a = es_row_ptr[e], z = es_row_ptr[e + 1];
    float mx = kNegInf;
    for (int32_t k = a; k < z; ++k) mx = nmax(mx, summary[es_row_cfg[k]]);
    float out = mx;
    if (!kT0) {
      float u = 0.f;
      for (int32_t k = a; k < z; ++k) u += exp_rel(summary[es_row_cfg[k]] - mx, t);
      out = mul_t(pgx_log(u), t) + mx;
    }
    ftov[e] = out - vtof[e];
  }
}

template <class T>
int upload(const std::vector<T>& v, T** dst) {
  *dst = nullptr;
  if (v.empty()) return PGX_OK;
  PGX_HIP(hipMalloc(reinterpret_cast<void**>(dst), v.size() * sizeof(T)));
  PGX_HIP(hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return PGX_OK;
}

}  // namespace
}  // namespace pgx

using namespace pgx;

namespace pgx {
// Execution shape of one run: mode 1 task loop / 2 phased, workgroup size of
// the task loop, factor-side variable sums, vector accesses.
struct RunCfg {
  int64_t mode = 0, bs = 0, fuse = 1, vec = 3, tile = 256;
};
}  // namespace pgx

namespace pgx {
// Factor tiles of one size, grouped by block kind.
struct TileSet {
  Tile* tiles = nullptr;
  int32_t kind_tile0[7] = {0, 0, 0, 0, 0, 0, 0};
  int32_t num_tiles = 0;
  int32_t size = 256;
};
constexpr int kTileSizes[3] = {64, 128, 256};
}  // namespace pgx

struct pgx_graph {
  Plan plan;
  TileSet tsets[3];  // tiles of 64, 128 and 256 factors (threads per workgroup)
  TileSet msets[3];  // the same with kBin2PerThread 2x2 factors per thread
  // shapes chosen by pgx_autotune, keyed by (batch, temperature == 0, converge)
  std::map<std::tuple<int32_t, int, int>, RunCfg> tuned;
  // runs captured as HIP graphs (option "graph"), keyed by every launch argument
  std::map<std::vector<int64_t>, hipGraphExec_t> graphs;
  std::vector<std::vector<int64_t>> graph_order;  // insertion order, for eviction
  DevGraph dev{};
  std::vector<void*> owned;
  Tile* tiles = nullptr;  // all factor tiles, grouped by factor type
  int32_t kind_tile0[7] = {0, 0, 0, 0, 0, 0, 0};  // tiles are sorted by block kind
  int32_t kinds_mask = 0;
  int32_t num_tiles = 0;
  int32_t fusable_plan = 0;
  int32_t bip_plan = 0;     // complete bipartite binary graph (k_bipartite applies)
  int32_t packed_plan = 0;  // 2x2 factors only, binary variables of <= 4 edges (Plan::packable)
  bool even2_plan = false;  // 2-state edges / variables at even offsets
  int32_t* zero_strides = nullptr;  // NHC zeros: chunk strides of the reference layout
  const int4* chunk_desc_ref = nullptr;  // chunk descriptors of the reference layout
  float* scratch = nullptr;  // message buffer + variable sums + chunk sums, grown on demand
  size_t scratch_bytes = 0;
  float* bscratch = nullptr;  // chunk sums for pgx_get_beliefs
  size_t bscratch_bytes = 0;
};

extern "C" {

const char* pgx_last_error(void) { return g_last_error.c_str(); }

int64_t pgx_set_option(const char* key, int64_t value) {
  std::lock_guard<std::mutex> lock(g_opt_mu);
  int64_t prev = -1;
  if (key && std::strcmp(key, "mode") == 0) {
    prev = g_opt_mode;
    g_opt_mode = value;
  } else if (key && std::strcmp(key, "block") == 0) {
    prev = g_opt_block;
    g_opt_block = value;
  } else if (key && std::strcmp(key, "profile") == 0) {
    prev = g_opt_profile;
    g_opt_profile = value;
  } else if (key && std::strcmp(key, "fuse") == 0) {
    prev = g_opt_fuse;
    g_opt_fuse = value;
  } else if (key && std::strcmp(key, "vec") == 0) {
    prev = g_opt_vec;
    g_opt_vec = value;
  } else if (key && std::strcmp(key, "graph") == 0) {
    prev = g_opt_graph;
    g_opt_graph = value;
  } else if (key && std::strcmp(key, "tile") == 0) {
    prev = g_opt_tile;
    g_opt_tile = value;
  } else if (key && std::strcmp(key, "planar") == 0) {
    prev = g_opt_planar;
    g_opt_planar = value;
  } else if (key && std::strcmp(key, "pack") == 0) {
    prev = g_opt_pack;
    g_opt_pack = value;
  } else if (key && std::strcmp(key, "merge") == 0) {
    prev = g_opt_merge;
    g_opt_merge = value;
  }
  return prev;
}

int pgx_profile_read(double* total_ms, int64_t* num_runs) {
  std::vector<std::pair<hipEvent_t, hipEvent_t>> evs;
  {
    std::lock_guard<std::mutex> lock(g_opt_mu);
    evs.swap(g_prof_events);
  }
  double total = 0.0;
  int rc = PGX_OK;
  for (auto& p : evs) {
    float ms = 0.f;
    hipError_t e = hipEventSynchronize(p.second);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, p.first, p.second);
    if (e != hipSuccess && rc == PGX_OK)
      rc = set_error(PGX_EHIP, std::string("pgx_profile_read: ") + hipGetErrorString(e));
    total += ms;
    (void)hipEventDestroy(p.first);
    (void)hipEventDestroy(p.second);
  }
  if (total_ms) *total_ms = total;
  if (num_runs) *num_runs = int64_t(evs.size());
  return rc;
}

int pgx_profile_kernels(double* total_ms, int64_t* launches) {
  std::vector<KernelRecord> evs;
  {
    std::lock_guard<std::mutex> lock(g_opt_mu);
    evs.swap(g_kernel_events);
  }
  for (int k = 0; k < PGX_KERNEL_COUNT; ++k) {
    if (total_ms) total_ms[k] = 0.0;
    if (launches) launches[k] = 0;
  }
  int rc = PGX_OK;
  for (auto& p : evs) {
    float ms = 0.f;
    hipError_t e = hipEventSynchronize(p.b);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, p.a, p.b);
    if (e != hipSuccess && rc == PGX_OK)
      rc = set_error(PGX_EHIP, std::string("pgx_profile_kernels: ") + hipGetErrorString(e));
    if (p.id >= 0 && p.id < PGX_KERNEL_COUNT) {
      if (total_ms) total_ms[p.id] += ms;
      if (launches) launches[p.id] += p.launches;
    }
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  return rc;
}

int pgx_graph_create(const pgx_wiring* wiring, pgx_graph** out) {
  if (!wiring || !out) return set_error(PGX_EINVAL, "null argument");
  *out = nullptr;
  pgx_graph* g = new pgx_graph();
  std::string err;
  int64_t merge, planar;
  {
    std::lock_guard<std::mutex> lock(g_opt_mu);
    merge = g_opt_merge;
    planar = g_opt_planar;
  }
  int rc = build_plan(*wiring, &g->plan, &err, merge, planar != 0);
  if (rc != PGX_OK) {
    delete g;
    return set_error(rc, "pgx_graph_create: " + err);
  }
  const Plan& P = g->plan;
  DevGraph& d = g->dev;
  auto up = [&](const auto& v, auto** dst) {
    int r2 = upload(v, dst);
    if (r2 == PGX_OK && *dst) g->owned.push_back(static_cast<void*>(*dst));
    return r2;
  };
  int32_t *var_vs0, *var_eptr, *var_adj_es, *edge_es0, *edge_vs0, *fac_e0, *tables, *heavy;
  int32_t *heavy_cptr, *heavy_pofs, *heavy_chunk_var, *heavy_chunk_k, *edge_var, *perm,
      *var_adj_ref, *heavy_gpofs, *vh_group_h, *vh_group_j, *heavy_col, *heavy_row,
      *heavy_stride, *chunk_desc, *chunk_desc_ref, *col_desc, *row_desc, *heavy_desc;
  Block* blocks;
  Task* tasks;
  if ((rc = up(P.var_vs0, &var_vs0)) || (rc = up(P.var_eptr, &var_eptr)) ||
      (rc = up(P.var_adj_es, &var_adj_es)) || (rc = up(P.edge_es0, &edge_es0)) ||
      (rc = up(P.edge_vs0, &edge_vs0)) || (rc = up(P.fac_e0, &fac_e0)) ||
      (rc = up(P.tables, &tables)) || (rc = up(P.blocks, &blocks)) ||
      (rc = up(P.tasks, &tasks)) || (rc = up(P.heavy_vars, &heavy)) ||
      (rc = up(P.heavy_cptr, &heavy_cptr)) || (rc = up(P.heavy_pofs, &heavy_pofs)) ||
      (rc = up(P.heavy_chunk_var, &heavy_chunk_var)) || (rc = up(P.heavy_chunk_k, &heavy_chunk_k)) ||
      (rc = up(P.edge_var, &edge_var)) || (rc = up(P.perm, &perm)) ||
      (rc = up(P.var_adj_es_ref, &var_adj_ref)) || (rc = up(P.heavy_gpofs, &heavy_gpofs)) ||
      (rc = up(P.vh_group_h, &vh_group_h)) || (rc = up(P.vh_group_j, &vh_group_j)) ||
      (rc = up(P.heavy_col_chunks, &heavy_col)) || (rc = up(P.heavy_row_chunks, &heavy_row)) ||
      (rc = up(P.heavy_chunk_stride, &heavy_stride)) || (rc = up(P.chunk_desc, &chunk_desc)) ||
      (rc = up(P.chunk_desc_ref, &chunk_desc_ref)) || (rc = up(P.col_desc, &col_desc)) ||
      (rc = up(P.row_desc, &row_desc)) || (rc = up(P.heavy_desc, &heavy_desc))) {
    pgx_graph_destroy(g);
    return rc;
  }
  d.var_vs0 = var_vs0;
  d.var_eptr = var_eptr;
  d.var_adj_es = var_adj_es;
  d.perm = perm;
  d.var_adj_es_ref = var_adj_ref;
  d.edge_es0 = edge_es0;
  d.edge_vs0 = edge_vs0;
  d.fac_e0 = fac_e0;
  d.tables = tables;
  d.blocks = blocks;
  d.tasks = tasks;
  d.heavy_vars = heavy;
  d.num_heavy = int32_t(P.heavy_vars.size());
  d.heavy_cptr = heavy_cptr;
  d.heavy_pofs = heavy_pofs;
  d.heavy_chunk_var = heavy_chunk_var;
  d.heavy_chunk_k = heavy_chunk_k;
  d.heavy_gpofs = heavy_gpofs;
  d.vh_group_h = vh_group_h;
  d.vh_group_j = vh_group_j;
  d.num_vh_groups = int32_t(P.vh_group_h.size());
  d.heavy_chunk_stride = heavy_stride;
  d.chunk_desc = reinterpret_cast<const int4*>(chunk_desc);
  d.col_desc = reinterpret_cast<const int4*>(col_desc);
  d.row_desc = reinterpret_cast<const int4*>(row_desc);
  d.heavy_desc = reinterpret_cast<const int4*>(heavy_desc);
  g->chunk_desc_ref = reinterpret_cast<const int4*>(chunk_desc_ref);
  {
    std::vector<int32_t> zeros(P.heavy_chunk_stride.size(), 0);
    if ((rc = up(zeros, &g->zero_strides))) {
      pgx_graph_destroy(g);
      return rc;
    }
  }
  d.heavy_col = heavy_col;
  d.heavy_row = heavy_row;
  d.num_heavy_col = int32_t(P.heavy_col_chunks.size());
  d.num_heavy_row = int32_t(P.heavy_row_chunks.size());
  d.num_heavy_chunks = int32_t(P.heavy_chunk_var.size());
  d.partial_size = P.partial_size;
  d.edge_var = edge_var;
  if (!P.var_pos4.empty()) {
    int32_t* vp4;
    if ((rc = up(P.var_pos4, &vp4))) {
      pgx_graph_destroy(g);
      return rc;
    }
    d.var_pos4 = reinterpret_cast<const int4*>(vp4);
  }
  if (!P.edge_pos4.empty()) {
    int32_t* ep4;
    if ((rc = up(P.edge_pos4, &ep4))) {
      pgx_graph_destroy(g);
      return rc;
    }
    d.edge_pos4 = reinterpret_cast<const int4*>(ep4);
  }
  if (P.bip_na > 0) {
    int32_t *bip_row, *bip_col;
    if ((rc = up(P.bip_row, &bip_row)) || (rc = up(P.bip_col, &bip_col))) {
      pgx_graph_destroy(g);
      return rc;
    }
    d.bip_row = reinterpret_cast<const int2*>(bip_row);
    d.bip_col = reinterpret_cast<const int2*>(bip_col);
    d.bip_na = P.bip_na;
    d.bip_nb = P.bip_nb;
    d.bip_affine = P.bip_affine ? 1 : 0;
    for (int k = 0; k < 8; ++k) d.bip_aff[k] = P.bip_aff[k];
    g->bip_plan = 1;
  }
  d.fusable = (P.pairwise_only && P.max_degree <= kFuseMaxDeg) ? 1 : 0;
  g->packed_plan = P.packable ? 1 : 0;
  g->even2_plan = P.even2;
  g->fusable_plan = d.fusable;
  d.num_vars = P.num_vars;
  d.num_edges = P.num_edges;
  d.num_tasks = int32_t(P.tasks.size());
  d.num_blocks = int32_t(P.blocks.size());
  d.num_es = P.num_es;
  d.num_vs = P.num_vs;
  d.num_lp = P.num_lp;
  // factor tiles, grouped by type so pgx_fac_to_var can launch one type
  // kinds 0-2 are Enum, 3 OR, 4 AND, 5 Pool, so each factor type's tiles
  // are contiguous too
  // tsets: one factor per thread (wide logical / large general Enum blocks:
  // a lane group / wavefront per factor); msets: the same, except that 2x2
  // Enum blocks take kBin2PerThread factors per thread (k_fac_phase only)
  for (int ts = 0; ts < 6; ++ts) {
    const bool multi = ts >= 3;
    TileSet& T = multi ? g->msets[ts - 3] : g->tsets[ts];
    T.size = kTileSizes[ts % 3];
    std::vector<Tile> tiles;
    for (int k = 0; k < 6; ++k) {
      T.kind_tile0[k] = int32_t(tiles.size());
      for (int32_t bi = 0; bi < int32_t(P.blocks.size()); ++bi) {
        const Block& b = P.blocks[bi];
        if (b.kind != k) continue;
        g->kinds_mask |= 1 << k;
        // wide logical blocks: kWideLanes threads per factor; large general
        // Enum factors: one wavefront each
        const int32_t lanes = (b.pad[0] & 4) ? kWideLanes : ((b.pad[0] & 8) ? 64 : 1);
        const bool m = multi && b.kind == kEnumBin2;
        const int32_t per = m ? T.size * kBin2PerThread : T.size / lanes;
        for (int32_t i0 = 0; i0 < b.count; i0 += per)
          tiles.push_back(Tile{bi, i0, std::min(b.count, i0 + per),
                               m ? -kBin2PerThread : (lanes == 1 ? 0 : lanes)});
      }
    }
    T.kind_tile0[6] = int32_t(tiles.size());
    T.num_tiles = int32_t(tiles.size());
    if ((rc = up(tiles, &T.tiles))) {
      pgx_graph_destroy(g);
      return rc;
    }
  }
  static_assert(kTileSizes[2] == kTileFactors, "the plugin API uses the largest tiles");
  g->tiles = g->tsets[2].tiles;
  for (int k = 0; k < 7; ++k) g->kind_tile0[k] = g->tsets[2].kind_tile0[k];
  g->num_tiles = g->tsets[2].num_tiles;
  *out = g;
  return PGX_OK;
}

int pgx_graph_destroy(pgx_graph* g) {
  if (!g) return PGX_OK;
  for (auto& kv : g->graphs) (void)hipGraphExecDestroy(kv.second);
  for (void* p : g->owned) (void)hipFree(p);
  if (g->scratch) (void)hipFree(g->scratch);
  if (g->bscratch) (void)hipFree(g->bscratch);
  delete g;
  return PGX_OK;
}

static void plan_info(const Plan& P, int64_t* o) {
  o[0] = P.num_es;
  o[1] = P.num_vs;
  o[2] = P.num_lp;
  o[3] = P.num_edges;
  o[4] = P.num_vars;
  o[5] = P.num_factors;
  o[6] = int64_t(P.tasks.size());
  o[7] = int64_t(P.blocks.size());
}

int pgx_wiring_check(const pgx_wiring* wiring, int64_t* o) {
  if (!wiring) return set_error(PGX_EINVAL, "null argument");
  Plan P;
  std::string err;
  int64_t merge, planar;
  {
    std::lock_guard<std::mutex> lock(g_opt_mu);
    merge = g_opt_merge;
    planar = g_opt_planar;
  }
  const int rc = build_plan(*wiring, &P, &err, merge, planar != 0);
  if (rc != PGX_OK) return set_error(rc, "pgx_wiring_check: " + err);
  if (o) plan_info(P, o);
  return PGX_OK;
}

int pgx_graph_info(const pgx_graph* g, int64_t* o) {
  if (!g || !o) return set_error(PGX_EINVAL, "null argument");
  const Plan& P = g->plan;
  o[0] = P.num_es;
  o[1] = P.num_vs;
  o[2] = P.num_lp;
  o[3] = P.num_edges;
  o[4] = P.num_vars;
  o[5] = P.num_factors;
  o[6] = int64_t(P.tasks.size());
  o[7] = int64_t(P.blocks.size());
  return PGX_OK;
}

}  // extern "C"

static int ensure_beliefs_scratch(pgx_graph* g, size_t bytes) {
  if (g->bscratch_bytes >= bytes) return PGX_OK;
  if (g->bscratch) PGX_HIP(hipFree(g->bscratch));
  g->bscratch = nullptr;
  g->bscratch_bytes = 0;
  PGX_HIP(hipMalloc(reinterpret_cast<void**>(&g->bscratch), bytes));
  g->bscratch_bytes = bytes;
  return PGX_OK;
}

static int ensure_scratch(pgx_graph* g, size_t bytes) {
  if (g->scratch_bytes >= bytes) return PGX_OK;
  // captured runs point into the old scratch
  for (auto& kv : g->graphs) (void)hipGraphExecDestroy(kv.second);
  g->graphs.clear();
  g->graph_order.clear();
  if (g->scratch) PGX_HIP(hipFree(g->scratch));
  g->scratch = nullptr;
  g->scratch_bytes = 0;
  PGX_HIP(hipMalloc(reinterpret_cast<void**>(&g->scratch), bytes));
  g->scratch_bytes = bytes;
  return PGX_OK;
}

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
static int launch_task_loop(pgx_graph* g, const RunBufs& r, int32_t batch, int bs,
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

template <bool kT0>
static void launch_fac_kinds(pgx_graph* g, const TileSet& T, const RunBufs& r, int32_t batch,
                             int32_t it, hipStream_t s, const KernelClock& clk) {
  const DevGraph& d = g->dev;
  const int32_t* k0 = T.kind_tile0;
  auto grid = [&](int k) { return dim3(k0[k + 1] - k0[k], batch); };
  const dim3 blk(T.size);
  const int32_t id0 = PGX_KERNEL_FAC_PHASE;
  if (k0[1] > k0[0])
    clk.run(id0 + 0, [&] { hipLaunchKernelGGL((k_fac_phase<kT0, 0>), grid(0), blk, 0, s, d, T.tiles + k0[0], r, it); });
  if (k0[2] > k0[1])
    clk.run(id0 + 1, [&] { hipLaunchKernelGGL((k_fac_phase<kT0, 1>), grid(1), blk, 0, s, d, T.tiles + k0[1], r, it); });
  if (k0[3] > k0[2])
    clk.run(id0 + 2, [&] { hipLaunchKernelGGL((k_fac_phase<kT0, 2>), grid(2), blk, 0, s, d, T.tiles + k0[2], r, it); });
  if (k0[4] > k0[3])
    clk.run(id0 + 3, [&] { hipLaunchKernelGGL((k_fac_phase<kT0, 3>), grid(3), blk, 0, s, d, T.tiles + k0[3], r, it); });
  if (k0[5] > k0[4])
    clk.run(id0 + 4, [&] { hipLaunchKernelGGL((k_fac_phase<kT0, 4>), grid(4), blk, 0, s, d, T.tiles + k0[4], r, it); });
  if (k0[6] > k0[5])
    clk.run(id0 + 5, [&] { hipLaunchKernelGGL((k_fac_phase<kT0, 5>), grid(5), blk, 0, s, d, T.tiles + k0[5], r, it); });
}

// Buffers of a complete bipartite run: the engine-layout buffer the entry
// normalisation writes (full), and two one-float message buffers in the
// other engine-layout buffer (half its size each).
static void bip_buffers(const DevGraph& d, const RunBufs& r, int32_t batch, float** full,
                        float** m0, float** m1) {
  *full = (r.n_it & 1) ? r.y : r.x;
  float* other = (r.n_it & 1) ? r.x : r.y;
  *m0 = other;
  *m1 = other + size_t(batch) * size_t(d.num_es / 2);
}

template <bool kT0>
static int launch_phased(pgx_graph* g, const RunBufs& r, int32_t batch, hipStream_t s,
                         const KernelClock& clk, int64_t tile, int64_t threads) {
  const DevGraph& d = g->dev;
  const TileSet& T = g->tsets[tile <= 64 ? 0 : (tile <= 128 ? 1 : 2)];
  const dim3 fgrid(T.num_tiles, batch);
  if (d.bip) {
    // chunk sums of the initial messages, their one-float form, then one
    // launch per iteration alternating between two message buffers and two
    // chunk-sum buffers (run_impl unpacks the last message buffer)
    float *first, *m0, *m1;
    bip_buffers(d, r, batch, &first, &m0, &m1);
    launch_var_phase(d, first, d.num_es, r.ev, r.ev_stride, r.vsum, r.partial, batch, s, &clk);
    hipLaunchKernelGGL(k_pair_pack, dim3(unsigned(std::min<int64_t>((d.num_es / 2 + 255) / 256, 2048)), batch),
                       dim3(256), 0, s, d, first, m0, d.num_es / 2);
    float* p0 = r.partial;
    float* p1 = r.partial + size_t(batch) * size_t(d.partial_size);
    clk.run_n(PGX_KERNEL_BIPARTITE, r.n_it, [&] {
      for (int32_t it = 0; it < r.n_it; ++it)
        launch_bipartite<kT0>(d, r, (it & 1) ? m1 : m0, (it & 1) ? m0 : m1, (it & 1) ? p1 : p0,
                              (it & 1) ? p0 : p1, it, batch, s, tile, threads);
    });
    PGX_HIP(hipGetLastError());
    return PGX_OK;
  }
  if (d.fusable) {
    for (int32_t it = 0; it < r.n_it; ++it) {
      clk.run(PGX_KERNEL_FAC_FUSED, [&] {
        if (g->kinds_mask == kKindsBin2)
          hipLaunchKernelGGL((k_fac_fused<kT0, kKindsBin2>), fgrid, dim3(T.size), 0, s, d,
                             T.tiles, r, it);
        else
          hipLaunchKernelGGL((k_fac_fused<kT0, kKindsPairwise>), fgrid, dim3(T.size), 0, s,
                             d, T.tiles, r, it);
      });
    }
    PGX_HIP(hipGetLastError());
    return PGX_OK;
  }
  for (int32_t it = 0; it < r.n_it; ++it) {
    const float* mo = ((r.n_it - it) & 1) ? r.y : r.x;
    if (d.num_vars > 0)
      launch_var_phase(d, mo, d.num_es, r.ev, r.ev_stride, r.vsum, r.partial, batch, s, &clk);
    launch_fac_kinds<kT0>(g, g->msets[tile <= 64 ? 0 : (tile <= 128 ? 1 : 2)], r, batch, it, s,
                          clk);
  }
  PGX_HIP(hipGetLastError());
  return PGX_OK;
}

static inline bool aligned(const void* p, uintptr_t a) {
  return (reinterpret_cast<uintptr_t>(p) & (a - 1)) == 0;
}

// Delta slots per (batch element, iteration) of the per-launch kernels:
// 32 when a batch element spans >= 256 workgroups (64-factor tiles), fewer
// if the slots would pass 16 MiB, else 1 (no slots).
static int32_t delta_slots(const pgx_graph* g, int32_t batch, int32_t n_it) {
  if (g->tsets[0].num_tiles < 256) return 1;
  const int64_t pairs = int64_t(batch) * n_it;
  int32_t n = 32;
  while (n > 1 && pairs * n * 4 > (int64_t(16) << 20)) n >>= 1;
  return n;
}

// Scratch of one run: message buffer(s), variable sums, partial sums, delta slots.
static size_t run_scratch_bytes(const pgx_graph* g, int32_t batch, int32_t n_it) {
  const DevGraph& d = g->dev;
  const size_t msg_bytes = size_t(batch) * size_t(d.num_es) * sizeof(float);
  const size_t vs_bytes = size_t(batch) * size_t(d.num_vs) * sizeof(float);
  const size_t part_bytes = size_t(batch) * size_t(d.partial_size) * sizeof(float);
  const int32_t ns = delta_slots(g, batch, n_it);
  const size_t slot_bytes = ns > 1 ? size_t(batch) * size_t(n_it) * size_t(ns) * 4 : 0;
  // complete bipartite graphs alternate between two chunk-sum buffers
  return msg_bytes * (d.perm != nullptr ? 2 : 1) + vs_bytes +
         part_bytes * (d.bip_na > 0 ? 2 : 1) + slot_bytes + 512;
}

template <bool kT0, int KINDS>
static int launch_coop_t(pgx_graph* g, const TileSet& T, const RunBufs& r, int32_t batch,
                         hipStream_t s) {
  const void* fn = reinterpret_cast<const void*>(&k_coop_fused<kT0, KINDS>);
  int per_cu = 0, cus = 0, dev = 0;
  PGX_HIP(hipGetDevice(&dev));
  PGX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  PGX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, T.size, 0));
  const int64_t units = int64_t(T.num_tiles) * batch;
  const int64_t cap = int64_t(std::max(per_cu, 1)) * cus;
  const unsigned grid = unsigned(std::max<int64_t>(1, std::min(units, cap)));
  DevGraph d = g->dev;
  const Tile* tiles = T.tiles;
  int32_t nt = T.num_tiles;
  RunBufs rr = r;
  int32_t bt = batch;
  void* args[] = {&d, &tiles, &nt, &rr, &bt};
  PGX_HIP(hipLaunchCooperativeKernel(fn, dim3(grid), dim3(T.size), args, 0, s));
  return PGX_OK;
}

// Cooperative shape (fusable graphs only).
template <bool kT0>
static int launch_coop(pgx_graph* g, const RunBufs& r, int32_t batch, hipStream_t s,
                       int64_t tile) {
  const TileSet& T = g->tsets[tile <= 64 ? 0 : (tile <= 128 ? 1 : 2)];
  if (g->kinds_mask == kKindsBin2) return launch_coop_t<kT0, kKindsBin2>(g, T, r, batch, s);
  return launch_coop_t<kT0, kKindsPairwise>(g, T, r, batch, s);
}

// Shared body of pgx_run and pgx_run_converge. With task_deltas set, the
// run uses the task-loop shape and stops every task at its convergence.
static int run_impl(pgx_graph* g, const float* lp, int64_t lp_stride, const float* ev,
                    int64_t ev_stride, const float* ftov_in, int64_t ftov_in_stride,
                    float* ftov_out, float* deltas, float* task_deltas, int32_t* iters_done,
                    int32_t batch, int32_t num_iters, float damping, float temperature, float tol,
                    void* stream, const RunCfg* force = nullptr) {
  const bool converge = task_deltas != nullptr;
  if (!g || !ftov_out || batch <= 0 || (!converge && !deltas) || (converge && !iters_done))
    return set_error(PGX_EINVAL, "pgx_run: null argument or batch <= 0");
  const DevGraph& d = g->dev;
  if (d.num_es > 0 && (!ftov_in || (!lp && d.num_lp > 0)))
    return set_error(PGX_EINVAL, "pgx_run: null input");
  if (d.num_vs > 0 && !ev) return set_error(PGX_EINVAL, "pgx_run: null evidence");
  if (ftov_in == ftov_out) return set_error(PGX_EINVAL, "pgx_run: ftov_out aliases ftov_in");
  if (!(temperature >= 0.f)) return set_error(PGX_EINVAL, "pgx_run: temperature must be >= 0");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int32_t n_it = std::max(num_iters, 1);
  const size_t msg_bytes = size_t(batch) * size_t(d.num_es) * sizeof(float);
  const size_t vs_bytes = size_t(batch) * size_t(d.num_vs) * sizeof(float);
  // planar graphs iterate in two engine-layout buffers and unpermute at the end
  const bool planar = d.perm != nullptr;
  int rc = ensure_scratch(g, run_scratch_bytes(g, batch, n_it));
  if (rc) return rc;
  RunBufs r;
  r.lp = lp;
  r.lp_stride = lp_stride;
  r.ev = ev;
  r.ev_stride = ev_stride;
  r.y = g->scratch;
  r.x = planar ? g->scratch + size_t(batch) * size_t(d.num_es) : ftov_out;
  r.vsum = g->scratch + size_t(batch) * size_t(d.num_es) * (planar ? 2 : 1);
  r.partial = r.vsum + size_t(batch) * size_t(d.num_vs);
  r.deltas = reinterpret_cast<uint32_t*>(deltas);
  r.dslot_n = converge ? 1 : delta_slots(g, batch, n_it);
  r.dslot = r.dslot_n > 1 ? reinterpret_cast<uint32_t*>(
                                r.partial + size_t(batch) * size_t(d.partial_size) *
                                                (d.bip_na > 0 ? 2 : 1))
                          : r.deltas;
  r.n_it = n_it;
  r.d = damping;
  r.omd = float(1.0 - double(damping));
  r.t = temperature;
  r.converge = converge ? 1 : 0;
  r.tol = tol;
  r.task_deltas = task_deltas;
  r.iters_done = iters_done;

  if (converge) {
    const int64_t nt = int64_t(batch) * d.num_tasks * n_it;
    if (nt > 0)
      hipLaunchKernelGGL(k_fill_u32, dim3(std::min<int64_t>((nt + 255) / 256, 4096)), dim3(256),
                         0, s, reinterpret_cast<uint32_t*>(task_deltas), nt, 0u);
  } else {
    const int64_t nd = int64_t(batch) * n_it;
    hipLaunchKernelGGL(k_fill_u32, dim3(std::min<int64_t>((nd + 255) / 256, 1024)), dim3(256), 0,
                       s, r.deltas, nd, 0u);
    if (r.dslot_n > 1)
      hipLaunchKernelGGL(k_fill_u32, dim3(std::min<int64_t>((nd * r.dslot_n + 255) / 256, 4096)),
                         dim3(256), 0, s, r.dslot, nd * r.dslot_n, 0u);
  }
  // initial messages land in the buffer iteration 0 reads
  float* first = (n_it & 1) ? r.y : r.x;
  if (d.num_edges > 0) {
    const dim3 grid(std::min<int32_t>((d.num_edges + 255) / 256, 4096), batch);
    hipLaunchKernelGGL(k_normalize_entry, grid, dim3(256), 0, s, d, ftov_in, ftov_in_stride,
                       first);
  }
  PGX_HIP(hipGetLastError());
  if (d.num_tasks == 0 || d.num_es == 0) {
    // no factors: every iteration maps the (empty) messages onto themselves
    if (first != r.x && d.num_es > 0)
      PGX_HIP(hipMemcpyAsync(r.x, first, msg_bytes, hipMemcpyDeviceToDevice, s));
    return PGX_OK;
  }
  int64_t mode, bs, prof, fuse, vec, tile, pack;
  {
    std::lock_guard<std::mutex> lock(g_opt_mu);
    pack = g_opt_pack;
    mode = g_opt_mode;
    bs = g_opt_block;
    prof = g_opt_profile;
    fuse = g_opt_fuse;
    vec = g_opt_vec;
    tile = g_opt_tile;
  }
  bool chosen = force != nullptr;  // shape set by the autotuner
  if (force) {
    mode = force->mode;
    bs = force->bs;
    fuse = force->fuse;
    vec = force->vec;
    tile = force->tile;
  } else if (mode == 0 && bs == 0 && tile == 0) {
    auto it = g->tuned.find(std::make_tuple(batch, temperature == 0.f ? 1 : 0, converge ? 1 : 0));
    if (it != g->tuned.end()) {
      mode = it->second.mode;
      bs = it->second.bs;
      fuse = it->second.fuse;
      vec = it->second.vec;
      tile = it->second.tile;
      chosen = true;
    }
  }
  if (tile == 0) tile = 256;
  const bool small_run = int64_t(batch) * d.num_es <= kFusedMaxStates;
  // k_bipartite numbers (tile, batch element) pairs in one grid dimension
  const int64_t bip_units = int64_t((d.bip_na + kBipRows - 1) / kBipRows) *
                            ((d.bip_nb + kVarChunk - 1) / kVarChunk) * batch;
  g->dev.bip = (g->bip_plan && fuse && bip_units <= int64_t(INT32_MAX)) ? 1 : 0;
  g->dev.even2 = ((vec & 1) && g->even2_plan && (d.num_es & 1) == 0 && (d.num_vs & 1) == 0 &&
                  (ev_stride & 1) == 0 && aligned(ev, 8) && aligned(ftov_out, 8))
                     ? 1
                     : 0;
  g->dev.lp_al4 = ((vec & 2) && lp_stride % 4 == 0 && aligned(lp, 16)) ? ((vec & 4) ? 2 : 1) : 0;
  if (converge) mode = 1;
  if (mode == 0) {
    // One workgroup per task pays when there are enough tasks to fill the
    // CUs and each is small enough that a CU's sweep over it is shorter than
    // a launch; otherwise every phase is spread over the whole GPU.
    const int64_t wgs = int64_t(d.num_tasks) * batch;
    const int64_t per_task = (int64_t(d.num_es) + d.num_vs) / std::max<int64_t>(d.num_tasks, 1);
    mode = (wgs >= 256 && per_task <= 4096) ? 1 : 2;
  }
  if (bs == 0) bs = 256;
  // Untuned, the per-phase shape recomputes variable sums on the factor side
  // (one launch per iteration, every variable's messages re-read per factor)
  // only for small runs, where the launch count dominates; past
  // kFusedMaxStates edge states its two-kernel iteration is faster.
  g->dev.fusable =
      (g->fusable_plan && fuse && (chosen || small_run || mode != 2)) ? 1 : 0;
  const bool t0 = (temperature == 0.f);
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  if (prof) {
    PGX_HIP(hipEventCreate(&ev0));
    PGX_HIP(hipEventCreate(&ev1));
    PGX_HIP(hipEventRecord(ev0, s));
  }
  KernelClock clk;
  clk.on = prof == 2;
  clk.s = s;
  if (mode == 3 && !g->dev.fusable) mode = 2;  // the cooperative shape needs a fusable graph
  if (mode == 1 && int64_t(d.num_tasks) * batch > int64_t(INT32_MAX)) {
    if (converge)
      return set_error(PGX_EUNSUP, "pgx_run_converge: tasks x batch exceeds one launch");
    mode = 2;
  }
  // graphs of 2x2 factors and variables of <= 4 edges: the per-phase shape
  // iterates on one float per message. The entry normalisation wrote full
  // pairs into `first`; pack them into the other buffer and swap the
  // buffers' roles, so iteration 0 reads the packed copy and the last
  // iteration writes r.x (unpacked into ftov_out below).
  g->dev.packed = (mode == 2 && !g->dev.bip && g->packed_plan && g->dev.even2 && pack) ? 1 : 0;
  if (g->dev.packed) {
    float* other = (first == r.x) ? r.y : r.x;
    hipLaunchKernelGGL(k_pair_pack,
                       dim3(unsigned(std::min<int64_t>((d.num_es / 2 + 255) / 256, 2048)), batch),
                       dim3(256), 0, s, g->dev, static_cast<const float*>(first), other,
                       d.num_es);
    std::swap(r.x, r.y);
  }
  if (mode == 1)
    clk.run(PGX_KERNEL_TASK_LOOP, [&] {
      rc = t0 ? launch_task_loop<true>(g, r, batch, int(bs), s)
              : launch_task_loop<false>(g, r, batch, int(bs), s);
    });
  else if (mode == 3)
    clk.run(PGX_KERNEL_FAC_FUSED, [&] {
      rc = t0 ? launch_coop<true>(g, r, batch, s, tile) : launch_coop<false>(g, r, batch, s, tile);
    });
  else {
    rc = t0 ? launch_phased<true>(g, r, batch, s, clk, tile, bs)
            : launch_phased<false>(g, r, batch, s, clk, tile, bs);
    if (rc == PGX_OK && r.dslot_n > 1) {  // the per-launch kernels wrote delta slots
      const int64_t nd = int64_t(batch) * n_it;
      hipLaunchKernelGGL(k_delta_reduce, dim3(std::min<int64_t>((nd + 255) / 256, 1024)),
                         dim3(256), 0, s, r.dslot, r.dslot_n, r.deltas, nd);
      PGX_HIP(hipGetLastError());
    }
  }
  if (rc == PGX_OK && planar) {
    const dim3 ugrid(unsigned(std::min<int64_t>((d.num_es + 255) / 256, 2048)), batch);
    if (mode == 2 && g->dev.bip) {  // the last one-float message buffer
      float *full, *m0, *m1;
      bip_buffers(g->dev, r, batch, &full, &m0, &m1);
      hipLaunchKernelGGL(k_pair_unpack, ugrid, dim3(256), 0, s, g->dev,
                         static_cast<const float*>((n_it & 1) ? m1 : m0), g->dev.num_es / 2,
                         ftov_out);
    } else if (g->dev.packed) {  // one float per message at the start of each batch slot
      hipLaunchKernelGGL(k_pair_unpack, ugrid, dim3(256), 0, s, g->dev,
                         static_cast<const float*>(r.x), g->dev.num_es, ftov_out);
    } else {
      hipLaunchKernelGGL(k_unpermute, ugrid, dim3(256), 0, s, d, static_cast<const float*>(r.x),
                         ftov_out);
    }
    PGX_HIP(hipGetLastError());
  }
  if (prof) {
    PGX_HIP(hipEventRecord(ev1, s));
    std::lock_guard<std::mutex> lock(g_opt_mu);
    g_prof_events.emplace_back(ev0, ev1);
  }
  return rc;
}

namespace pgx {
const Plan& graph_plan(const pgx_graph* g) { return g->plan; }
}  // namespace pgx

extern "C" {

int pgx_run(pgx_graph* g, const float* lp, int64_t lp_stride, const float* ev,
            int64_t ev_stride, const float* ftov_in, int64_t ftov_in_stride, float* ftov_out,
            float* deltas, int32_t batch, int32_t num_iters, float damping, float temperature,
            void* stream) {
  if (g && batch > kMaxBatchPerLaunch) {  // the batch is a grid dimension: run it in parts
    const int64_t n_it = std::max(num_iters, 1);
    for (int32_t b0 = 0; b0 < batch; b0 += kMaxBatchPerLaunch) {
      const int32_t nb = std::min(kMaxBatchPerLaunch, batch - b0);
      const int rc = pgx_run(g, lp ? lp + b0 * lp_stride : lp, lp_stride,
                             ev ? ev + b0 * ev_stride : ev, ev_stride,
                             ftov_in ? ftov_in + b0 * ftov_in_stride : ftov_in, ftov_in_stride,
                             ftov_out + int64_t(b0) * g->dev.num_es,
                             deltas ? deltas + b0 * n_it : deltas, nb, num_iters, damping,
                             temperature, stream);
      if (rc) return rc;
    }
    return PGX_OK;
  }
  int64_t use_graph, prof;
  std::vector<int64_t> key;
  {
    std::lock_guard<std::mutex> lock(g_opt_mu);
    use_graph = g_opt_graph;
    prof = g_opt_profile;
    key = {g_opt_mode, g_opt_block, g_opt_fuse, g_opt_vec, g_opt_tile, g_opt_pack};
  }
  if (!use_graph || prof || !g || batch <= 0)
    return run_impl(g, lp, lp_stride, ev, ev_stride, ftov_in, ftov_in_stride, ftov_out, deltas,
                    nullptr, nullptr, batch, num_iters, damping, temperature, 0.f, stream);
  // Replay a captured run when every launch argument matches (scratch included).
  int rc = ensure_scratch(g, run_scratch_bytes(g, batch, std::max(num_iters, 1)));
  if (rc) return rc;
  uint32_t dbits, tbits;
  std::memcpy(&dbits, &damping, 4);
  std::memcpy(&tbits, &temperature, 4);
  const int64_t ptrs[] = {int64_t(reinterpret_cast<uintptr_t>(lp)), lp_stride,
                          int64_t(reinterpret_cast<uintptr_t>(ev)), ev_stride,
                          int64_t(reinterpret_cast<uintptr_t>(ftov_in)), ftov_in_stride,
                          int64_t(reinterpret_cast<uintptr_t>(ftov_out)),
                          int64_t(reinterpret_cast<uintptr_t>(deltas)),
                          int64_t(reinterpret_cast<uintptr_t>(g->scratch)),
                          int64_t(reinterpret_cast<uintptr_t>(stream)), batch, num_iters,
                          int64_t(dbits), int64_t(tbits)};
  key.insert(key.end(), std::begin(ptrs), std::end(ptrs));
  hipStream_t s = static_cast<hipStream_t>(stream);
  auto it = g->graphs.find(key);
  if (it == g->graphs.end()) {
    // the legacy default stream cannot be captured, and a stream the caller
    // is capturing already records the launches itself
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (s == nullptr || hipStreamIsCapturing(s, &cs) != hipSuccess ||
        cs != hipStreamCaptureStatusNone) {
      (void)hipGetLastError();
      return run_impl(g, lp, lp_stride, ev, ev_stride, ftov_in, ftov_in_stride, ftov_out, deltas,
                      nullptr, nullptr, batch, num_iters, damping, temperature, 0.f, stream);
    }
    PGX_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    rc = run_impl(g, lp, lp_stride, ev, ev_stride, ftov_in, ftov_in_stride, ftov_out, deltas,
                  nullptr, nullptr, batch, num_iters, damping, temperature, 0.f, stream);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(s, &graph);
    hipGraphExec_t exec = nullptr;
    const bool ok = rc == PGX_OK && e == hipSuccess && graph != nullptr &&
                    hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess;
    if (graph) (void)hipGraphDestroy(graph);
    if (!ok) {
      // a run that cannot be captured (e.g. the cooperative shape) runs as is
      (void)hipGetLastError();
      return run_impl(g, lp, lp_stride, ev, ev_stride, ftov_in, ftov_in_stride, ftov_out,
                      deltas, nullptr, nullptr, batch, num_iters, damping, temperature, 0.f,
                      stream);
    }
    constexpr size_t kMaxGraphs = 8;
    if (g->graph_order.size() >= kMaxGraphs) {
      auto old = g->graphs.find(g->graph_order.front());
      if (old != g->graphs.end()) {
        (void)hipGraphExecDestroy(old->second);
        g->graphs.erase(old);
      }
      g->graph_order.erase(g->graph_order.begin());
    }
    it = g->graphs.emplace(key, exec).first;
    g->graph_order.push_back(key);
  }
  PGX_HIP(hipGraphLaunch(it->second, s));
  return PGX_OK;
}

int pgx_run_converge(pgx_graph* g, const float* lp, int64_t lp_stride, const float* ev,
                     int64_t ev_stride, const float* ftov_in, int64_t ftov_in_stride,
                     float* ftov_out, float* task_deltas, int32_t* iters_done, int32_t batch,
                     int32_t max_iters, float damping, float temperature, float tol,
                     void* stream) {
  if (!task_deltas) return set_error(PGX_EINVAL, "pgx_run_converge: task_deltas is null");
  if (g && ftov_out && iters_done && batch > kMaxBatchPerLaunch) {
    const int64_t n_it = std::max(max_iters, 1), nt = g->dev.num_tasks;
    for (int32_t b0 = 0; b0 < batch; b0 += kMaxBatchPerLaunch) {
      const int32_t nb = std::min(kMaxBatchPerLaunch, batch - b0);
      const int rc = pgx_run_converge(
          g, lp ? lp + b0 * lp_stride : lp, lp_stride, ev ? ev + b0 * ev_stride : ev, ev_stride,
          ftov_in ? ftov_in + b0 * ftov_in_stride : ftov_in, ftov_in_stride,
          ftov_out + int64_t(b0) * g->dev.num_es, task_deltas + b0 * nt * n_it,
          iters_done + b0 * nt, nb, max_iters, damping, temperature, tol, stream);
      if (rc) return rc;
    }
    return PGX_OK;
  }
  return run_impl(g, lp, lp_stride, ev, ev_stride, ftov_in, ftov_in_stride, ftov_out, nullptr,
                  task_deltas, iters_done, batch, max_iters, damping, temperature, tol, stream);
}

int pgx_autotune(pgx_graph* g, const float* lp, int64_t lp_stride, const float* ev,
                 int64_t ev_stride, const float* ftov_in, int64_t ftov_in_stride, int32_t batch,
                 float damping, float temperature, int32_t converge, int32_t probe_iters,
                 void* stream, int64_t* chosen) {
  if (!g || batch <= 0 || batch > kMaxBatchPerLaunch)
    return set_error(PGX_EINVAL, "pgx_autotune: batch must be in [1, 65535]");
  const DevGraph& d = g->dev;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int32_t n_it = std::max(probe_iters, 2);
  std::vector<RunCfg> cands;
  const bool can_fuse = g->fusable_plan || g->bip_plan;
  const int64_t fuse_opts[2] = {can_fuse ? 1 : 0, 0};
  const int nf = can_fuse ? 2 : 1;
  for (int fi = 0; fi < nf; ++fi)
    for (int64_t b : {256, 512, 1024}) cands.push_back(RunCfg{1, b, fuse_opts[fi], 3, 256});
  if (!converge && g->bip_plan)  // k_bipartite tiles: (columns, threads)
    for (const auto& ct : {std::make_pair(64, 256), std::make_pair(64, 512),
                           std::make_pair(32, 256), std::make_pair(128, 512)})
      cands.push_back(RunCfg{2, ct.second, 1, 3, ct.first});
  if (!converge)
    for (int fi = 0; fi < nf; ++fi) {
      if (g->bip_plan && fuse_opts[fi]) continue;  // covered above
      for (int64_t t : {64, 128, 256}) cands.push_back(RunCfg{2, 0, fuse_opts[fi], 3, t});
      cands.push_back(RunCfg{2, 0, fuse_opts[fi], 0, 256});
    }
  if (!converge && g->fusable_plan)
    for (int64_t t : {64, 256}) cands.push_back(RunCfg{3, 0, 1, 3, t});
  float *out = nullptr, *dl = nullptr;
  int32_t* it_done = nullptr;
  const size_t out_bytes = size_t(batch) * size_t(std::max<int64_t>(d.num_es, 1)) * 4;
  const size_t dl_bytes = size_t(batch) * size_t(std::max(d.num_tasks, 1)) * n_it * 4;
  PGX_HIP(hipMalloc(reinterpret_cast<void**>(&out), out_bytes));
  if (hipMalloc(reinterpret_cast<void**>(&dl), dl_bytes) != hipSuccess ||
      hipMalloc(reinterpret_cast<void**>(&it_done), size_t(batch) * std::max(d.num_tasks, 1) * 4) !=
          hipSuccess) {
    (void)hipFree(out);
    if (dl) (void)hipFree(dl);
    return set_error(PGX_ENOMEM, "pgx_autotune: cannot allocate probe buffers");
  }
  hipEvent_t e0, e1;
  int rc = PGX_OK;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    rc = set_error(PGX_EHIP, "pgx_autotune: hipEventCreate failed");
  RunCfg best;
  float best_ms = 3.4e38f;
  for (const RunCfg& c : cands) {
    if (rc) break;
    float ms = 3.4e38f;
    for (int rep = 0; rep < 3 && !rc; ++rep) {  // first repetition warms up
      (void)hipEventRecord(e0, s);
      rc = converge ? run_impl(g, lp, lp_stride, ev, ev_stride, ftov_in, ftov_in_stride, out,
                               nullptr, dl, it_done, batch, n_it, damping, temperature, 0.f,
                               stream, &c)
                    : run_impl(g, lp, lp_stride, ev, ev_stride, ftov_in, ftov_in_stride, out, dl,
                               nullptr, nullptr, batch, n_it, damping, temperature, 0.f, stream,
                               &c);
      (void)hipEventRecord(e1, s);
      if (rc) break;
      if (hipEventSynchronize(e1) != hipSuccess) {
        rc = set_error(PGX_EHIP, "pgx_autotune: probe run failed");
        break;
      }
      float t = 0.f;
      (void)hipEventElapsedTime(&t, e0, e1);
      if (rep > 0) ms = std::min(ms, t);
    }
    if (!rc && ms < best_ms) {
      best_ms = ms;
      best = c;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(out);
  (void)hipFree(dl);
  (void)hipFree(it_done);
  if (rc) return rc;
  g->tuned[std::make_tuple(batch, temperature == 0.f ? 1 : 0, converge ? 1 : 0)] = best;
  if (chosen) {
    chosen[0] = best.mode;
    chosen[1] = best.bs;
    chosen[2] = best.fuse;
    chosen[3] = best.vec;
    chosen[4] = best.tile;
  }
  return PGX_OK;
}

int pgx_graph_tasks(const pgx_graph* g, int32_t* out) {
  if (!g || !out) return set_error(PGX_EINVAL, "pgx_graph_tasks: null argument");
  for (size_t i = 0; i < g->plan.tasks.size(); ++i) {
    const Task& t = g->plan.tasks[i];
    out[4 * i] = t.v0;
    out[4 * i + 1] = t.v1;
    out[4 * i + 2] = t.b0;
    out[4 * i + 3] = t.b1;
  }
  return PGX_OK;
}

int pgx_get_beliefs(pgx_graph* g, const float* ev, int64_t ev_stride, const float* ftov,
                    int64_t ftov_stride, float* beliefs, int32_t batch, void* stream) {
  if (!g || !beliefs || batch <= 0) return set_error(PGX_EINVAL, "pgx_get_beliefs: bad argument");
  if (batch > kMaxBatchPerLaunch) {
    for (int32_t b0 = 0; b0 < batch; b0 += kMaxBatchPerLaunch) {
      const int32_t nb = std::min(kMaxBatchPerLaunch, batch - b0);
      const int rc = pgx_get_beliefs(g, ev ? ev + b0 * ev_stride : ev, ev_stride,
                                     ftov ? ftov + b0 * ftov_stride : ftov, ftov_stride,
                                     beliefs + int64_t(b0) * g->dev.num_vs, nb, stream);
      if (rc) return rc;
    }
    return PGX_OK;
  }
  const DevGraph& d = g->dev;
  if (d.num_vars == 0) return PGX_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const size_t part_bytes = size_t(batch) * size_t(d.partial_size) * 4;
  if (part_bytes > 0) {
    const int rc = ensure_beliefs_scratch(g, part_bytes);
    if (rc) return rc;
  }
  int64_t vec;
  {
    std::lock_guard<std::mutex> lock(g_opt_mu);
    vec = g_opt_vec;
  }
  g->dev.even2 = ((vec & 1) && g->even2_plan && (d.num_vs & 1) == 0 && (ev_stride & 1) == 0 &&
                  (ftov_stride & 1) == 0 && aligned(ev, 8) && aligned(ftov, 8) &&
                  aligned(beliefs, 8))
                     ? 1
                     : 0;
  // the caller's messages are in the reference layout
  DevGraph dref = g->dev;
  dref.var_adj_es = dref.var_adj_es_ref;
  // the row split and chunk strides hold for engine positions only: every
  // chunk in order, positions loaded
  dref.heavy_col = nullptr;
  dref.col_desc = nullptr;
  dref.chunk_desc = g->chunk_desc_ref;
  dref.heavy_chunk_stride = g->zero_strides;
  dref.num_heavy_col = dref.num_heavy_chunks;
  dref.num_heavy_row = 0;
  dref.var_pos4 = nullptr;  // engine positions
  dref.edge_pos4 = nullptr;
  launch_var_phase(dref, ftov, ftov_stride, ev, ev_stride, beliefs, g->bscratch, batch, s);
  PGX_HIP(hipGetLastError());
  return PGX_OK;
}

int pgx_decode_map(pgx_graph* g, const float* beliefs, int32_t* map_states, int32_t batch,
                   void* stream) {
  if (!g || !beliefs || !map_states || batch <= 0)
    return set_error(PGX_EINVAL, "pgx_decode_map: bad argument");
  if (batch > kMaxBatchPerLaunch) {
    for (int32_t b0 = 0; b0 < batch; b0 += kMaxBatchPerLaunch) {
      const int32_t nb = std::min(kMaxBatchPerLaunch, batch - b0);
      const int rc = pgx_decode_map(g, beliefs + int64_t(b0) * g->dev.num_vs,
                                    map_states + int64_t(b0) * g->dev.num_vars, nb, stream);
      if (rc) return rc;
    }
    return PGX_OK;
  }
  const DevGraph& d = g->dev;
  if (d.num_vars == 0) return PGX_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const dim3 grid((d.num_vars + 255) / 256, batch);
  hipLaunchKernelGGL(k_decode, grid, dim3(256), 0, s, d, beliefs, map_states);
  PGX_HIP(hipGetLastError());
  return PGX_OK;
}

int pgx_fac_to_var(pgx_graph* g, int32_t factor_type, const float* vtof, const float* lp,
                   float* ftov, float temperature, int32_t normalize, void* stream) {
  if (!g || !vtof || !ftov || factor_type < 0 || factor_type > 3)
    return set_error(PGX_EINVAL, "pgx_fac_to_var: bad argument");
  if (factor_type == PGX_ENUM && !normalize)
    return set_error(PGX_EUNSUP, "pgx_fac_to_var: Enum factors have no unnormalised form here");
  if (!(temperature >= 0.f)) return set_error(PGX_EINVAL, "pgx_fac_to_var: temperature < 0");
  const int k_lo = factor_type == PGX_ENUM ? 0 : factor_type + 2;
  const int k_hi = factor_type == PGX_ENUM ? 3 : factor_type + 3;
  if (g->kind_tile0[k_hi] == g->kind_tile0[k_lo]) return PGX_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int64_t m0 = g->plan.msgs_range[factor_type][0], m1 = g->plan.msgs_range[factor_type][1];
  if (m1 > m0) PGX_HIP(hipMemsetAsync(ftov + m0, 0, size_t(m1 - m0) * sizeof(float), s));
  for (int k = k_lo; k < k_hi; ++k) {
    const int32_t a = g->kind_tile0[k], z = g->kind_tile0[k + 1];
    if (z == a) continue;
    const dim3 grid(z - a);
    const Tile* tl = g->tiles + a;
    const bool t0 = temperature == 0.f;
#define PGX_OP_CASE(K)                                                                      \
  case K:                                                                                   \
    if (t0)                                                                                 \
      hipLaunchKernelGGL((k_fac_op<true, K>), grid, dim3(kPhaseThreads), 0, s, g->dev, tl,  \
                         vtof, lp, ftov, temperature, normalize);                           \
    else                                                                                    \
      hipLaunchKernelGGL((k_fac_op<false, K>), grid, dim3(kPhaseThreads), 0, s, g->dev, tl, \
                         vtof, lp, ftov, temperature, normalize);                           \
    break;
    switch (k) {
      PGX_OP_CASE(0)
      PGX_OP_CASE(1)
      PGX_OP_CASE(2)
      PGX_OP_CASE(3)
      PGX_OP_CASE(4)
      PGX_OP_CASE(5)
    }
#undef PGX_OP_CASE
  }
  PGX_HIP(hipGetLastError());
  return PGX_OK;
}

int pgx_enum_rows_fac_to_var(const int32_t* cfg_row_ptr, const int32_t* row_es,
                             const int32_t* es_row_ptr, const int32_t* es_row_cfg,
                             int64_t num_configs, int64_t num_es, const float* vtof,
                             const float* lp, float* summary, float* ftov, float temperature,
                             void* stream) {
  if (num_configs < 0 || num_es < 0 || !(temperature >= 0.f))
    return set_error(PGX_EINVAL, "pgx_enum_rows_fac_to_var: bad sizes or temperature");
  if (num_es == 0) return PGX_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (num_configs > 0)
    hipLaunchKernelGGL(k_enum_rows_summary, dim3(std::min<int64_t>((num_configs + 255) / 256, 4096)),
                       dim3(256), 0, s, cfg_row_ptr, row_es, num_configs, vtof, lp, summary);
  const dim3 grid(std::min<int64_t>((num_es + 255) / 256, 4096));
  if (temperature == 0.f)
    hipLaunchKernelGGL((k_enum_rows_out<true>), grid, dim3(256), 0, s, es_row_ptr, es_row_cfg,
                       num_es, vtof, summary, ftov, temperature);
  else
    hipLaunchKernelGGL((k_enum_rows_out<false>), grid, dim3(256), 0, s, es_row_ptr, es_row_cfg,
                       num_es, vtof, summary, ftov, temperature);
  PGX_HIP(hipGetLastError());
  return PGX_OK;
}

}  // extern "C"
