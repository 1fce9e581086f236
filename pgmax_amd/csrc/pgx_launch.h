// Host-side declarations shared by the engine's translation units: engine
// options, per-kernel profiling clocks, the buffers of one run, factor tiles,
// the compiled graph, and the launchers each kernel file provides. The
// kernels are split over several files so they compile in parallel; each
// kernel is launched only from host code in its own file.
#ifndef PGX_LAUNCH_H_
#define PGX_LAUNCH_H_

#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <utility>
#include <vector>

#include "../../include/pgx.h"
#include "pgx_device.h"
#include "pgx_internal.h"
#include "pgx_plan.h"

namespace pgx {

// Engine options (pgx_set_option), guarded by g_opt_mu.
extern int64_t g_opt_mode, g_opt_block, g_opt_profile, g_opt_fuse, g_opt_vec, g_opt_merge,
    g_opt_planar, g_opt_tile, g_opt_graph, g_opt_pack, g_opt_lanes, g_opt_xcd, g_opt_share,
    g_opt_mpnn_rows;
extern std::mutex g_opt_mu;
extern std::vector<std::pair<hipEvent_t, hipEvent_t>> g_prof_events;
// per-kernel records (option "profile" = 2): kernel id, start, end
struct KernelRecord {
  int32_t id;
  hipEvent_t a, b;
  int32_t launches;  // back-to-back launches between a and b
};
extern std::vector<KernelRecord> g_kernel_events;

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

// Block coordinates renumbered so that each XCD runs runs of `chunk`
// consecutive workgroups: XCD x takes workgroups [r * kXcds * chunk + x *
// chunk, + chunk) of each round r. Consecutive factor tiles share variables
// (and consecutive variable blocks share factors), so those shared lines are
// fetched into one XCD's L2, while all XCDs still advance through the grid
// together (the imbalance of whole contiguous ranges per XCD does not
// arise). The last, partial round keeps the linear order. chunk <= 1:
// unchanged.
__device__ __forceinline__ void xcd_chunked(uint32_t chunk, uint32_t& bx, uint32_t& by) {
  bx = blockIdx.x;
  by = blockIdx.y;
  if (chunk <= 1) return;
  const uint64_t nx = gridDim.x, n = nx * gridDim.y;
  const uint64_t p = uint64_t(blockIdx.y) * nx + blockIdx.x;
  const uint64_t round = uint64_t(kXcds) * chunk;
  const uint64_t full = n / round * round;
  uint64_t q = p;
  if (p < full) {
    const uint64_t x = p % kXcds, j = p / kXcds;
    q = (j / chunk) * round + x * chunk + (j % chunk);
  }
  bx = uint32_t(q % nx);
  by = uint32_t(q / nx);
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
  int32_t xcd = 0;  // per-phase kernels: XCD run length in workgroups (xcd_chunked)
  // Lane-group tiles of square pairwise factors whose factors all hold the
  // tile's first factor's log-potential table bit for bit (k_tile_share):
  // share[b * share_bstride + t] != 0 for lane tile t of the run's tile set
  // and batch element b (share_bstride 0: one log-potential array for every
  // element); null: no tile is read that way.
  int32_t* share = nullptr;
  int64_t share_bstride = 0;
  int32_t* share_mem = nullptr;  // scratch for the flags
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

struct Tile {
  int32_t block, i0, i1, pad;
};

// Execution shape of one run: mode 1 task loop / 2 phased, workgroup size of
// the task loop, factor-side variable sums, vector accesses.
struct RunCfg {
  int64_t mode = 0, bs = 0, fuse = 1, vec = 3, tile = 256, lanes = 1, xcd = 0;
};

// Tile groups: the six block kinds, then (per-phase tile sets only) square
// pairwise Enum blocks of 3..8 states, once as lane-group tiles
// (k_pair_lanes) and once one factor per thread (k_fac_phase); a run
// launches one of the two.
constexpr int kTileGroups = 8;
constexpr int kPairLanesGroup = 6;
constexpr int kPairThreadGroup = 7;

// Factor tiles of one size, grouped by tile group.
struct TileSet {
  Tile* tiles = nullptr;
  int32_t kind_tile0[kTileGroups + 1] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  int32_t num_tiles = 0;
  int32_t size = 256;
};
constexpr int kTileSizes[3] = {64, 128, 256};
constexpr int kLaneFusedPasses[3] = {4, 8, 16};

constexpr int kBipRows = kVarChunk;  // rows per tile of k_bipartite: one column chunk

// Buffers of a complete bipartite run: the engine-layout buffer the entry
// normalisation writes (full), and two one-float message buffers in the
// other engine-layout buffer (half its size each).
inline void bip_buffers(const DevGraph& d, const RunBufs& r, int32_t batch, float** full,
                        float** m0, float** m1) {
  *full = (r.n_it & 1) ? r.y : r.x;
  float* other = (r.n_it & 1) ? r.x : r.y;
  *m0 = other;
  *m1 = other + size_t(batch) * size_t(d.num_es / 2);
}

}  // namespace pgx

struct pgx_graph {
  pgx::Plan plan;
  pgx::TileSet tsets[3];  // tiles of 64, 128 and 256 factors (threads per workgroup)
  pgx::TileSet msets[3];  // the same with kBin2PerThread 2x2 factors per thread
  // lane-group tiles of every block, for graphs of 2 x 2 and square pairwise
  // blocks whose variables have at most 4 edges (k_pair_lanes_fused; empty
  // otherwise), with kLaneFusedPasses passes per wavefront (tile option 64,
  // 128, 256 -> 4, 8, 16)
  pgx::TileSet lsets[3];
  // shapes chosen by pgx_autotune, keyed by (batch, temperature == 0, converge)
  std::map<std::tuple<int32_t, int, int>, pgx::RunCfg> tuned;
  // runs captured as HIP graphs (option "graph"), keyed by every launch argument
  std::map<std::vector<int64_t>, hipGraphExec_t> graphs;
  std::vector<std::vector<int64_t>> graph_order;  // insertion order, for eviction
  pgx::DevGraph dev{};
  std::vector<void*> owned;
  pgx::Tile* tiles = nullptr;  // all factor tiles, grouped by factor type
  int32_t kind_tile0[pgx::kTileGroups + 1] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // by tile group
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

namespace pgx {

// pgx_phase.hip: the per-phase (one launch per phase and iteration) shapes.
void launch_var_phase(const DevGraph& d, const float* m, int64_t m_stride, const float* ev,
                      int64_t ev_stride, float* out, float* partial, int32_t batch,
                      hipStream_t s, const KernelClock* clk = nullptr, int32_t xcd = 0);
template <bool kT0>
int launch_phased(pgx_graph* g, const RunBufs& r, int32_t batch, hipStream_t s,
                  const KernelClock& clk, int64_t tile, int64_t threads, bool lanes);
template <bool kT0>
int launch_coop(pgx_graph* g, const RunBufs& r, int32_t batch, hipStream_t s, int64_t tile);
// one float per binary message: engine-layout pairs -> packed, packed -> caller layout
void launch_pair_pack(const DevGraph& d, const float* full, float* packed, int64_t stride,
                      int32_t batch, hipStream_t s);
void launch_pair_unpack(const DevGraph& d, const float* packed, int64_t stride, float* out,
                        int32_t batch, hipStream_t s);
// the plugin API: one block kind's update on given vtof, for `batch` elements
// (vtof and ftov num_es floats apart, lp lp_stride apart)
int launch_fac_op(pgx_graph* g, int kind, const float* vtof, const float* lp, float* ftov,
                  float temperature, int32_t normalize, hipStream_t s, int32_t batch,
                  int64_t lp_stride);
// pgx_fac_to_var over a batch (pgx_engine.hip)
int fac_to_var_batched(pgx_graph* g, int32_t factor_type, const float* vtof, const float* lp,
                       int64_t lp_stride, float* ftov, float temperature, int32_t normalize,
                       int32_t batch, hipStream_t s);

// pgx_task.hip: one workgroup per (task, batch element), every iteration.
template <bool kT0>
int launch_task_loop(pgx_graph* g, const RunBufs& r, int32_t batch, int bs, hipStream_t s);

// pgx_bip.hip: complete bipartite binary graphs, one launch per iteration.
template <bool kT0>
void launch_bipartite(const DevGraph& d, const RunBufs& r, const float* msg_in, float* msg_out,
                      const float* pin, float* pout, int32_t it, int32_t batch, hipStream_t s,
                      int64_t cols, int64_t threads);

}  // namespace pgx

#endif  // PGX_LAUNCH_H_
