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
#include "pgx_launch.h"

namespace pgx {

thread_local std::string g_last_error;

int set_error(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

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
int64_t g_opt_lanes = 1;     // 1: square pairwise tables on S lanes per factor (per-phase shape)
int64_t g_opt_xcd = 0;       // per-phase kernels: workgroups per XCD run (xcd_chunked), 0 off
int64_t g_opt_share = 1;     // 1: lane tiles of identical pairwise tables read one table
int64_t g_opt_mpnn_rows = 1; // 1: warm-start prediction in variable-major row order
std::mutex g_opt_mu;
std::vector<std::pair<hipEvent_t, hipEvent_t>> g_prof_events;
std::vector<KernelRecord> g_kernel_events;

namespace {

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

// MAP state and marginals of every variable from its log-beliefs in one pass
// over them: the first maximal state (a NaN counts as maximal and stops the
// scan, as jnp.argmax takes it; pgmax/infer/inferer.py:251-264) and
// exp(b - m) / sum of exp(b - m) over the variable's states, m the
// NaN-propagating max and the sum in state order (jax.nn.softmax, as
// pgmax/infer/bp.py:263-288 applies it). Either output may be null.
__global__ void k_map_marginals(DevGraph g, const float* beliefs, int32_t* map, float* marg) {
  const int32_t b = blockIdx.y;
  const int32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= g.num_vars) return;
  const float* x = beliefs + int64_t(b) * g.num_vs;
  const int32_t a = g.var_vs0[v], z = g.var_vs0[v + 1];
  int32_t best = 0;
  float m = x[a];
  for (int32_t k = a + 1; k < z && m == m; ++k)
    if (x[k] > m || x[k] != x[k]) {
      m = x[k];
      best = k - a;
    }
  if (map) map[int64_t(b) * g.num_vars + v] = best;
  if (marg) {
    float sum = 0.f;
    for (int32_t k = a; k < z; ++k) sum += pgx_exp(x[k] - m);
    float* y = marg + int64_t(b) * g.num_vs;
    for (int32_t k = a; k < z; ++k) y[k] = pgx_exp(x[k] - m) / sum;
  }
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
      for (int32_t k = a; k < z; ++k) u += exp_le0(summary[es_row_cfg[k]] - mx, t);
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
  } else if (key && std::strcmp(key, "lanes") == 0) {
    prev = g_opt_lanes;
    g_opt_lanes = value;
  } else if (key && std::strcmp(key, "xcd") == 0) {
    prev = g_opt_xcd;
    g_opt_xcd = value;
  } else if (key && std::strcmp(key, "share") == 0) {
    prev = g_opt_share;
    g_opt_share = value;
  } else if (key && std::strcmp(key, "merge") == 0) {
    prev = g_opt_merge;
    g_opt_merge = value;
  } else if (key && std::strcmp(key, "mpnn_rows") == 0) {
    prev = g_opt_mpnn_rows;
    g_opt_mpnn_rows = value;
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
    d.bip_unary = P.bip_unary;
    d.bip_pb = P.bip_pb;
    d.bip_uc0 = P.bip_uc0;
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
  // a lane group / wavefront per factor); msets (the per-phase kernels):
  // the same, except that 2x2 Enum blocks take kBin2PerThread factors per
  // thread, and square pairwise blocks of 3..8 states go to two groups of
  // their own, one of which runs per launch (option "lanes"): S lanes per
  // factor (T.size / 64 passes of kPhaseThreads / 64 * (64 / S) factors per
  // tile) or one factor per thread
  auto lane_pair = [](const Block& b) {
    return b.kind == kEnumPair && b.s0 == b.s1 && b.s0 >= 3 && b.s0 <= 8;
  };
  for (int ts = 0; ts < 6; ++ts) {
    const bool multi = ts >= 3;
    TileSet& T = multi ? g->msets[ts - 3] : g->tsets[ts];
    T.size = kTileSizes[ts % 3];
    std::vector<Tile> tiles;
    for (int k = 0; k < kTileGroups; ++k) {
      T.kind_tile0[k] = int32_t(tiles.size());
      for (int32_t bi = 0; bi < int32_t(P.blocks.size()); ++bi) {
        const Block& b = P.blocks[bi];
        if (k == kPairLanesGroup) {
          if (!multi || !lane_pair(b)) continue;
          const int32_t per = (T.size / 64) * (kPhaseThreads / 64) * (64 / b.s0);
          for (int32_t i0 = 0; i0 < b.count; i0 += per)
            tiles.push_back(Tile{bi, i0, std::min(b.count, i0 + per), b.s0});
          continue;
        }
        if (k == kPairThreadGroup) {
          if (!multi || !lane_pair(b)) continue;
          for (int32_t i0 = 0; i0 < b.count; i0 += T.size)
            tiles.push_back(Tile{bi, i0, std::min(b.count, i0 + T.size), 0});
          continue;
        }
        if (b.kind != k) continue;
        g->kinds_mask |= 1 << k;
        if (multi && lane_pair(b)) continue;
        // wide logical blocks: kWideLanes threads per factor; logical blocks
        // of at most 256 parents and large general Enum factors: one
        // wavefront each
        const int32_t lanes =
            (b.pad[0] & 4) ? kWideLanes : ((b.pad[0] & (8 | 16)) ? 64 : 1);
        const bool m = multi && b.kind == kEnumBin2;
        const int32_t per = m ? T.size * kBin2PerThread : T.size / lanes;
        for (int32_t i0 = 0; i0 < b.count; i0 += per)
          tiles.push_back(Tile{bi, i0, std::min(b.count, i0 + per),
                               m ? -kBin2PerThread : (lanes == 1 ? 0 : lanes)});
      }
    }
    T.kind_tile0[kTileGroups] = int32_t(tiles.size());
    T.num_tiles = int32_t(tiles.size());
    if ((rc = up(tiles, &T.tiles))) {
      pgx_graph_destroy(g);
      return rc;
    }
  }
  // fused lane-group tiles: every block a 2 x 2 or square pairwise block of
  // 3..8 states, every variable of at most 4 edges with its descriptor
  bool lane_fusable = !P.blocks.empty() && !P.var_pos4.empty() && P.max_degree <= 4;
  for (const Block& b : P.blocks)
    lane_fusable = lane_fusable && (b.kind == kEnumBin2 || lane_pair(b));
  for (int ts = 0; ts < 3 && lane_fusable; ++ts) {
    TileSet& T = g->lsets[ts];
    T.size = kTileSizes[ts];
    std::vector<Tile> tiles;
    for (int32_t bi = 0; bi < int32_t(P.blocks.size()); ++bi) {
      const Block& b = P.blocks[bi];
      const int32_t s = b.kind == kEnumBin2 ? 2 : b.s0;
      const int32_t per = kLaneFusedPasses[ts] * (kPhaseThreads / 64) * (64 / s);
      for (int32_t i0 = 0; i0 < b.count; i0 += per)
        tiles.push_back(Tile{bi, i0, std::min(b.count, i0 + per), s});
    }
    T.num_tiles = int32_t(tiles.size());
    T.kind_tile0[kTileGroups] = T.num_tiles;
    if ((rc = up(tiles, &T.tiles))) {
      pgx_graph_destroy(g);
      return rc;
    }
  }
  static_assert(kTileSizes[2] == kTileFactors, "the plugin API uses the largest tiles");
  g->tiles = g->tsets[2].tiles;
  for (int k = 0; k <= kTileGroups; ++k) g->kind_tile0[k] = g->tsets[2].kind_tile0[k];
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

// Flags of the lane tiles (RunBufs::share), for the largest lane group of
// the three tile sets and the largest set of fused lane tiles.
static size_t share_bytes(const pgx_graph* g, int32_t batch) {
  int64_t n = 0;
  for (const TileSet& T : g->msets)
    n = std::max<int64_t>(n, T.kind_tile0[kPairLanesGroup + 1] - T.kind_tile0[kPairLanesGroup]);
  for (const TileSet& T : g->lsets) n = std::max<int64_t>(n, T.num_tiles);
  return size_t(batch) * size_t(n) * sizeof(int32_t);
}

// Scratch of one run: message buffer(s), variable sums, partial sums, delta
// slots, lane-tile flags.
static size_t run_scratch_bytes(const pgx_graph* g, int32_t batch, int32_t n_it) {
  const DevGraph& d = g->dev;
  const size_t msg_bytes = size_t(batch) * size_t(d.num_es) * sizeof(float);
  const size_t vs_bytes = size_t(batch) * size_t(d.num_vs) * sizeof(float);
  const size_t part_bytes = size_t(batch) * size_t(d.partial_size) * sizeof(float);
  const int32_t ns = delta_slots(g, batch, n_it);
  const size_t slot_bytes = ns > 1 ? size_t(batch) * size_t(n_it) * size_t(ns) * 4 : 0;
  // complete bipartite graphs alternate between two chunk-sum buffers
  return msg_bytes * (d.perm != nullptr ? 2 : 1) + vs_bytes +
         part_bytes * (d.bip_na > 0 ? 2 : 1) + slot_bytes + share_bytes(g, batch) + 512;
}

// Shared body of pgx_run and pgx_run_converge. With task_deltas set, the
// run uses the task-loop shape and stops every task at its convergence.
// fixed_scratch (non-null): the run's scratch lies there (the caller checked
// that run_scratch_bytes fits) instead of the graph's own, and the run makes
// no allocation of any kind (no profiling events either).
static int run_impl(pgx_graph* g, const float* lp, int64_t lp_stride, const float* ev,
                    int64_t ev_stride, const float* ftov_in, int64_t ftov_in_stride,
                    float* ftov_out, float* deltas, float* task_deltas, int32_t* iters_done,
                    int32_t batch, int32_t num_iters, float damping, float temperature, float tol,
                    void* stream, const RunCfg* force = nullptr,
                    float* fixed_scratch = nullptr) {
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
  int rc = fixed_scratch ? PGX_OK : ensure_scratch(g, run_scratch_bytes(g, batch, n_it));
  if (rc) return rc;
  float* const scr = fixed_scratch ? fixed_scratch : g->scratch;
  RunBufs r;
  r.lp = lp;
  r.lp_stride = lp_stride;
  r.ev = ev;
  r.ev_stride = ev_stride;
  r.y = scr;
  r.x = planar ? scr + size_t(batch) * size_t(d.num_es) : ftov_out;
  r.vsum = scr + size_t(batch) * size_t(d.num_es) * (planar ? 2 : 1);
  r.partial = r.vsum + size_t(batch) * size_t(d.num_vs);
  r.deltas = reinterpret_cast<uint32_t*>(deltas);
  r.dslot_n = converge ? 1 : delta_slots(g, batch, n_it);
  float* const after_partial =
      r.partial + size_t(batch) * size_t(d.partial_size) * (d.bip_na > 0 ? 2 : 1);
  r.dslot = r.dslot_n > 1 ? reinterpret_cast<uint32_t*>(after_partial) : r.deltas;
  {
    const int32_t ns = delta_slots(g, batch, n_it);  // the slots run_scratch_bytes sized
    r.share_mem = reinterpret_cast<int32_t*>(
        after_partial + (ns > 1 ? size_t(batch) * size_t(n_it) * size_t(ns) : 0));
  }
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
  int64_t mode, bs, prof, fuse, vec, tile, pack, lanes, xcd, share;
  {
    std::lock_guard<std::mutex> lock(g_opt_mu);
    pack = g_opt_pack;
    lanes = g_opt_lanes;
    xcd = g_opt_xcd;
    share = g_opt_share;
    mode = g_opt_mode;
    bs = g_opt_block;
    prof = g_opt_profile;
    fuse = g_opt_fuse;
    vec = g_opt_vec;
    tile = g_opt_tile;
  }
  if (fixed_scratch) prof = 0;  // profiling events would be allocations
  bool chosen = force != nullptr;  // shape set by the autotuner
  if (force) {
    mode = force->mode;
    bs = force->bs;
    fuse = force->fuse;
    vec = force->vec;
    tile = force->tile;
    lanes = force->lanes;
    xcd = force->xcd;
  } else if (mode == 0 && bs == 0 && tile == 0) {
    auto it = g->tuned.find(std::make_tuple(batch, temperature == 0.f ? 1 : 0, converge ? 1 : 0));
    if (it != g->tuned.end()) {
      mode = it->second.mode;
      bs = it->second.bs;
      fuse = it->second.fuse;
      vec = it->second.vec;
      tile = it->second.tile;
      lanes = it->second.lanes;
      xcd = it->second.xcd;
      chosen = true;
    }
  }
  if (tile == 0) tile = 256;
  r.xcd = int32_t(std::max<int64_t>(0, std::min<int64_t>(xcd, 1024)));
  const bool small_run = int64_t(batch) * d.num_es <= kFusedMaxStates;
  // k_bipartite numbers (tile, batch element) pairs in one grid dimension
  const int64_t bip_units = int64_t((d.bip_na + d.bip_unary + kBipRows - 1) / kBipRows) *
                            ((d.bip_nb + d.bip_unary + kVarChunk - 1) / kVarChunk) * batch;
  g->dev.bip = (g->bip_plan && fuse && bip_units <= int64_t(INT32_MAX)) ? 1 : 0;
  g->dev.even2 = ((vec & 1) && g->even2_plan && (d.num_es & 1) == 0 && (d.num_vs & 1) == 0 &&
                  (ev_stride & 1) == 0 && aligned(ev, 8) && aligned(ftov_out, 8))
                     ? 1
                     : 0;
  g->dev.lp_al4 = ((vec & 2) && lp_stride % 4 == 0 && aligned(lp, 16)) ? ((vec & 4) ? 2 : 1) : 0;
  g->dev.lp_nt = (vec & 4) ? 1 : 0;
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
      (g->fusable_plan && fuse && (fuse == 2 || chosen || small_run || mode != 2)) ? 1 : 0;
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
    launch_pair_pack(g->dev, first, other, d.num_es, batch, s);
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
    // reading shared tables pays when the one pass that finds them is short
    // next to the run (it reads every table once)
    if (!(share && n_it >= 8)) r.share_mem = nullptr;
    rc = t0 ? launch_phased<true>(g, r, batch, s, clk, tile, bs, lanes != 0)
            : launch_phased<false>(g, r, batch, s, clk, tile, bs, lanes != 0);
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
      launch_pair_unpack(g->dev, (n_it & 1) ? m1 : m0, g->dev.num_es / 2, ftov_out, batch, s);
    } else if (g->dev.packed) {  // one float per message at the start of each batch slot
      launch_pair_unpack(g->dev, r.x, g->dev.num_es, ftov_out, batch, s);
    } else {
      hipLaunchKernelGGL(k_unpermute, ugrid, dim3(256), 0, s, d, static_cast<const float*>(r.x),
                         ftov_out);
    }
    PGX_HIP(hipGetLastError());
  }
  if (prof) {
    PGX_HIP(hipEventRecord(ev1, s));
    std::lock_guard<std::mutex> lock(g_opt_mu);

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
    key = {g_opt_mode, g_opt_block, g_opt_fuse,  g_opt_vec, g_opt_tile,
           g_opt_pack, g_opt_lanes, g_opt_xcd, g_opt_share};
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

}  // extern "C"

// ---- allocation-free run ----------------------------------------------------
// The scratch of pgx_run_noalloc: one device array on every GPU the library
// runs on, held until the process ends. In host code its name is only a
// stand-in; its device address on the current GPU comes from
// hipGetSymbolAddress, taken once per device at the first call there.
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
static __device__ __align__(256) float scratch[kNoAllocScratchBytes / sizeof(float)];

namespace {
constexpr int kNoAllocMaxDevices = 64;
std::mutex g_noalloc_mu;
float* g_noalloc_scratch[kNoAllocMaxDevices] = {};
}  // namespace

// The current device's scratch address (looked up on the first call there).
static int noalloc_scratch(float** out) {
  int dev = 0;
  PGX_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoAllocMaxDevices)
    return set_error(PGX_EUNSUP, "pgx_run_noalloc: device index beyond 63");
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  if (!g_noalloc_scratch[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return set_error(PGX_EHIP, std::string("pgx_run_noalloc: hipGetSymbolAddress: ") +
                                     hipGetErrorString(e));
    g_noalloc_scratch[dev] = static_cast<float*>(ptr);
  }
  *out = g_noalloc_scratch[dev];
  return PGX_OK;
}

extern "C" {

int pgx_run_noalloc(pgx_graph* g, const float* lp, int64_t lp_stride, const float* ev,
                    int64_t ev_stride, const float* ftov_in, int64_t ftov_in_stride,
                    float* ftov_out, float* deltas, int32_t batch, int32_t num_iters,
                    float damping, float temperature, void* stream) {
  if (!g || batch <= 0) return set_error(PGX_EINVAL, "pgx_run_noalloc: null graph or batch <= 0");
  if (batch > kMaxBatchPerLaunch)
    return set_error(PGX_EINVAL, "pgx_run_noalloc: batch above 65535 (one launch per call)");
  const size_t need = run_scratch_bytes(g, batch, std::max(num_iters, 1));
  if (need > kNoAllocScratchBytes)
    return set_error(PGX_EINVAL, "pgx_run_noalloc: this run needs " + std::to_string(need) +
                                     " bytes of scratch, more than the " +
                                     std::to_string(kNoAllocScratchBytes) +
                                     " the library holds per GPU (use pgx_run)");
  float* scr = nullptr;
  const int rc = noalloc_scratch(&scr);
  if (rc) return rc;
  return run_impl(g, lp, lp_stride, ev, ev_stride, ftov_in, ftov_in_stride, ftov_out, deltas,
                  nullptr, nullptr, batch, num_iters, damping, temperature, 0.f, stream, nullptr,
                  scr);
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
      // square pairwise tables: also thread per factor instead of S lanes
      if (!fuse_opts[fi] && g->msets[2].kind_tile0[kPairLanesGroup + 1] >
                                g->msets[2].kind_tile0[kPairLanesGroup])
        for (int64_t t : {128, 256}) cands.push_back(RunCfg{2, 0, 0, 3, t, 0});
      // runs of consecutive tiles per XCD (enough tiles for several rounds)
      if (!fuse_opts[fi] && g->tsets[0].num_tiles >= 64 * 8)
        for (int64_t x : {4, 16})
          for (int64_t ln : {1, 0}) cands.push_back(RunCfg{2, 0, 0, 3, 256, ln, x});
      // fused lane groups: also one factor per thread, and XCD runs
      if (fuse_opts[fi] && g->lsets[0].num_tiles > 0) {
        cands.push_back(RunCfg{2, 0, 1, 3, 256, 0, 0});
        for (int64_t x : {4, 16}) cands.push_back(RunCfg{2, 0, 1, 3, 256, 1, x});
      }
    }
  if (!converge && g->fusable_plan)
    for (int64_t t : {64, 256}) cands.push_back(RunCfg{3, 0, 1, 3, t});
  // the task loop last, so its probes can stop early against the best shape
  std::stable_partition(cands.begin(), cands.end(), [](const RunCfg& c) { return c.mode != 1; });
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
      // a shape whose first run takes 4x the best shape's time cannot win
      if (rep == 0 && best_ms < 3.4e38f && t > 4.f * best_ms) break;
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
    chosen[5] = best.lanes;
    chosen[6] = best.xcd;
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

int pgx_beliefs_decode(pgx_graph* g, const float* ev, int64_t ev_stride, const float* ftov,
                       int64_t ftov_stride, float* beliefs, int32_t* map_states,
                       float* marginals, int32_t batch, void* stream) {
  if (!g || !beliefs || batch <= 0) return set_error(PGX_EINVAL, "pgx_beliefs_decode: bad argument");
  if (batch > kMaxBatchPerLaunch) {
    for (int32_t b0 = 0; b0 < batch; b0 += kMaxBatchPerLaunch) {
      const int32_t nb = std::min(kMaxBatchPerLaunch, batch - b0);
      const int64_t vs0 = int64_t(b0) * g->dev.num_vs;
      const int rc = pgx_beliefs_decode(
          g, ev ? ev + b0 * ev_stride : ev, ev_stride, ftov ? ftov + b0 * ftov_stride : ftov,
          ftov_stride, beliefs + vs0, map_states ? map_states + int64_t(b0) * g->dev.num_vars : nullptr,
          marginals ? marginals + vs0 : nullptr, nb, stream);
      if (rc) return rc;
    }
    return PGX_OK;
  }
  int rc = pgx_get_beliefs(g, ev, ev_stride, ftov, ftov_stride, beliefs, batch, stream);
  if (rc || (!map_states && !marginals) || g->dev.num_vars == 0) return rc;
  const dim3 grid((g->dev.num_vars + 255) / 256, batch);
  hipLaunchKernelGGL(k_map_marginals, grid, dim3(256), 0, static_cast<hipStream_t>(stream), g->dev,
                     beliefs, map_states, marginals);
  PGX_HIP(hipGetLastError());
  return PGX_OK;
}

}  // extern "C"

namespace pgx {

int fac_to_var_batched(pgx_graph* g, int32_t factor_type, const float* vtof, const float* lp,
                       int64_t lp_stride, float* ftov, float temperature, int32_t normalize,
                       int32_t batch, hipStream_t s) {
  const int k_lo = factor_type == PGX_ENUM ? 0 : factor_type + 2;
  const int k_hi = factor_type == PGX_ENUM ? 3 : factor_type + 3;
  if (g->kind_tile0[k_hi] == g->kind_tile0[k_lo]) return PGX_OK;
  const int64_t m0 = g->plan.msgs_range[factor_type][0], m1 = g->plan.msgs_range[factor_type][1];
  // entries a logical / Pool update does not write stay 0
  // (pgmax/factor/logical.py:760-768); a fill kernel rather than a memset.
  // Enum updates write every edge state of their factors.
  if (factor_type != PGX_ENUM && m1 > m0)
    for (int32_t b = 0; b < batch; ++b)
      hipLaunchKernelGGL(k_fill_u32, dim3(unsigned(std::min<int64_t>((m1 - m0 + 255) / 256, 4096))),
                         dim3(256), 0, s,
                         reinterpret_cast<uint32_t*>(ftov + b * g->dev.num_es + m0), m1 - m0, 0u);
  for (int k = k_lo; k < k_hi; ++k) {
    const int rc = launch_fac_op(g, k, vtof, lp, ftov, temperature, normalize, s, batch,
                                 lp ? lp_stride : 0);
    if (rc) return rc;
  }
  PGX_HIP(hipGetLastError());
  return PGX_OK;
}

}  // namespace pgx

extern "C" {

int pgx_fac_to_var(pgx_graph* g, int32_t factor_type, const float* vtof, const float* lp,
                   float* ftov, float temperature, int32_t normalize, void* stream) {
  if (!g || !vtof || !ftov || factor_type < 0 || factor_type > 3)
    return set_error(PGX_EINVAL, "pgx_fac_to_var: bad argument");
  if (!(temperature >= 0.f)) return set_error(PGX_EINVAL, "pgx_fac_to_var: temperature < 0");
  if (!lp && factor_type == PGX_ENUM)
    return set_error(PGX_EINVAL, "pgx_fac_to_var: Enum updates need log-potentials");
  return fac_to_var_batched(g, factor_type, vtof, lp, 0, ftov, temperature, normalize, 1,
                            static_cast<hipStream_t>(stream));
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
