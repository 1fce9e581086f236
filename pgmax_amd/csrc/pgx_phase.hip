// Per-phase shape: every BP iteration is a few launches over the whole
// (batched) graph -- the variable phase, then one factor-phase launch per
// block kind -- or one launch for graphs whose factor side recomputes
// variable sums. Also the one-float message pack / unpack passes and the
// plugin API's one-type factor update.

#include <hip/hip_cooperative_groups.h>

#include <algorithm>
#include <type_traits>

#include "pgx_launch.h"

namespace pgx {
namespace {

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
                int32_t col_blocks, int32_t xcd) {
  uint32_t bx, by;
  xcd_chunked(uint32_t(xcd), bx, by);
  const int32_t b = int32_t(by);
  const int32_t blk_x = int32_t(bx);
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

// Phased shape: factor phase over tiles of <= kTileFactors factors.
template <bool kT0, int KIND>
__device__ __forceinline__ void fac_phase_body(const DevGraph& g, const Tile* tiles,
                                               const RunBufs& r, int32_t it, float t) {
  uint32_t bx, by;
  xcd_chunked(uint32_t(r.xcd), bx, by);
  const int32_t b = int32_t(by);
  const Tile tile = tiles[bx];
  const Block blk = g.blocks[tile.block];
  const float* xb = r.x + b * g.num_es;
  const float* yb = r.y + b * g.num_es;
  const float* mo = ((r.n_it - it) & 1) ? yb : xb;
  float* mn = const_cast<float*>(((r.n_it - 1 - it) & 1) ? yb : xb);
  EngineIO io{r.vsum + b * g.num_vs, mo, mn, r.lp + b * r.lp_stride, r.d, r.omd, g.edge_vs0,
              g.even2 != 0, g.packed != 0};
  float dl = 0.f;
  if constexpr (KIND == kOr || KIND == kAnd) {
    if (tile.pad == 64) {  // one wavefront per factor
      const int32_t i = tile.i0 + int32_t(threadIdx.x) / 64;
      if (i < tile.i1)
        logical_factor_wave<EngineIO>(g, blk.f0 + i, blk.off, io, t, true, dl,
                                      int32_t(threadIdx.x) % 64);
      block_delta_store<kPhaseThreads>(dl, delta_slot(r, b, it));
      return;
    }
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

// Square pairwise Enum blocks (S x S, 3 <= S <= 8; the fused form also
// 2 x 2) on S lanes per factor (pair_lanes): the tile's factors go to the
// workgroup's wavefronts in passes of 64 / S factors per wavefront;
// wavefronts run independently.
template <bool kT0, int S, class L>
__device__ __forceinline__ void pair_lanes_body(const L& ld, const Tile& tile, const EngineIO& io,
                                                float t, float* lds, float& dl) {
  constexpr int kFw = 64 / S;  // factors per wavefront and pass
  constexpr int kWaves = kPhaseThreads / 64;
  const int32_t lane = int32_t(threadIdx.x) & 63, w = int32_t(threadIdx.x) >> 6;
  const int32_t fl = lane / S, p = lane % S;  // fl == kFw: lanes past the whole groups
  float* xs = lds + fl * 2 * S;
  float* sm = lds + (kFw + 1) * 2 * S + fl * S * (S + 1);
  const int32_t n = tile.i1 - tile.i0;
  const int32_t passes = (n + kWaves * kFw - 1) / (kWaves * kFw);
  if (ld.lp_fix >= 0) {
    // every factor of the tile holds one table: its row p, clipped, once
    float rowc[S];
    ldv<S>(io.lp + ld.b.c0 + S * S * ld.lp_fix + p * S, rowc);
#pragma unroll
    for (int q = 0; q < S; ++q) rowc[q] = clip_lp(rowc[q]);
    pair_lanes<kT0, S, L, true>(ld, tile.i0 + w * kFw + fl, kWaves * kFw, passes, tile.i1,
                                fl < kFw, p, io, t, dl, xs, sm, rowc);
  } else {
    pair_lanes<kT0, S>(ld, tile.i0 + w * kFw + fl, kWaves * kFw, passes, tile.i1, fl < kFw, p,
                       io, t, dl, xs, sm);
  }
}

// kFused: variable sums formed on the factor side (PairFusedLoader, no
// variable phase); else read from the variable phase's sums.
template <bool kT0, bool kFused>
__device__ __forceinline__ void pair_lanes_phase(const DevGraph& g, const Tile* tiles,
                                                 const RunBufs& r, int32_t it, float t) {
  __shared__ float lds[kPhaseThreads / 64][kPairLanesLds];
  uint32_t bx, by;
  xcd_chunked(uint32_t(r.xcd), bx, by);
  const int32_t b = int32_t(by);
  const Tile tile = tiles[bx];
  const Block blk = g.blocks[tile.block];
  const float* xb = r.x + b * g.num_es;
  const float* yb = r.y + b * g.num_es;
  const float* mo = ((r.n_it - it) & 1) ? yb : xb;
  float* mn = const_cast<float*>(((r.n_it - 1 - it) & 1) ? yb : xb);
  EngineIO io{kFused ? nullptr : r.vsum + b * g.num_vs, mo, mn, r.lp + b * r.lp_stride, r.d,
              r.omd, g.edge_vs0, g.even2 != 0, false};
  float dl = 0.f;
  float* wl = lds[threadIdx.x >> 6];
  // every factor of the tile holds the tile's first factor's table
  const int32_t lp_fix =
      (r.share && r.share[int64_t(b) * r.share_bstride + int32_t(bx)]) ? tile.i0 : -1;
  auto body = [&](auto sc) {
    constexpr int S = decltype(sc)::value;
    if constexpr (kFused) {
      const PairFusedLoader ld{g, io, r.ev + b * r.ev_stride, blk, lp_fix};
      pair_lanes_body<kT0, S>(ld, tile, io, t, wl, dl);
    } else {
      const PairVsumLoader ld{g, io, blk, lp_fix};
      pair_lanes_body<kT0, S>(ld, tile, io, t, wl, dl);
    }
  };
  switch (tile.pad) {  // the block's state count; uniform: a tile lies in one block
    case 2:
      if constexpr (kFused) body(std::integral_constant<int, 2>{});
      break;
    case 3: body(std::integral_constant<int, 3>{}); break;
    case 4: body(std::integral_constant<int, 4>{}); break;
    case 5: body(std::integral_constant<int, 5>{}); break;
    case 6: body(std::integral_constant<int, 6>{}); break;
    case 7: body(std::integral_constant<int, 7>{}); break;
    default: body(std::integral_constant<int, 8>{}); break;
  }
  block_delta_store<kPhaseThreads>(dl, delta_slot(r, b, it));
}

// Lane-group tile t of batch element b: do all its factors hold the tile's
// first factor's log-potential table, bit for bit (as every factor of a
// PairwiseFactorGroup built from one 2-D matrix does)? Then k_pair_lanes
// reads that one table for every factor of the tile: the same values, so the
// same results, from one table's bytes instead of one per factor. One pass
// per run reads every table once.
__global__ void __launch_bounds__(256)
    k_tile_share(DevGraph g, const Tile* tiles, const float* lp, int64_t lp_stride,
                 int32_t* share, int32_t num_tiles) {
  const int32_t t = blockIdx.x, b = blockIdx.y;
  const Tile tile = tiles[t];
  const Block blk = g.blocks[tile.block];
  const int32_t ss = tile.pad * tile.pad;
  const uint32_t* tab =
      reinterpret_cast<const uint32_t*>(lp + b * lp_stride + blk.c0) + int64_t(tile.i0) * ss;
  const int32_t n = (tile.i1 - tile.i0) * ss;  // a lane tile holds a few hundred factors
  int32_t j = ss + int32_t(threadIdx.x), q = j % ss;
  const int32_t step = 256 % ss;
  int32_t same = 1;
  for (; j < n; j += 256) {
    same &= tab[j] == tab[q] ? 1 : 0;
    q += step;
    if (q >= ss) q -= ss;
  }
  same = __syncthreads_and(same);
  if (threadIdx.x == 0) share[int64_t(b) * num_tiles + t] = same;
}

template <bool kT0>
__global__ void __launch_bounds__(kPhaseThreads)
    k_pair_lanes(DevGraph g, const Tile* tiles, RunBufs r, int32_t it) {
  if (!kT0 && r.t == 1.f)
    pair_lanes_phase<kT0, false>(g, tiles, r, it, 1.f);
  else
    pair_lanes_phase<kT0, false>(g, tiles, r, it, r.t);
}

// The whole iteration of a graph of 2 x 2 and square pairwise blocks whose
// variables have at most 4 edges, in one launch (no variable phase).
template <bool kT0>
__global__ void __launch_bounds__(kPhaseThreads)
    k_pair_lanes_fused(DevGraph g, const Tile* tiles, RunBufs r, int32_t it) {
  if (!kT0 && r.t == 1.f)
    pair_lanes_phase<kT0, true>(g, tiles, r, it, 1.f);
  else
    pair_lanes_phase<kT0, true>(g, tiles, r, it, r.t);
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

// One factor type's update on given vtof (plugin API).
template <bool kT0, int KIND>
__global__ void __launch_bounds__(kPhaseThreads)
    k_fac_op(DevGraph g, const Tile* tiles, const float* vtof, const float* lp, float* out,
             float t, int32_t normalize, int64_t lp_stride) {
  const Tile tile = tiles[blockIdx.x];
  const Block blk = g.blocks[tile.block];
  const int64_t b = blockIdx.y;  // batch element: vtof and out of num_es floats each
  OpIO io{vtof + b * g.num_es, out + b * g.num_es, lp + b * lp_stride};
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

}  // namespace

static inline dim3 var_grid(const DevGraph& d, int32_t batch, int32_t* light_blocks,
                            int32_t* col_blocks) {
  *light_blocks = (d.num_vars + kPhaseThreads - 1) / kPhaseThreads;
  *col_blocks = (d.num_heavy_col + kPhaseThreads - 1) / kPhaseThreads;
  const int32_t per = kPhaseThreads / kRowLanes;
  const int32_t row_blocks = (d.num_heavy_row + per - 1) / per;
  return dim3(*light_blocks + *col_blocks + row_blocks, batch);
}

void launch_var_phase(const DevGraph& d, const float* m, int64_t m_stride, const float* ev,
                      int64_t ev_stride, float* out, float* partial, int32_t batch,
                      hipStream_t s, const KernelClock* clk, int32_t xcd) {
  int32_t light_blocks, col_blocks;
  const dim3 vgrid = var_grid(d, batch, &light_blocks, &col_blocks);
  const KernelClock off;
  const KernelClock& c = clk ? *clk : off;
  c.run(PGX_KERNEL_VAR_PHASE, [&] {
    hipLaunchKernelGGL(k_var_phase, vgrid, dim3(kPhaseThreads), 0, s, d, m, m_stride, ev,
                       ev_stride, out, partial, light_blocks, col_blocks, xcd);
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

void launch_pair_pack(const DevGraph& d, const float* full, float* packed, int64_t stride,
                      int32_t batch, hipStream_t s) {
  hipLaunchKernelGGL(k_pair_pack,
                     dim3(unsigned(std::min<int64_t>((d.num_es / 2 + 255) / 256, 2048)), batch),
                     dim3(256), 0, s, d, full, packed, stride);
}

void launch_pair_unpack(const DevGraph& d, const float* packed, int64_t stride, float* out,
                        int32_t batch, hipStream_t s) {
  const dim3 ugrid(unsigned(std::min<int64_t>((d.num_es + 255) / 256, 2048)), batch);
  hipLaunchKernelGGL(k_pair_unpack, ugrid, dim3(256), 0, s, d, packed, stride, out);
}

template <bool kT0>
static void launch_fac_kinds(pgx_graph* g, const TileSet& T, const RunBufs& r, int32_t batch,
                             int32_t it, hipStream_t s, const KernelClock& clk, bool lanes) {
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
  // square pairwise blocks on lane groups or one factor per thread (timed
  // as the pairwise Enum kind)
  constexpr int kL = kPairLanesGroup, kT = kPairThreadGroup;
  if (lanes && k0[kL + 1] > k0[kL])
    clk.run(id0 + kEnumPair, [&] {
      hipLaunchKernelGGL((k_pair_lanes<kT0>), grid(kL), dim3(kPhaseThreads), 0, s, d,
                         T.tiles + k0[kL], r, it);
    });
  if (!lanes && k0[kT + 1] > k0[kT])
    clk.run(id0 + kEnumPair, [&] {
      hipLaunchKernelGGL((k_fac_phase<kT0, kEnumPair>), grid(kT), blk, 0, s, d, T.tiles + k0[kT],
                         r, it);
    });
}

// The run's buffers with RunBufs::share set for the nl lane tiles at
// `tiles` (one k_tile_share pass), or unset when the run does not look for
// shared tables (r.share_mem null) or there are no such tiles.
static RunBufs with_shared_tables(const DevGraph& d, const Tile* tiles, int32_t nl,
                                  const RunBufs& r, int32_t batch, hipStream_t s) {
  RunBufs rr = r;
  rr.share = nullptr;
  rr.share_bstride = 0;
  if (!r.share_mem || nl <= 0 || !r.lp) return rr;
  const int32_t nb = r.lp_stride ? batch : 1;
  hipLaunchKernelGGL(k_tile_share, dim3(nl, nb), dim3(256), 0, s, d, tiles, r.lp, r.lp_stride,
                     r.share_mem, nl);
  rr.share = r.share_mem;
  rr.share_bstride = r.lp_stride ? nl : 0;
  return rr;
}

template <bool kT0>
int launch_phased(pgx_graph* g, const RunBufs& r, int32_t batch, hipStream_t s,
                  const KernelClock& clk, int64_t tile, int64_t threads, bool lanes) {
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
  if (d.fusable && lanes && !d.packed && g->lsets[0].num_tiles > 0) {
    const TileSet& L = g->lsets[tile <= 64 ? 0 : (tile <= 128 ? 1 : 2)];
    const dim3 lgrid(L.num_tiles, batch);
    const RunBufs rr = with_shared_tables(d, L.tiles, L.num_tiles, r, batch, s);
    for (int32_t it = 0; it < rr.n_it; ++it)
      clk.run(PGX_KERNEL_FAC_FUSED, [&] {
        hipLaunchKernelGGL((k_pair_lanes_fused<kT0>), lgrid, dim3(kPhaseThreads), 0, s, d,
                           L.tiles, rr, it);
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
  const TileSet& M = g->msets[tile <= 64 ? 0 : (tile <= 128 ? 1 : 2)];
  const int32_t* lk = M.kind_tile0 + kPairLanesGroup;
  const RunBufs rr = lanes ? with_shared_tables(d, M.tiles + lk[0], lk[1] - lk[0], r, batch, s)
                           : with_shared_tables(d, nullptr, 0, r, batch, s);
  for (int32_t it = 0; it < rr.n_it; ++it) {
    const float* mo = ((rr.n_it - it) & 1) ? rr.y : rr.x;
    if (d.num_vars > 0)
      launch_var_phase(d, mo, d.num_es, rr.ev, rr.ev_stride, rr.vsum, rr.partial, batch, s, &clk,
                       rr.xcd);
    launch_fac_kinds<kT0>(g, M, rr, batch, it, s, clk, lanes);
  }
  PGX_HIP(hipGetLastError());
  return PGX_OK;
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
int launch_coop(pgx_graph* g, const RunBufs& r, int32_t batch, hipStream_t s,
                       int64_t tile) {
  const TileSet& T = g->tsets[tile <= 64 ? 0 : (tile <= 128 ? 1 : 2)];
  if (g->kinds_mask == kKindsBin2) return launch_coop_t<kT0, kKindsBin2>(g, T, r, batch, s);
  return launch_coop_t<kT0, kKindsPairwise>(g, T, r, batch, s);
}

int launch_fac_op(pgx_graph* g, int kind, const float* vtof, const float* lp, float* ftov,
                  float temperature, int32_t normalize, hipStream_t s, int32_t batch,
                  int64_t lp_stride) {
  const int32_t a = g->kind_tile0[kind], z = g->kind_tile0[kind + 1];
  if (z == a) return PGX_OK;
  const dim3 grid(z - a, batch);
  const Tile* tl = g->tiles + a;
  const bool t0 = temperature == 0.f;
#define PGX_OP_CASE(K)                                                                      \
  case K:                                                                                   \
    if (t0)                                                                                 \
      hipLaunchKernelGGL((k_fac_op<true, K>), grid, dim3(kPhaseThreads), 0, s, g->dev, tl,  \
                         vtof, lp, ftov, temperature, normalize, lp_stride);                \
    else                                                                                    \
      hipLaunchKernelGGL((k_fac_op<false, K>), grid, dim3(kPhaseThreads), 0, s, g->dev, tl, \
                         vtof, lp, ftov, temperature, normalize, lp_stride);                \
    break;
  switch (kind) {
    PGX_OP_CASE(0)
    PGX_OP_CASE(1)
    PGX_OP_CASE(2)
    PGX_OP_CASE(3)
    PGX_OP_CASE(4)
    PGX_OP_CASE(5)
    default:
      return set_error(PGX_EINVAL, "launch_fac_op: bad block kind");
  }
#undef PGX_OP_CASE
  PGX_HIP(hipGetLastError());
  return PGX_OK;
}

template int launch_phased<true>(pgx_graph*, const RunBufs&, int32_t, hipStream_t,
                                 const KernelClock&, int64_t, int64_t, bool);
template int launch_phased<false>(pgx_graph*, const RunBufs&, int32_t, hipStream_t,
                                  const KernelClock&, int64_t, int64_t, bool);
template int launch_coop<true>(pgx_graph*, const RunBufs&, int32_t, hipStream_t, int64_t);
template int launch_coop<false>(pgx_graph*, const RunBufs&, int32_t, hipStream_t, int64_t);

}  // namespace pgx
