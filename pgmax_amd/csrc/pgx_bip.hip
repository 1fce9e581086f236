// Complete bipartite binary graphs (a binary RBM): the whole BP iteration
// in one launch per iteration, messages as one float per binary edge.

#include "pgx_launch.h"

namespace pgx {
namespace {

// The update of a unary factor with configs [0] and [1] (a unit's bias in
// the reference's RBM form) from its edge's vtof: enum_general's arithmetic
// for one variable, two configs -- summary = (0 + vtof) + lp, the max (T = 0)
// or T-log-sum-exp over the one config of each state, ftov = out - vtof.
template <bool kT0>
__device__ __forceinline__ void unary2_ftov(float v0, float v1, float l0, float l1, float t,
                                            float& f0, float& f1) {
  const float s0 = (0.f + v0) + clip_lp(l0), s1 = (0.f + v1) + clip_lp(l1);
  float o0 = nmax(kNegInf, s0), o1 = nmax(kNegInf, s1);
  if (!kT0) {
    const float u0 = 0.f + exp_le0(s0 - o0, t), u1 = 0.f + exp_le0(s1 - o1, t);
    o0 = mul_t(pgx_log(u0), t) + o0;
    o1 = mul_t(pgx_log(u1), t) + o1;
  }
  f0 = o0 - v0;
  f1 = o1 - v1;
}

// Complete bipartite binary graphs (Plan::bip_*): the whole iteration in
// one launch with no separate variable phase. Workgroups tile a grid of
// virtual rows x columns -- row i >= U the row unit i - U, column j >= U the
// column unit j - U, and with unary factors (U = 1) column 0 the row units'
// unary factors and row 0 the column units' -- so that a unit's edges in
// edge order (its unary factor, then its pairwise factors) run along its
// virtual row or column and every 32-edge run of its sum lies in one tile.
// A workgroup owns kBipRows x kCols cells of one batch element. It forms its
// units' sums from the previous iteration's chunk sums (the order of
// k_var_combine), updates the tile's factors (enum_bin2's and
// enum_general's arithmetic), and sums the new messages of each chunk in
// edge order from 0 (the order of k_var_phase) into part_out. Reading and
// writing different chunk-sum buffers keeps the workgroups of one launch
// independent. Thread t takes column t % kCols of rows t / kCols,
// t / kCols + kThreads / kCols, ...

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
  __shared__ float2 srow[kBipRows][kCols + 1];  // new messages of the rows' edges (padded)
  __shared__ float2 scol[kBipRows][kCols];      // new messages of the columns' edges
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
  const Block blk = g.blocks[g.bip_pb];
  const int32_t na = g.bip_na, nb = g.bip_nb, U = g.bip_unary;
  const int32_t a0 = rt * kBipRows, c0 = ct * kCols;  // first virtual row and column
  const int32_t nr = min(kBipRows, na + U - a0), ncl = min(kCols, nb + U - c0);
  const int32_t tid = threadIdx.x;
  const int32_t col = tid % kCols, row0 = tid / kCols;
  // messages as one float per edge (pair_encode): unary factors' edges at
  // [0, na + nb), the pairwise factors' first edges at [m0, m0 + n), second
  // edges at [m0 + n, m0 + 2n) of each batch element's num_es / 2 floats
  const int64_t npair = g.num_es / 2;
  const int32_t m0 = blk.es0 >> 1, n = blk.count;
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
  // loads. Variables of more chunks use chunk_total2. (Virtual row 0 and
  // column 0 of the unary form are no unit: their lanes load unit 0's data
  // and leave it unused.)
  constexpr int kSumLoads = 8;
  const float* pin = part_in + b * g.partial_size;
  const float* ev = r.ev + b * r.ev_stride;
  const int32_t nch_row = (nb + U + kVarChunk - 1) / kVarChunk;
  const int32_t nch_col = (na + U + kVarChunk - 1) / kVarChunk;
  const bool fast = max(nch_row, nch_col) <= 2 * kSumLoads;  // uniform
  const bool is_row = tid < 2 * kBipRows;
  const bool is_col = !is_row && tid < 2 * kBipRows + 2 * kCols;
  const int32_t vi = (is_row ? tid : tid - 2 * kBipRows) >> 1, half = tid & 1;
  const int32_t vg = is_row ? a0 + min(vi, nr - 1) : c0 + min(is_col ? vi : 0, ncl - 1);
  const int32_t ug = max(vg - U, 0);  // the unit
  int2 vinfo;
  if (g.bip_affine) {  // no table load ahead of the chunk-sum loads
    vinfo = is_row ? make_int2(g.bip_aff[0] + ug * g.bip_aff[1], g.bip_aff[2] + ug * g.bip_aff[3])
                   : make_int2(g.bip_aff[4] + ug * g.bip_aff[5], g.bip_aff[6] + ug * g.bip_aff[7]);
  } else {
    vinfo = is_row ? g.bip_row[ug] : g.bip_col[ug];
  }
  const int32_t nch = is_row ? nch_row : nch_col;
  float2 cv[kSumLoads];
#pragma unroll
  for (int u = 0; u < kSumLoads; ++u)
    cv[u] = ld2(pin + vinfo.y + 2 * min(half * kSumLoads + u, nch - 1));
  const float2 evv = make_float2(ev[vinfo.x], ev[vinfo.x + 1]);
  // the tile's old messages and raw log-potentials; cells of virtual row 0
  // or column 0 hold a unary factor (its message in ua, table in lq.xy)
  float ua[kPer], ub[kPer];  // encoded old messages
  float4 lq[kPer];
  const int32_t cc = min(col, ncl - 1);
  const int32_t vc = c0 + cc;  // virtual column of this thread
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const int32_t vr = a0 + min(row0 + kStep * k, nr - 1);
    if (vr >= U && vc >= U) {
      const int32_t f = (vr - U) * nb + (vc - U);
      ua[k] = mo[m0 + f];
      ub[k] = mo[m0 + n + f];
      if ((blk.pad[0] & 1) && g.lp_al4) {
        lq[k] = ld4(lp + blk.c0 + 4 * f);
      } else {
        const float* q = lp + blk.c0 + 4 * f;
        lq[k] = make_float4(q[0], q[1], q[2], q[3]);
      }
    } else {
      // unary factor of row unit vr - 1 (column 0) or column unit vc - 1
      // (row 0); cell (0, 0) loads the first one and stays idle
      const int32_t u = vr >= U ? vr - 1 : na + max(vc - 1, 0);
      ua[k] = mo[u];
      ub[k] = 0.f;
      const float* q = lp + g.bip_uc0 + 2 * u;
      lq[k] = make_float4(q[0], q[1], 0.f, 0.f);
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
    const int2 ri = g.bip_row[max(a0 + tid - U, 0)];
    float t0, t1;
    chunk_total2(pin + ri.y, nch_row, t0, t1);
    vrow[tid] = make_float2(t0 + ev[ri.x], t1 + ev[ri.x + 1]);
  } else if (tid >= 64 && tid - 64 < ncl) {
    const int2 ci = g.bip_col[max(c0 + tid - 64 - U, 0)];
    float t0, t1;
    chunk_total2(pin + ci.y, nch_col, t0, t1);
    vcol[tid - 64] = make_float2(t0 + ev[ci.x], t1 + ev[ci.x + 1]);
  }
  __syncthreads();
  float dl = 0.f;
  const float2 vcs = col_ok ? vcol[col] : make_float2(0.f, 0.f);
  auto update = [&](float t) {
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const int32_t ra = row0 + kStep * k;
      if (col_ok && ra < nr) {
        const int32_t vr = a0 + ra, vcl = c0 + col;
        const float2 oa = pair_decode(ua[k]);
        if (vr >= U && vcl >= U) {
          const int32_t f = (vr - U) * nb + (vcl - U);
          const float2 vrs = vrow[ra];
          const float2 ob = pair_decode(ub[k]);
          const float4 lc = make_float4(clip_lp(lq[k].x), clip_lp(lq[k].y), clip_lp(lq[k].z),
                                        clip_lp(lq[k].w));
          float fa0, fa1, fb0, fb1;
          bin2_ftov<kT0>(vrs.x - oa.x, vrs.y - oa.y, vcs.x - ob.x, vcs.y - ob.y, lc, t, fa0, fa1,
                         fb0, fb1);
          float na0, na1, nb0, nb1;
          damp_norm2(r.d, r.omd, oa.x, oa.y, fa0, fa1, na0, na1, dl);
          damp_norm2(r.d, r.omd, ob.x, ob.y, fb0, fb1, nb0, nb1, dl);
          mn[m0 + f] = pair_encode(na0, na1);
          mn[m0 + n + f] = pair_encode(nb0, nb1);
          srow[ra][col] = make_float2(na0, na1);
          scol[ra][col] = make_float2(nb0, nb1);
        } else if (vr >= U || vcl >= U) {
          // a unary factor: of row unit vr - 1 in column 0 (its message
          // enters the row's chunk), of column unit vcl - 1 in row 0
          const bool of_row = vr >= U;
          const float2 vs = of_row ? vrow[ra] : vcs;
          const int32_t u = of_row ? vr - 1 : na + vcl - 1;
          float f0, f1, r0, r1;
          unary2_ftov<kT0>(vs.x - oa.x, vs.y - oa.y, lq[k].x, lq[k].y, t, f0, f1);
          damp_norm2(r.d, r.omd, oa.x, oa.y, f0, f1, r0, r1, dl);
          mn[u] = pair_encode(r0, r1);
          if (of_row)
            srow[ra][col] = make_float2(r0, r1);
          else
            scol[ra][col] = make_float2(r0, r1);
        }
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
    if (ra < nr && len > 0 && a0 + ra >= U) {
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
      st2(pout + g.bip_row[a0 + ra - U].y + 2 * ((c0 + cb) / kVarChunk), s0, s1);
    }
  } else if (tid < 2 * kCols && tid - kCols < ncl && c0 + (tid - kCols) >= U) {
    const int32_t c = tid - kCols;  // column chunks: thread = column
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
    st2(pout + g.bip_col[c0 + c - U].y + 2 * rt, s0, s1);
  }
  block_delta_store<kThreads>(dl, delta_slot(r, b, it));
}

}  // namespace

template <bool kT0, int kCols, int kThreads>
static void launch_bipartite_t(const DevGraph& d, const RunBufs& r, const float* msg_in,
                               float* msg_out, const float* pin, float* pout, int32_t it,
                               int32_t batch, hipStream_t s) {
  const int32_t col_tiles = (d.bip_nb + d.bip_unary + kCols - 1) / kCols;
  const int32_t row_tiles = (d.bip_na + d.bip_unary + kBipRows - 1) / kBipRows;
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
void launch_bipartite(const DevGraph& d, const RunBufs& r, const float* msg_in, float* msg_out,
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

template void launch_bipartite<true>(const DevGraph&, const RunBufs&, const float*, float*,
                                     const float*, float*, int32_t, int32_t, hipStream_t,
                                     int64_t, int64_t);
template void launch_bipartite<false>(const DevGraph&, const RunBufs&, const float*, float*,
                                      const float*, float*, int32_t, int32_t, hipStream_t,
                                      int64_t, int64_t);

}  // namespace pgx
