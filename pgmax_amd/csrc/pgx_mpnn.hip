// Learned warm start: a 3-layer message-passing network over the factor graph
// that predicts initial factor-to-variable messages for graphs of binary
// variables and pairwise factors. No counterpart in the reference; its
// output is what BPArrays.ftov_msgs takes (pgmax/infer/bp_state.py:172-255).
//
// Model (per edge e of factor f and variable v; o(e) = the factor's other
// edge, u = its variable):
//   x_e  = [ev_v, ev_u, J, h_v, h_u, 1, 0, 0]    (log-odds evidence, coupling
//          and fields of the oriented 2x2 log-potential table)
//   h0_e = relu(x_e W0 + b0)                               64 wide
//   l = 1..3:  z_e = [h_e, sum_{e' at v} h_e' - h_e, h_o(e)]   192 wide
//              h_e = relu(z_e W_l + b_l)
//   y_e  = h3_e . w_h + b_h,  message = [min(0, -y_e), min(0, y_e)]
// Activations are stored in BF16; the three 192x64 layers run on MFMA
// (v_mfma_f32_16x16x32_bf16, FP32 accumulate); sums, the input layer, the
// head and the optimiser are FP32. Training minimises the mean squared error
// of y_e against the log-ratio of converged messages.
//
// Row r of every activation array is edge e = r mod E of batch element
// b = r / E. Every reduction runs in a fixed order, so results are
// deterministic for a given launch configuration.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/pgx.h"
#include "pgx_internal.h"
#include "pgx_launch.h"
#include "pgx_math.h"
#include "pgx_plan.h"

namespace pgx {
namespace {

constexpr int kH = 64;          // hidden width
constexpr int kKZ = 3 * kH;     // layer input width
constexpr int kNin = 8;         // input features
constexpr int kLayers = 3;
constexpr int kTileRows = 64;   // rows per MFMA tile
constexpr int kThreads = 256;
constexpr int kOffW0 = 0;
constexpr int kOffB0 = kNin * kH;
constexpr int kOffL = kOffB0 + kH;
constexpr int kLayerSize = kKZ * kH + kH;
constexpr int kOffWh = kOffL + kLayers * kLayerSize;
constexpr int kOffBh = kOffWh + kH;
constexpr int kNumParams = kOffBh + 1;
static_assert(kNumParams == PGX_MPNN_NUM_PARAMS, "parameter layout");
constexpr int kLayerPart = kKZ * kH + kH;  // per-workgroup partial of dW_l, db_l
constexpr int kEmbedPart = kNin * kH + kH;
constexpr int kHeadPart = kH + 1;
constexpr int kMaxLayerBlocks = 1024;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint16_t u16;

#define PGX_HIP(call)                                                              \
  do {                                                                             \
    hipError_t _e = (call);                                                        \
    if (_e != hipSuccess)                                                          \
      return set_error(PGX_EHIP, std::string(#call) + ": " + hipGetErrorString(_e)); \
  } while (0)

typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

// FP32 -> BF16, round to nearest even (NaN stays NaN): v_cvt_pk_bf16_f32,
// which gives the bits of the integer rounding (add 0x7fff plus the kept
// LSB, keep the upper half) for every non-NaN input (pgx_check_numerics,
// which = 0, compares the two over all 2^32 patterns).
__device__ __forceinline__ u16 f2bf(float x) { return __builtin_bit_cast(u16, __bf16(x)); }
// Two values as one word of BF16 (lo in the low half).
__device__ __forceinline__ uint32_t pack_bf16(float lo, float hi) {
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(f32x2{lo, hi}, bf16x2));
}
__device__ __forceinline__ float bf2f(u16 h) { return __uint_as_float(uint32_t(h) << 16); }
__device__ __forceinline__ float ld(const u16* p) { return bf2f(*p); }
__device__ __forceinline__ float ld(const float* p) { return *p; }

// D += A B for a 16x32 A and 32x16 B in BF16. Lane l holds A[l % 16][8(l/16) .. +7],
// B[8(l/16) .. +7][l % 16] and D[4(l/16) .. +3][l % 16].
__device__ __forceinline__ f32x4 mfma(u32x4 a, u32x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                 __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}

struct MpGraph {
  const int32_t* edge_var;  // [E]
  const int32_t* fac_lp0;   // [F] first log-potential of each factor
  const int32_t* var_eptr;  // [NV + 1]
  const int32_t* var_adj;   // [E] edges of each variable, in edge order
  int64_t E, NV;
};

// Input features of edge e of batch element b (oracle/mpnn.py:features
// restates them in FP64).
// ... from the factor's table L[0..3] and the evidence V = (ev_v0, ev_v1,
// ev_u0, ev_u1) of the edge's variable v and the other edge's u; odd: e is
// the factor's second edge (it sees the transposed table).
__device__ __forceinline__ void features_from(const float* L, const float* V, bool odd,
                                              float* x) {
  const float t00 = L[0], t11 = L[3];
  const float t01 = odd ? L[2] : L[1];
  const float t10 = odd ? L[1] : L[2];
  x[0] = V[1] - V[0];
  x[1] = V[3] - V[2];
  x[2] = (((t00 + t11) - t01) - t10) * 0.25f;
  x[3] = (((t10 + t11) - t00) - t01) * 0.5f;
  x[4] = (((t01 + t11) - t00) - t10) * 0.5f;
  x[5] = 1.f;
  x[6] = 0.f;
  x[7] = 0.f;
}
__device__ __forceinline__ void features_be(const MpGraph& g, int64_t b, int64_t e,
                                            const float* lp, int64_t lp_stride, const float* ev,
                                            int64_t ev_stride, float* x) {
  const float* L = lp + b * lp_stride + g.fac_lp0[e >> 1];
  const float* Vb = ev + b * ev_stride;
  const int32_t v = g.edge_var[e], u = g.edge_var[e ^ 1];
  const float Lr[4] = {L[0], L[1], L[2], L[3]};
  const float V[4] = {Vb[2 * v], Vb[2 * v + 1], Vb[2 * u], Vb[2 * u + 1]};
  features_from(Lr, V, (e & 1) != 0, x);
}

// Input features of row r.
__device__ __forceinline__ void features(const MpGraph& g, int64_t r, const float* lp,
                                         int64_t lp_stride, const float* ev, int64_t ev_stride,
                                         float* x) {
  const int64_t b = r / g.E;
  features_be(g, b, r - b * g.E, lp, lp_stride, ev, ev_stride, x);
}

// Head of one row, in the order every head kernel uses: quarter q (columns
// 16q .. 16q + 15) is summed in column order, then y = ((p0 + p1) + p2) + p3
// plus the bias. head_part gives one quarter from 16 BF16 values and their
// 16 weights.
__device__ __forceinline__ float head_part(const u16* h16, const float* w16) {
  float p = 0.f;
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    const u32x4 v = *reinterpret_cast<const u32x4*>(h16 + 8 * c);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      p += bf2f(u16(v[k] & 0xffffu)) * w16[8 * c + 2 * k];
      p += bf2f(u16(v[k] >> 16)) * w16[8 * c + 2 * k + 1];
    }
  }
  return p;
}

// ((p0 + p1) + p2) + p3 over the four lanes of a quad (DPP quad_perm
// broadcasts of lanes 0, 1, 2, 3); every lane of the quad gets the sum.
__device__ __forceinline__ float quad_sum(float x) {
  const int xi = __float_as_int(x);
  const float p0 = __int_as_float(__builtin_amdgcn_mov_dpp(xi, 0x00, 0xf, 0xf, false));
  const float p1 = __int_as_float(__builtin_amdgcn_mov_dpp(xi, 0x55, 0xf, 0xf, false));
  const float p2 = __int_as_float(__builtin_amdgcn_mov_dpp(xi, 0xaa, 0xf, 0xf, false));
  const float p3 = __int_as_float(__builtin_amdgcn_mov_dpp(xi, 0xff, 0xf, 0xf, false));
  return ((p0 + p1) + p2) + p3;
}

// Orders LDS accesses of the lanes of one wavefront: the compiler keeps LDS
// accesses on either side of it, and the wave's LDS instructions run in order.
__device__ __forceinline__ void wave_lds_fence() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Eight consecutive values of an activation row as FP32.
__device__ __forceinline__ void load8(const u16* p, float* x) {
  const u32x4 v = *reinterpret_cast<const u32x4*>(p);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    x[2 * k] = bf2f(u16(v[k] & 0xffffu));
    x[2 * k + 1] = bf2f(u16(v[k] >> 16));
  }
}
__device__ __forceinline__ void load8(const float* p, float* x) {
  const f32x4 a = *reinterpret_cast<const f32x4*>(p);
  const f32x4 c = *reinterpret_cast<const f32x4*>(p + 4);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    x[k] = a[k];
    x[4 + k] = c[k];
  }
}

// W_l in BF16, transposed ([64][192], for the forward) and as is ([192][64]).
__global__ void k_mp_cast(const float* params, u16* wt, u16* wr) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= kLayers * kKZ * kH) return;
  const int l = t / (kKZ * kH), ij = t % (kKZ * kH), i = ij / kH, j = ij % kH;
  const u16 w = f2bf(params[kOffL + l * kLayerSize + ij]);
  wr[t] = w;
  wt[l * kKZ * kH + j * kKZ + i] = w;
}

// h0 = relu(x W0 + b0); thread = edge (all 64 units: the weights, staged in
// LDS, are read at the same address by the whole wavefront), batch elements
// along y, each workgroup walking its share of them (embed_grid). The edge's
// wiring is loaded once; each element's table and evidence are loaded while
// the previous element computes. Features 5..7 are the constants 1, 0, 0:
// the unit adds its weight of feature 5 and skips the other two. The
// workgroup's 256 rows go out through LDS, so each store instruction writes
// 1 KB of consecutive rows. Row p holds edge row_edge[p] (edge p without
// row_edge).
constexpr int kEmbedPitch = kH / 2 + 4;  // LDS row pitch (words) of the output rows

__global__ void __launch_bounds__(kThreads)
    k_mp_embed(MpGraph g, const float* __restrict__ params, const float* __restrict__ lp,
               int64_t lp_stride, const float* __restrict__ ev, int64_t ev_stride, int64_t B,
               const int32_t* __restrict__ row_edge, u16* __restrict__ h0) {
  static_assert(kOffW0 == 0 && kOffB0 == kNin * kH, "W0 then b0 at the front of the parameters");
  __shared__ __attribute__((aligned(16))) float sP[kNin * kH + kH];
  __shared__ __attribute__((aligned(16))) uint32_t sO[kThreads * kEmbedPitch];
  for (int i = threadIdx.x; i < kNin * kH + kH; i += kThreads) sP[i] = params[i];
  __syncthreads();
  for (int64_t e0 = blockIdx.x * int64_t(kThreads); e0 < g.E;
       e0 += int64_t(gridDim.x) * kThreads) {
    const int64_t e = e0 + threadIdx.x;
    const bool ok = e < g.E;
    const int nrows = int(std::min<int64_t>(kThreads, g.E - e0));
    // the edge's wiring is the same for every batch element: loaded once;
    // each element's table and evidence are loaded one element ahead
    int64_t lp0 = 0, v2 = 0, u2 = 0;
    bool odd = false;
    if (ok) {
      const int64_t ee = row_edge ? row_edge[e] : e;
      lp0 = g.fac_lp0[ee >> 1];
      v2 = 2 * int64_t(g.edge_var[ee]);
      u2 = 2 * int64_t(g.edge_var[ee ^ 1]);
      odd = (ee & 1) != 0;
    }
    auto fetch = [&](int64_t b, float* L, float* V) {
      if (!ok || b >= B) return;
      const float* Lb = lp + b * lp_stride + lp0;
      const float* Vb = ev + b * ev_stride;
      L[0] = Lb[0];
      L[1] = Lb[1];
      L[2] = Lb[2];
      L[3] = Lb[3];
      V[0] = Vb[v2];
      V[1] = Vb[v2 + 1];
      V[2] = Vb[u2];
      V[3] = Vb[u2 + 1];
    };
    float Lc[4] = {0.f, 0.f, 0.f, 0.f}, Vc[4] = {0.f, 0.f, 0.f, 0.f};
    fetch(blockIdx.y, Lc, Vc);
    for (int64_t b = blockIdx.y; b < B; b += gridDim.y) {
      float Ln[4] = {0.f, 0.f, 0.f, 0.f}, Vn[4] = {0.f, 0.f, 0.f, 0.f};
      fetch(b + gridDim.y, Ln, Vn);
      if (ok) {
        float x[kNin];
        features_from(Lc, Vc, odd, x);
        uint32_t* o = sO + threadIdx.x * kEmbedPitch;
#pragma unroll 1
        for (int c = 0; c < kH / 8; ++c) {
          u32x4 out;
#pragma unroll
          for (int k2 = 0; k2 < 4; ++k2) {
            const int j = 8 * c + 2 * k2;
            float a0 = sP[kOffB0 + j], a1 = sP[kOffB0 + j + 1];
#pragma unroll
            for (int k = 0; k < 5; ++k) {
              a0 += x[k] * sP[k * kH + j];
              a1 += x[k] * sP[k * kH + j + 1];
            }
            a0 += sP[5 * kH + j];
            a1 += sP[5 * kH + j + 1];
            out[k2] = pack_bf16(fmaxf(a0, 0.f), fmaxf(a1, 0.f));
          }
          *reinterpret_cast<u32x4*>(o + 4 * c) = out;
        }
      }
      __syncthreads();
      u16* dst = h0 + (b * g.E + e0) * kH;
      for (int i = threadIdx.x; i < nrows * (kH / 8); i += kThreads) {
        const int row = i >> 3, c = i & 7;
        *reinterpret_cast<u32x4*>(dst + row * kH + 8 * c) =
            *reinterpret_cast<const u32x4*>(sO + row * kEmbedPitch + 4 * c);
      }
      __syncthreads();
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        Lc[k] = Ln[k];
        Vc[k] = Vn[k];
      }
    }
  }
}

// S[b, v, :] = sum over the variable's edges of h[b, e, :] (in edge order);
// thread = (variable, 8 columns), batch elements along y.
template <typename T>
__global__ void __launch_bounds__(kThreads)
    k_mp_varsum(MpGraph g, const T* h, int64_t B, float* S) {
  for (int64_t t = blockIdx.x * int64_t(kThreads) + threadIdx.x; t < g.NV * 8;
       t += int64_t(gridDim.x) * kThreads) {
    const int64_t v = t >> 3;
    const int j0 = int(t & 7) * 8;
    const int32_t p0 = g.var_eptr[v], p1 = g.var_eptr[v + 1];
    for (int64_t b = blockIdx.y; b < B; b += gridDim.y) {
      const T* hb = h + b * g.E * kH + j0;
      float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
      int32_t p = p0;
      for (; p + 4 <= p1; p += 4) {  // four rows in flight, summed in edge order
        float x[4][8];
#pragma unroll
        for (int u = 0; u < 4; ++u) load8(hb + int64_t(g.var_adj[p + u]) * kH, x[u]);
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
          for (int k = 0; k < 8; ++k) s[k] += x[u][k];
      }
      for (; p < p1; ++p) {
        float x[8];
        load8(hb + int64_t(g.var_adj[p]) * kH, x);
#pragma unroll
        for (int k = 0; k < 8; ++k) s[k] += x[k];
      }
      float* d = S + (b * g.NV + v) * kH + j0;
      *reinterpret_cast<f32x4*>(d) = f32x4{s[0], s[1], s[2], s[3]};
      *reinterpret_cast<f32x4*>(d + 4) = f32x4{s[4], s[5], s[6], s[7]};
    }
  }
}

// Row `row` (< 64) of the layer input z for global row r, 8 values from
// column 8c (c < 24), as BF16.
__device__ __forceinline__ u32x4 z_chunk(const MpGraph& g, const u16* hin, const float* S,
                                         int64_t r, int c) {
  const int part = c >> 3, cc = (c & 7) * 8;
  const int64_t b = r / g.E;
  const int64_t e = r - b * g.E;
  if (part == 0) return *reinterpret_cast<const u32x4*>(hin + r * kH + cc);
  if (part == 2) return *reinterpret_cast<const u32x4*>(hin + (b * g.E + (e ^ 1)) * kH + cc);
  const u32x4 hh = *reinterpret_cast<const u32x4*>(hin + r * kH + cc);
  const float* s = S + (b * g.NV + g.edge_var[e]) * kH + cc;
  u32x4 out;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    out[k] = pack_bf16(s[2 * k] - bf2f(u16(hh[k] & 0xffffu)),
                       s[2 * k + 1] - bf2f(u16(hh[k] >> 16)));
  }
  return out;
}

constexpr int kZPad = kKZ + 8;  // LDS row pitch (elements) of 192-wide tiles
constexpr int kHPad = kH + 8;   // LDS row pitch of 64-wide tiles

// h_out = relu(z W_l + b_l). Every wavefront works alone on tiles of 16 rows
// (persistent, tiles strided over all wavefronts of the grid): lane l gathers
// columns 16(l % 4) .. +15 of row l / 4 (its activation, its variable's sum)
// into the wave's own LDS tile, writes its activation a second time as the
// factor part of row l / 4 ^ 1 (the factor's other edge, in the same tile as
// E is even), and issues the next tile's gathers before the MFMAs of this one.
// With kHead the head runs on the tile too: y = h_out . w_h + b_h gives the
// messages [min(0, -y), min(0, y)] and h_out is not stored.
constexpr int kWaveRows = 16;
constexpr int kFwdWaves = 8;  // wavefronts per workgroup (LDS: 76.8 KB)

struct FwdRow {
  u32x4 h0, h1;  // activation columns 16q .. +15
  f32x4 s[4];    // variable-sum columns 16q .. +15
};

template <bool kHead>
__global__ void __launch_bounds__(kFwdWaves * 64)
    k_mp_layer_fwd(MpGraph g, const u16* hin, const float* S, const u16* wt,
                   const float* bias, u16* hout, int64_t R, const float* params, float* msgs,
                   int64_t msgs_stride) {
  __shared__ __attribute__((aligned(16))) u16 sW[kH * kZPad];
  __shared__ __attribute__((aligned(16))) u16 sZ[kFwdWaves * kWaveRows * kZPad];
  for (int q = threadIdx.x; q < kH * (kKZ / 8); q += kFwdWaves * 64) {
    const int n = q / (kKZ / 8), c = q % (kKZ / 8);
    *reinterpret_cast<u32x4*>(&sW[n * kZPad + c * 8]) =
        *reinterpret_cast<const u32x4*>(wt + n * kKZ + c * 8);
  }
  __syncthreads();
  const int w = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
  const int l = threadIdx.x & 63, row = l >> 2, q = l & 3;
  u16* z = sZ + w * (kWaveRows * kZPad);
  float bcol[4];
#pragma unroll
  for (int n = 0; n < 4; ++n) bcol[n] = bias[16 * n + (l & 15)];
  float whq[16];  // head weights of this lane's 16 columns
#pragma unroll
  for (int k = 0; k < 16; ++k) whq[k] = kHead ? params[kOffWh + 16 * q + k] : 0.f;
  const int64_t ntiles = (R + kWaveRows - 1) / kWaveRows;
  const int64_t stride = int64_t(gridDim.x) * kFwdWaves;
  // this lane's row of tile t: global row, batch element and edge; false past R
  auto locate = [&](int64_t t, int64_t* r, int64_t* b, int64_t* e) -> bool {
    *r = t * kWaveRows + row;
    if (t >= ntiles || *r >= R) return false;
    const int64_t r0 = t * kWaveRows;
    int64_t bb = r0 / g.E, ee = r0 - bb * g.E + row;
    while (ee >= g.E) {
      ee -= g.E;
      ++bb;
    }
    *b = bb;
    *e = ee;
    return true;
  };
  auto gather = [&](bool ok, int64_t r, int64_t b, int32_t v, FwdRow* x) {
    if (ok) {
      const u32x4* hp = reinterpret_cast<const u32x4*>(hin + r * kH + 16 * q);
      x->h0 = hp[0];
      x->h1 = hp[1];
      const f32x4* sp = reinterpret_cast<const f32x4*>(S + (b * g.NV + v) * kH + 16 * q);
#pragma unroll
      for (int k = 0; k < 4; ++k) x->s[k] = sp[k];
    } else {
      x->h0 = x->h1 = u32x4{0u, 0u, 0u, 0u};
#pragma unroll
      for (int k = 0; k < 4; ++k) x->s[k] = f32x4{0.f, 0.f, 0.f, 0.f};
    }
  };
  int64_t tc = int64_t(blockIdx.x) * kFwdWaves + w;
  int64_t rc = 0, bcur = 0, ec = 0;
  bool okc = locate(tc, &rc, &bcur, &ec);
  FwdRow cur;
  gather(okc, rc, bcur, okc ? g.edge_var[ec] : 0, &cur);
  int64_t rn = 0, bn = 0, en = 0;
  bool okn = locate(tc + stride, &rn, &bn, &en);
  int32_t vn = okn ? g.edge_var[en] : 0;
  for (; tc < ntiles; tc += stride) {
    // this tile's input rows into LDS
    {
      u16* zr = z + row * kZPad + 16 * q;
      u16* zo = z + (row ^ 1) * kZPad + 2 * kH + 16 * q;
      *reinterpret_cast<u32x4*>(zr) = cur.h0;
      *reinterpret_cast<u32x4*>(zr + 8) = cur.h1;
      *reinterpret_cast<u32x4*>(zo) = cur.h0;
      *reinterpret_cast<u32x4*>(zo + 8) = cur.h1;
      u32x4 d[2];
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const uint32_t hw = k < 4 ? cur.h0[k] : cur.h1[k - 4];
        const float s0 = cur.s[k >> 1][(2 * k) & 3], s1 = cur.s[k >> 1][(2 * k + 1) & 3];
        d[k >> 2][k & 3] = pack_bf16(s0 - bf2f(u16(hw & 0xffffu)), s1 - bf2f(u16(hw >> 16)));
      }
      *reinterpret_cast<u32x4*>(zr + kH) = d[0];
      *reinterpret_cast<u32x4*>(zr + kH + 8) = d[1];
    }
    const int64_t r_out = rc;
    const bool ok_out = okc;
    const int64_t b_out = bcur, e_out = ec;
    // next tile's gathers, and the edge of the one after
    okc = okn;
    rc = rn;
    bcur = bn;
    ec = en;
    gather(okc, rc, bcur, vn, &cur);
    okn = locate(tc + 2 * stride, &rn, &bn, &en);
    vn = okn ? g.edge_var[en] : 0;
    wave_lds_fence();
    f32x4 acc[4];
#pragma unroll
    for (int n = 0; n < 4; ++n) acc[n] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < kKZ / 32; ++ks) {
      const int kk = ks * 32 + 8 * (l >> 4);
      const u32x4 a = *reinterpret_cast<const u32x4*>(&z[(l & 15) * kZPad + kk]);
#pragma unroll
      for (int n = 0; n < 4; ++n) {
        const u32x4 bb = *reinterpret_cast<const u32x4*>(&sW[(16 * n + (l & 15)) * kZPad + kk]);
        acc[n] = mfma(a, bb, acc[n]);
      }
    }
    wave_lds_fence();
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      const int col = 16 * n + (l & 15);
#pragma unroll
      for (int i = 0; i < 4; ++i)
        z[(4 * (l >> 4) + i) * kZPad + col] = f2bf(fmaxf(acc[n][i] + bcol[n], 0.f));
    }
    wave_lds_fence();
    if (kHead) {
      const float y = quad_sum(head_part(z + row * kZPad + 16 * q, whq)) + params[kOffBh];
      if (ok_out && q < 2) msgs[b_out * msgs_stride + 2 * e_out + q] = fminf(0.f, q ? y : -y);
    } else if (ok_out) {
      const u16* zr = z + row * kZPad + 16 * q;
      u32x4* dst = reinterpret_cast<u32x4*>(hout + r_out * kH + 16 * q);
      dst[0] = *reinterpret_cast<const u32x4*>(zr);
      dst[1] = *reinterpret_cast<const u32x4*>(zr + 8);
    }
    wave_lds_fence();
  }
}

// Prediction in variable-major row order: row p of a batch element is edge
// var_adj[p], so the edges of each variable are consecutive rows. Tiles of at
// most 16 rows hold whole variables (variables of degree <= 16), and each
// variable's sum over its edges is taken from the tile itself, in edge order,
// so there is no separate sum pass and no sum array. Per row: the row of the
// factor's other edge (gathered), the edge (for the messages) and the
// variable's first row in the tile and degree (seg = first | degree << 8).
// Per tile: its first row, its row count and the rows that end a variable
// (rows | ends << 16). Otherwise as k_mp_layer_fwd: one wavefront per tile,
// the next tiles' gathers in flight during this one's MFMAs.
constexpr int kVmMaxDegree = kWaveRows;

struct VmRow {
  u32x4 h0, h1;  // activation columns 16q .. +15
  u32x4 o0, o1;  // the factor's other edge, same columns
};

template <bool kHead>
__global__ void __launch_bounds__(kFwdWaves * 64)
    k_mp_layer_vm(const int4* __restrict__ rows, const int2* __restrict__ tiles, int64_t T,
                  int64_t E, const u16* hin, const u16* wt, const float* bias, u16* hout,
                  int64_t B, const float* params, float* msgs, int64_t msgs_stride) {
  __shared__ __attribute__((aligned(16))) u16 sW[kH * kZPad];
  __shared__ __attribute__((aligned(16))) u16 sZ[kFwdWaves * kWaveRows * kZPad];
  for (int q = threadIdx.x; q < kH * (kKZ / 8); q += kFwdWaves * 64) {
    const int n = q / (kKZ / 8), c = q % (kKZ / 8);
    *reinterpret_cast<u32x4*>(&sW[n * kZPad + c * 8]) =
        *reinterpret_cast<const u32x4*>(wt + n * kKZ + c * 8);
  }
  __syncthreads();
  const int w = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
  const int l = threadIdx.x & 63, row = l >> 2, q = l & 3;
  u16* z = sZ + w * (kWaveRows * kZPad);
  float bcol[4];
#pragma unroll
  for (int n = 0; n < 4; ++n) bcol[n] = bias[16 * n + (l & 15)];
  float whq[16];  // head weights of this lane's 16 columns
#pragma unroll
  for (int k = 0; k < 16; ++k) whq[k] = kHead ? params[kOffWh + 16 * q + k] : 0.f;
  const int64_t ntiles = B * T;
  const int64_t stride = int64_t(gridDim.x) * kFwdWaves;
  struct Loc {
    bool ok;
    int64_t b;      // batch element
    int32_t p;      // row within the batch element
    int4 info;      // other row, edge, seg
    uint32_t ends;  // the tile's rows that end a variable (bit r: row r)
  };
  auto locate = [&](int64_t t) -> Loc {
    Loc c{false, 0, 0, make_int4(0, 0, 0, 0), 0u};
    if (t >= ntiles) return c;
    const int64_t b = t / T;
    const int2 d = tiles[t - b * T];
    c.ends = uint32_t(d.y) >> 16;
    if (row >= (d.y & 0xffff)) return c;
    c.ok = true;
    c.b = b;
    c.p = d.x + row;
    c.info = rows[c.p];
    return c;
  };
  auto gather = [&](const Loc& c, VmRow* x) {
    if (c.ok) {
      const u16* hb = hin + c.b * E * kH + 16 * q;
      const u32x4* hp = reinterpret_cast<const u32x4*>(hb + int64_t(c.p) * kH);
      const u32x4* op = reinterpret_cast<const u32x4*>(hb + int64_t(c.info.x) * kH);
      x->h0 = hp[0];
      x->h1 = hp[1];
      x->o0 = op[0];
      x->o1 = op[1];
    } else {
      x->h0 = x->h1 = x->o0 = x->o1 = u32x4{0u, 0u, 0u, 0u};
    }
  };
  // Workgroups go round-robin over the 8 XCDs; numbering them XCD by XCD
  // gives each XCD a run of consecutive tiles, so the rows of a variable's
  // neighbours (the factors' other edges, mostly a few tiles away) come
  // through the same L2.
  int wg = blockIdx.x;
  if ((gridDim.x & 7) == 0) wg = (blockIdx.x & 7) * (gridDim.x >> 3) + (blockIdx.x >> 3);
  u16* zr = z + row * kZPad + 16 * q;
  // this tile's gathered rows into LDS
  auto stage = [&](const VmRow& x) {
    *reinterpret_cast<u32x4*>(zr) = x.h0;
    *reinterpret_cast<u32x4*>(zr + 8) = x.h1;
    *reinterpret_cast<u32x4*>(zr + 2 * kH) = x.o0;
    *reinterpret_cast<u32x4*>(zr + 2 * kH + 8) = x.o1;
  };
  // the rest of the tile whose rows are lo: variable part, MFMAs, output
  auto finish = [&](const Loc& lo) {
    wave_lds_fence();
    {
      // Each variable's sum over its rows (edge order, from 0) minus each
      // row's own value. Lane l takes column l of all 16 rows: running sums
      // restart after every row that ends a variable (bit r of lo.ends),
      // and a backward sweep hands each row its variable's total.
      float h[kWaveRows], pre[kWaveRows];
#pragma unroll
      for (int r = 0; r < kWaveRows; ++r) h[r] = bf2f(z[r * kZPad + l]);
      float run = 0.f;
#pragma unroll
      for (int r = 0; r < kWaveRows; ++r) {
        run = ((r > 0 && ((lo.ends >> (r - 1)) & 1u)) ? 0.f : run) + h[r];
        pre[r] = run;
      }
      float tot = 0.f;
#pragma unroll
      for (int r = kWaveRows - 1; r >= 0; --r) {
        tot = ((lo.ends >> r) & 1u) ? pre[r] : tot;
        z[r * kZPad + kH + l] = f2bf(tot - h[r]);
      }
    }
    wave_lds_fence();
    f32x4 acc[4];
#pragma unroll
    for (int n = 0; n < 4; ++n) acc[n] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < kKZ / 32; ++ks) {
      const int kk = ks * 32 + 8 * (l >> 4);
      const u32x4 a = *reinterpret_cast<const u32x4*>(&z[(l & 15) * kZPad + kk]);
#pragma unroll
      for (int n = 0; n < 4; ++n) {
        const u32x4 bb = *reinterpret_cast<const u32x4*>(&sW[(16 * n + (l & 15)) * kZPad + kk]);
        acc[n] = mfma(a, bb, acc[n]);
      }
    }
    wave_lds_fence();
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      const int col = 16 * n + (l & 15);
#pragma unroll
      for (int i = 0; i < 4; ++i)
        z[(4 * (l >> 4) + i) * kZPad + col] = f2bf(fmaxf(acc[n][i] + bcol[n], 0.f));
    }
    wave_lds_fence();
    if (kHead) {
      const float y = quad_sum(head_part(zr, whq)) + params[kOffBh];
      if (lo.ok && q < 2)
        msgs[lo.b * msgs_stride + 2 * int64_t(lo.info.y) + q] = fminf(0.f, q ? y : -y);
    } else if (lo.ok) {
      u32x4* dst = reinterpret_cast<u32x4*>(hout + (lo.b * E + lo.p) * kH + 16 * q);
      dst[0] = *reinterpret_cast<const u32x4*>(zr);
      dst[1] = *reinterpret_cast<const u32x4*>(zr + 8);
    }
    wave_lds_fence();
  };
  // Two register sets take turns, so two tiles' gathers are in flight while
  // one tile computes; the rows of the tile after those are located a step
  // earlier still.
  int64_t tc = int64_t(wg) * kFwdWaves + w;
  Loc la = locate(tc), lb = locate(tc + stride);
  VmRow ra, rb;
  gather(la, &ra);
  gather(lb, &rb);
  Loc ln = locate(tc + 2 * stride);
  while (tc < ntiles) {
    stage(ra);
    const Loc lo_a = la;
    la = ln;
    gather(la, &ra);  // tile tc + 2 stride
    ln = locate(tc + 3 * stride);
    finish(lo_a);
    tc += stride;
    if (tc >= ntiles) break;
    stage(rb);
    const Loc lo_b = lb;
    lb = ln;
    gather(lb, &rb);  // tile tc + 2 stride
    ln = locate(tc + 3 * stride);
    finish(lo_b);
    tc += stride;
  }
}

// y = h3 . w_h + b_h; optional messages, and with a target the loss
// partial (per workgroup: sum of (y - y*)^2 / n) and dy = 2 (y - y*) / n.
__global__ void __launch_bounds__(kThreads)
    k_mp_head(const u16* h3, const float* params, int64_t R, int64_t E, float* msgs,
              int64_t msgs_stride, const float* target, int64_t target_stride, float* dy,
              float* loss_part, float inv_n) {
  __shared__ float red[kThreads];
  float err2 = 0.f;
  for (int64_t r = blockIdx.x * int64_t(kThreads) + threadIdx.x; r < R;
       r += int64_t(gridDim.x) * kThreads) {
    const u16* hr = h3 + r * kH;
    float p[4];
#pragma unroll
    for (int qq = 0; qq < 4; ++qq) p[qq] = head_part(hr + 16 * qq, params + kOffWh + 16 * qq);
    const float y = (((p[0] + p[1]) + p[2]) + p[3]) + params[kOffBh];
    const int64_t b = r / E, e = r - b * E;
    if (msgs) {
      float* m = msgs + b * msgs_stride + 2 * e;
      m[0] = fminf(0.f, -y);
      m[1] = fminf(0.f, y);
    }
    if (target) {
      const float* t = target + b * target_stride + 2 * e;
      const float err = y - (t[1] - t[0]);
      dy[r] = 2.f * err * inv_n;
      err2 += err * err;
    }
  }
  if (!target) return;
  red[threadIdx.x] = err2;
  __syncthreads();
  for (int s = kThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss_part[blockIdx.x] = red[0] * inv_n;
}

// g3 = dy w_h^T and per-workgroup partials of (dw_h, db_h).
__global__ void __launch_bounds__(kThreads)
    k_mp_head_bwd(const u16* h3, const float* params, const float* dy, int64_t R, float* g,
                  float* part) {
  __shared__ float red[kThreads];
  __shared__ float redb[4];
  const int rg = threadIdx.x >> 6, j = threadIdx.x & 63;
  const float wh = params[kOffWh + j];
  float aw = 0.f, ab = 0.f;
  for (int64_t r = blockIdx.x * 4 + rg; r < R; r += int64_t(gridDim.x) * 4) {
    const float d = dy[r];
    g[r * kH + j] = d * wh;
    aw += d * bf2f(h3[r * kH + j]);
    ab += d;
  }
  red[threadIdx.x] = aw;
  if (j == 0) redb[rg] = ab;
  __syncthreads();
  if (rg == 0) {
    float s = red[j];
    for (int q = 1; q < 4; ++q) s += red[q * 64 + j];
    part[blockIdx.x * kHeadPart + j] = s;
    if (j == 0) part[blockIdx.x * kHeadPart + kH] = ((redb[0] + redb[1]) + redb[2]) + redb[3];
  }
}

// Backward of one layer over tiles of 64 rows:
//   dpre = g_out * [h_out > 0]
//   dz = dpre W_l^T -> gin (self part), dV (variable-sum part), dF (factor part)
//   dW_l += z^T dpre, db_l += sum(dpre)   (per-workgroup partials)
__global__ void __launch_bounds__(kThreads)
    k_mp_layer_bwd(MpGraph g, const u16* hin, const float* S, const u16* hout,
                   const float* gout, const u16* wr, float* gin, float* dV, float* dF,
                   float* part, int64_t R) {
  __shared__ __attribute__((aligned(16))) u16 sW[kKZ * kHPad];       // W[i][j]
  __shared__ __attribute__((aligned(16))) u16 sZt[kKZ * kHPad];      // z^T[i][row]
  __shared__ __attribute__((aligned(16))) u16 sD[kTileRows * kHPad];  // dpre[row][j]
  __shared__ __attribute__((aligned(16))) u16 sDt[kH * kHPad];        // dpre^T[j][row]
  for (int q = threadIdx.x; q < kKZ * (kH / 8); q += kThreads) {
    const int i = q / (kH / 8), c = q % (kH / 8);
    *reinterpret_cast<u32x4*>(&sW[i * kHPad + c * 8]) =
        *reinterpret_cast<const u32x4*>(wr + i * kH + c * 8);
  }
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  f32x4 accW[12];
#pragma unroll
  for (int q = 0; q < 12; ++q) accW[q] = f32x4{0.f, 0.f, 0.f, 0.f};
  float db = 0.f;
  const int64_t ntiles = (R + kTileRows - 1) / kTileRows;
  for (int64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int64_t r0 = tile * kTileRows;
    __syncthreads();
    for (int q = threadIdx.x; q < kTileRows * (kKZ / 8); q += kThreads) {
      const int row = q % kTileRows, c = q / kTileRows;
      const int64_t r = r0 + row;
      u32x4 val = {0u, 0u, 0u, 0u};
      if (r < R) val = z_chunk(g, hin, S, r, c);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        sZt[(c * 8 + 2 * k) * kHPad + row] = u16(val[k] & 0xffffu);
        sZt[(c * 8 + 2 * k + 1) * kHPad + row] = u16(val[k] >> 16);
      }
    }
    for (int q = threadIdx.x; q < kTileRows * kH; q += kThreads) {
      const int row = q >> 6, j = q & 63;
      const int64_t r = r0 + row;
      u16 d = 0;
      if (r < R && bf2f(hout[r * kH + j]) > 0.f) d = f2bf(gout[r * kH + j]);
      sD[row * kHPad + j] = d;
      sDt[j * kHPad + row] = d;
    }
    __syncthreads();
    // dz = dpre W^T: this wave's 16 rows x 192 columns
    f32x4 acc[12];
#pragma unroll
    for (int nb = 0; nb < 12; ++nb) acc[nb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < kH / 32; ++ks) {
      const int kk = ks * 32 + 8 * (l >> 4);
      const u32x4 a = *reinterpret_cast<const u32x4*>(&sD[(16 * w + (l & 15)) * kHPad + kk]);
#pragma unroll
      for (int nb = 0; nb < 12; ++nb) {
        const u32x4 bb = *reinterpret_cast<const u32x4*>(&sW[(16 * nb + (l & 15)) * kHPad + kk]);
        acc[nb] = mfma(a, bb, acc[nb]);
      }
    }
#pragma unroll
    for (int nb = 0; nb < 12; ++nb) {
      const int col = 16 * nb + (l & 15);
      float* dst = col < kH ? gin : (col < 2 * kH ? dV : dF);
      const int cj = col & 63;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int64_t r = r0 + 16 * w + 4 * (l >> 4) + i;
        if (r < R) dst[r * kH + cj] = acc[nb][i];
      }
    }
    // dW += z^T dpre: this wave's 48 rows of W x 64 columns
#pragma unroll
    for (int ks = 0; ks < kTileRows / 32; ++ks) {
      const int kk = ks * 32 + 8 * (l >> 4);
#pragma unroll
      for (int ib = 0; ib < 3; ++ib) {
        const u32x4 a =
            *reinterpret_cast<const u32x4*>(&sZt[(48 * w + 16 * ib + (l & 15)) * kHPad + kk]);
#pragma unroll
        for (int jb = 0; jb < 4; ++jb) {
          const u32x4 bb = *reinterpret_cast<const u32x4*>(&sDt[(16 * jb + (l & 15)) * kHPad + kk]);
          accW[ib * 4 + jb] = mfma(a, bb, accW[ib * 4 + jb]);
        }
      }
    }
    if (threadIdx.x < kH)
      for (int row = 0; row < kTileRows; ++row) db += bf2f(sD[row * kHPad + threadIdx.x]);
  }
  float* p = part + int64_t(blockIdx.x) * kLayerPart;
#pragma unroll
  for (int ib = 0; ib < 3; ++ib)
#pragma unroll
    for (int jb = 0; jb < 4; ++jb)
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int wi = 48 * w + 16 * ib + 4 * (l >> 4) + i, wj = 16 * jb + (l & 15);
        p[wi * kH + wj] = accW[ib * 4 + jb][i];
      }
  if (threadIdx.x < kH) p[kKZ * kH + threadIdx.x] = db;
}

// gin[r] = d_self[r] + T[v(r)] - dV[r] + dF[o(r)]   (gin holds d_self on entry)
__global__ void __launch_bounds__(kThreads)
    k_mp_combine(MpGraph g, float* gin, const float* T, const float* dV, const float* dF,
                 int64_t R) {
  for (int64_t t = blockIdx.x * int64_t(kThreads) + threadIdx.x; t < R * kH;
       t += int64_t(gridDim.x) * kThreads) {
    const int64_t r = t >> 6;
    const int j = int(t & 63);
    const int64_t b = r / g.E, e = r - b * g.E;
    const float tv = T[(b * g.NV + g.edge_var[e]) * kH + j];
    gin[r * kH + j] =
        ((gin[r * kH + j] + tv) - dV[r * kH + j]) + dF[(b * g.E + (e ^ 1)) * kH + j];
  }
}

// dW0, db0 partials of the input layer (features recomputed).
__global__ void __launch_bounds__(kThreads)
    k_mp_embed_bwd(MpGraph g, const float* lp, int64_t lp_stride, const float* ev,
                   int64_t ev_stride, const u16* h0, const float* g0, int64_t R, float* part) {
  __shared__ float red[kThreads * (kNin + 1)];
  const int rg = threadIdx.x >> 6, j = threadIdx.x & 63;
  float aw[kNin], ab = 0.f;
#pragma unroll
  for (int k = 0; k < kNin; ++k) aw[k] = 0.f;
  for (int64_t r = blockIdx.x * 4 + rg; r < R; r += int64_t(gridDim.x) * 4) {
    const float dp = bf2f(h0[r * kH + j]) > 0.f ? g0[r * kH + j] : 0.f;
    float x[kNin];
    features(g, r, lp, lp_stride, ev, ev_stride, x);
#pragma unroll
    for (int k = 0; k < kNin; ++k) aw[k] += x[k] * dp;
    ab += dp;
  }
#pragma unroll
  for (int k = 0; k < kNin; ++k) red[k * kThreads + threadIdx.x] = aw[k];
  red[kNin * kThreads + threadIdx.x] = ab;
  __syncthreads();
  if (rg == 0) {
    float* p = part + int64_t(blockIdx.x) * kEmbedPart;
    for (int k = 0; k <= kNin; ++k) {
      const float* q = red + k * kThreads;
      const float s = ((q[j] + q[64 + j]) + q[128 + j]) + q[192 + j];
      p[(k < kNin ? k * kH : kNin * kH) + j] = s;
    }
  }
}

// out[i] = sum over workgroups b (in order) of part[b * stride + i].
__global__ void k_mp_reduce(const float* part, int32_t nparts, int32_t stride, int32_t count,
                            float* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  float s = 0.f;
  for (int32_t b = 0; b < nparts; ++b) s += part[int64_t(b) * stride + i];
  out[i] = s;
}

// Adam (Kingma & Ba 2015, Algorithm 1).
__global__ void k_mp_adam(float* p, const float* g, float* m, float* v, int64_t n, float lr,
                          float b1, float b2, float eps, float bc1, float bc2) {
  const int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

// pgx_check_numerics: one bit pattern per thread.
__global__ void k_check_numerics(int32_t which, uint32_t first, uint32_t count,
                                 unsigned long long* bad) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const uint32_t u = first + i;
  bool miss = false;
  if (which == 0) {
#pragma unroll
    for (int sgn = 0; sgn < 2; ++sgn) {
      const uint32_t v = sgn ? (u ^ 0x80000000u) : u;
      if ((v & 0x7f800000u) == 0x7f800000u && (v & 0x7fffffu)) continue;  // NaN
      const uint32_t ref = (v + 0x7fffu + ((v >> 16) & 1u)) >> 16;
      miss |= uint32_t(f2bf(__uint_as_float(v))) != ref;
    }
  } else {
    const float f = __uint_as_float(u) - 1.f;
    miss = __float_as_uint(div_2pf(f)) != __float_as_uint(f / (2.f + f));
  }
  if (miss) atomicAdd(bad, 1ull);
}

template <typename T>
int upload(const std::vector<T>& v, T** dst, std::vector<void*>* owned) {
  *dst = nullptr;
  if (v.empty()) return PGX_OK;
  PGX_HIP(hipMalloc(reinterpret_cast<void**>(dst), v.size() * sizeof(T)));
  owned->push_back(*dst);
  PGX_HIP(hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return PGX_OK;
}

}  // namespace
}  // namespace pgx

using namespace pgx;

struct pgx_mpnn {
  MpGraph g{};
  // variable-major prediction (null when a variable has more than 16 edges):
  // per row (other row, edge, seg), per tile (first row, rows | ends << 16)
  const int4* vm_rows = nullptr;
  const int2* vm_tiles = nullptr;
  int64_t vm_T = 0;
  std::vector<void*> owned;
  // scratch, grown on demand
  char* buf = nullptr;
  size_t buf_bytes = 0;
  int64_t F = 0;
};

namespace {

int grow(pgx_mpnn* n, size_t bytes) {
  if (bytes <= n->buf_bytes) return PGX_OK;
  if (n->buf) PGX_HIP(hipFree(n->buf));
  n->buf = nullptr;
  n->buf_bytes = 0;
  if (hipMalloc(reinterpret_cast<void**>(&n->buf), bytes) != hipSuccess) {
    (void)hipGetLastError();
    return set_error(PGX_ENOMEM, "pgx_mpnn: cannot allocate " + std::to_string(bytes) + " bytes");
  }
  n->buf_bytes = bytes;
  return PGX_OK;
}

struct Carve {
  char* p;
  template <typename T>
  T* take(size_t count) {
    T* r = reinterpret_cast<T*>(p);
    p += (count * sizeof(T) + 255) & ~size_t(255);
    return r;
  }
};

inline size_t al256(size_t b) { return (b + 255) & ~size_t(255); }

// Grid for n items of the grid-stride kernels (at most 65,536 workgroups).
inline unsigned blocks_for(int64_t n) {
  return unsigned(std::max<int64_t>(1, std::min<int64_t>((n + kThreads - 1) / kThreads, 65536)));
}

int check_args(pgx_mpnn* n, const float* params, const float* lp, const float* ev, int32_t batch) {
  if (!n || !params || !lp || !ev || batch <= 0)
    return set_error(PGX_EINVAL, "pgx_mpnn: null argument or batch <= 0");
  const int64_t rows = int64_t(batch) * n->g.E;
  if (rows * kH >= (int64_t(1) << 40)) return set_error(PGX_EINVAL, "pgx_mpnn: batch too large");
  return PGX_OK;
}

// Grid of the kernels over (n items, B batch elements): x covers the items,
// y the batch elements (both loop past the hardware limits).
inline dim3 item_batch_grid(int64_t n, int64_t B) {
  return dim3(blocks_for(n), unsigned(std::max<int64_t>(1, std::min<int64_t>(B, 65535))));
}

// Grid of k_mp_embed: as item_batch_grid, but each workgroup walks up to 16
// batch elements (their loads one element ahead) once there are enough
// workgroups to fill the chip.
inline dim3 embed_grid(int64_t E, int64_t B) {
  const int64_t xb = std::max<int64_t>(1, std::min<int64_t>((E + kThreads - 1) / kThreads, 65536));
  const int64_t walk16 = (B + 15) / 16;  // rows of workgroups when each walks 16 elements
  const int64_t yb = xb * walk16 >= 4096 ? walk16 : B;
  return dim3(unsigned(xb), unsigned(std::max<int64_t>(1, std::min<int64_t>(yb, 65535))));
}

// Workgroups of k_mp_layer_fwd: two per CU of 256 fit its LDS, every
// wavefront strides over the tiles.
inline unsigned fwd_blocks(int64_t R) {
  const int64_t tiles = (R + kWaveRows - 1) / kWaveRows;
  const int64_t groups = (tiles + kFwdWaves - 1) / kFwdWaves;
  return unsigned(std::max<int64_t>(1, std::min<int64_t>(groups, 512)));
}

// Forward pass; h[0..3] and S[0..2] receive the activations and variable sums.
// With msgs the last layer also runs the head into msgs (stride msgs_stride)
// and h[3] is not written.
void forward(pgx_mpnn* n, const float* params, const float* lp, int64_t lp_stride,
             const float* ev, int64_t ev_stride, int64_t B, u16* wt, u16** h, float** S,
             hipStream_t s, float* msgs = nullptr, int64_t msgs_stride = 0) {
  const MpGraph& g = n->g;
  const int64_t R = B * g.E;
  hipLaunchKernelGGL(k_mp_embed, embed_grid(g.E, B), dim3(kThreads), 0, s, g, params, lp,
                     lp_stride, ev, ev_stride, B, static_cast<const int32_t*>(nullptr), h[0]);
  const unsigned lb = fwd_blocks(R);
  for (int l = 0; l < kLayers; ++l) {
    hipLaunchKernelGGL(k_mp_varsum<u16>, item_batch_grid(g.NV * 8, B), dim3(kThreads), 0, s, g,
                       static_cast<const u16*>(h[l]), B, S[l]);
    const float* bias = params + kOffL + l * kLayerSize + kKZ * kH;
    if (msgs && l == kLayers - 1)
      hipLaunchKernelGGL(k_mp_layer_fwd<true>, dim3(lb), dim3(kFwdWaves * 64), 0, s, g,
                         static_cast<const u16*>(h[l]), static_cast<const float*>(S[l]),
                         static_cast<const u16*>(wt + l * kKZ * kH), bias,
                         static_cast<u16*>(nullptr), R, params, msgs, msgs_stride);
    else
      hipLaunchKernelGGL(k_mp_layer_fwd<false>, dim3(lb), dim3(kFwdWaves * 64), 0, s, g,
                         static_cast<const u16*>(h[l]), static_cast<const float*>(S[l]),
                         static_cast<const u16*>(wt + l * kKZ * kH), bias, h[l + 1], R, params,
                         static_cast<float*>(nullptr), int64_t(0));
  }
}

}  // namespace

extern "C" {

int pgx_mpnn_create(const pgx_graph* graph, pgx_mpnn** out) {
  if (!graph || !out) return set_error(PGX_EINVAL, "pgx_mpnn_create: null argument");
  *out = nullptr;
  const Plan& P = graph_plan(graph);
  const int64_t E = P.num_edges, NV = P.num_vars, F = P.num_factors;
  if (E == 0) return set_error(PGX_EUNSUP, "pgx_mpnn_create: graph has no edges");
  if (P.num_es != 2 * E || P.num_vs != 2 * NV)
    return set_error(PGX_EUNSUP, "pgx_mpnn_create: every variable must be binary");
  std::vector<int32_t> fac_lp0(F, -1);
  for (const Block& b : P.blocks) {
    if (b.kind != kEnumBin2)
      return set_error(PGX_EUNSUP, "pgx_mpnn_create: every factor must be a pairwise Enum factor "
                                   "with its full 2x2 table");
    for (int32_t i = 0; i < b.count; ++i) fac_lp0[b.f0 + i] = b.c0 + 4 * i;
  }
  for (int64_t f = 0; f < F; ++f)
    if (fac_lp0[f] < 0 || P.fac_e0[f] != 2 * f)
      return set_error(PGX_EUNSUP, "pgx_mpnn_create: factor " + std::to_string(f) +
                                       " is not a pairwise binary factor");
  std::vector<int32_t> var_adj(E);
  for (int64_t p = 0; p < E; ++p) var_adj[p] = P.var_adj_es_ref[p] / 2;
  pgx_mpnn* n = new pgx_mpnn();
  int32_t *edge_var, *flp, *eptr, *adj;
  int rc;
  if ((rc = upload(P.edge_var, &edge_var, &n->owned)) || (rc = upload(fac_lp0, &flp, &n->owned)) ||
      (rc = upload(P.var_eptr, &eptr, &n->owned)) || (rc = upload(var_adj, &adj, &n->owned))) {
    pgx_mpnn_destroy(n);
    return rc;
  }
  n->g = MpGraph{edge_var, flp, eptr, adj, E, NV};
  n->F = F;
  // variable-major rows: row p is edge var_adj[p]; whole variables per tile
  std::vector<int32_t> pos(E);
  for (int64_t p = 0; p < E; ++p) pos[var_adj[p]] = int32_t(p);
  std::vector<int4> vrows(E);
  std::vector<int2> vtiles;
  bool vm = true;
  int32_t t0 = 0, fill = 0;
  uint32_t ends = 0;  // rows of the open tile that end a variable
  for (int64_t v = 0; v < NV && vm; ++v) {
    const int32_t p0 = P.var_eptr[v], deg = P.var_eptr[v + 1] - p0;
    if (deg == 0) continue;
    if (deg > kVmMaxDegree) {
      vm = false;
      break;
    }
    if (fill + deg > kWaveRows) {
      vtiles.push_back(make_int2(t0, int32_t(uint32_t(fill) | (ends << 16))));
      t0 = p0;
      fill = 0;
      ends = 0;
    }
    for (int32_t k = 0; k < deg; ++k)
      vrows[p0 + k] = make_int4(pos[var_adj[p0 + k] ^ 1], var_adj[p0 + k], fill | (deg << 8), 0);
    fill += deg;
    ends |= 1u << (fill - 1);
  }
  if (vm && fill > 0) vtiles.push_back(make_int2(t0, int32_t(uint32_t(fill) | (ends << 16))));
  if (vm) {
    int4* dr;
    int2* dt;
    if ((rc = upload(vrows, &dr, &n->owned)) || (rc = upload(vtiles, &dt, &n->owned))) {
      pgx_mpnn_destroy(n);
      return rc;
    }
    n->vm_rows = dr;
    n->vm_tiles = dt;
    n->vm_T = int64_t(vtiles.size());
  }
  *out = n;
  return PGX_OK;
}

int pgx_mpnn_destroy(pgx_mpnn* n) {
  if (!n) return PGX_OK;
  for (void* p : n->owned) (void)hipFree(p);
  if (n->buf) (void)hipFree(n->buf);
  delete n;
  return PGX_OK;
}

int pgx_mpnn_predict(pgx_mpnn* n, const float* params, const float* lp, int64_t lp_stride,
                     const float* ev, int64_t ev_stride, float* ftov_out, int32_t batch,
                     void* stream) {
  int rc = check_args(n, params, lp, ev, batch);
  if (rc) return rc;
  if (!ftov_out) return set_error(PGX_EINVAL, "pgx_mpnn_predict: ftov_out is null");
  bool vm;
  {
    std::lock_guard<std::mutex> lock(g_opt_mu);
    vm = g_opt_mpnn_rows != 0 && n->vm_rows != nullptr;
  }
  const MpGraph& g = n->g;
  const int64_t B = batch, R = B * g.E;
  const size_t need = al256(size_t(2) * kLayers * kKZ * kH * 2) + 2 * al256(size_t(R) * kH * 2) +
                      (vm ? 0 : al256(size_t(B * g.NV) * kH * 4)) + 1024;
  if ((rc = grow(n, need))) return rc;
  Carve c{n->buf};
  u16* wt = c.take<u16>(kLayers * kKZ * kH);
  u16* wr = c.take<u16>(kLayers * kKZ * kH);
  u16* ha = c.take<u16>(size_t(R) * kH);
  u16* hb = c.take<u16>(size_t(R) * kH);
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_mp_cast, dim3(blocks_for(kLayers * kKZ * kH)), dim3(kThreads), 0, s, params,
                     wt, wr);
  u16* h[4] = {ha, hb, ha, hb};
  if (!vm) {
    float* S = c.take<float>(size_t(B * g.NV) * kH);
    float* Sv[3] = {S, S, S};
    forward(n, params, lp, lp_stride, ev, ev_stride, B, wt, h, Sv, s, ftov_out, 2 * g.E);
  } else {
    hipLaunchKernelGGL(k_mp_embed, embed_grid(g.E, B), dim3(kThreads), 0, s, g, params, lp,
                       lp_stride, ev, ev_stride, B, g.var_adj, h[0]);
    const unsigned lb = fwd_blocks(B * n->vm_T * kWaveRows);
    for (int l = 0; l < kLayers; ++l) {
      const float* bias = params + kOffL + l * kLayerSize + kKZ * kH;
      const u16* w_l = wt + l * kKZ * kH;
      if (l == kLayers - 1)
        hipLaunchKernelGGL(k_mp_layer_vm<true>, dim3(lb), dim3(kFwdWaves * 64), 0, s, n->vm_rows,
                           n->vm_tiles, n->vm_T, g.E, static_cast<const u16*>(h[l]), w_l, bias,
                           static_cast<u16*>(nullptr), B, params, ftov_out, 2 * g.E);
      else
        hipLaunchKernelGGL(k_mp_layer_vm<false>, dim3(lb), dim3(kFwdWaves * 64), 0, s,
                           n->vm_rows, n->vm_tiles, n->vm_T, g.E, static_cast<const u16*>(h[l]),
                           w_l, bias, h[l + 1], B, params, static_cast<float*>(nullptr),
                           int64_t(0));
    }
  }
  PGX_HIP(hipGetLastError());
  return PGX_OK;
}

int pgx_mpnn_grad(pgx_mpnn* n, const float* params, const float* lp, int64_t lp_stride,
                  const float* ev, int64_t ev_stride, const float* target, int64_t target_stride,
                  float* grads, float* loss, int32_t batch, void* stream) {
  int rc = check_args(n, params, lp, ev, batch);
  if (rc) return rc;
  if (!target || !grads || !loss) return set_error(PGX_EINVAL, "pgx_mpnn_grad: null argument");
  const int64_t B = batch, R = B * n->g.E, BV = B * n->g.NV;
  const unsigned lb = unsigned(std::min<int64_t>((R + kTileRows - 1) / kTileRows, kMaxLayerBlocks));
  const unsigned hb_blocks = unsigned(std::min<int64_t>((R + 3) / 4, 1024));
  const unsigned loss_blocks = blocks_for(R);
  const size_t need = 2 * al256(size_t(kLayers) * kKZ * kH * 2) + 4 * al256(size_t(R) * kH * 2) +
                      3 * al256(size_t(BV) * kH * 4) + 4 * al256(size_t(R) * kH * 4) +
                      al256(size_t(R) * 4) + al256(size_t(BV) * kH * 4) +
                      al256(size_t(lb) * kLayerPart * 4) +
                      al256(size_t(std::max<unsigned>(hb_blocks, loss_blocks)) *
                            std::max(kEmbedPart, kHeadPart) * 4) + 4096;
  if ((rc = grow(n, need))) return rc;
  Carve c{n->buf};
  u16* wt = c.take<u16>(kLayers * kKZ * kH);
  u16* wr = c.take<u16>(kLayers * kKZ * kH);
  u16* h[4];
  for (auto& x : h) x = c.take<u16>(size_t(R) * kH);
  float* S[3];
  for (auto& x : S) x = c.take<float>(size_t(BV) * kH);
  float* ga = c.take<float>(size_t(R) * kH);
  float* gb = c.take<float>(size_t(R) * kH);
  float* dV = c.take<float>(size_t(R) * kH);
  float* dF = c.take<float>(size_t(R) * kH);
  float* dy = c.take<float>(size_t(R));
  float* T = c.take<float>(size_t(BV) * kH);
  float* lpart = c.take<float>(size_t(lb) * kLayerPart);
  float* spart = c.take<float>(size_t(std::max<unsigned>(hb_blocks, loss_blocks)) *
                               std::max(kEmbedPart, kHeadPart));
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(k_mp_cast, dim3(blocks_for(kLayers * kKZ * kH)), dim3(kThreads), 0, s, params,
                     wt, wr);
  forward(n, params, lp, lp_stride, ev, ev_stride, B, wt, h, S, s);
  // head: loss and dy
  hipLaunchKernelGGL(k_mp_head, dim3(loss_blocks), dim3(kThreads), 0, s, h[3], params, R, n->g.E,
                     static_cast<float*>(nullptr), int64_t(0), target, target_stride, dy, spart,
                     1.f / float(R));
  hipLaunchKernelGGL(k_mp_reduce, dim3(1), dim3(kThreads), 0, s, spart, int32_t(loss_blocks), 1, 1,
                     loss);
  hipLaunchKernelGGL(k_mp_head_bwd, dim3(hb_blocks), dim3(kThreads), 0, s, h[3], params, dy, R, ga,
                     spart);
  hipLaunchKernelGGL(k_mp_reduce, dim3(1), dim3(kThreads), 0, s, spart, int32_t(hb_blocks),
                     kHeadPart, kHeadPart, grads + kOffWh);
  float* gout = ga;
  float* gin = gb;
  for (int l = kLayers - 1; l >= 0; --l) {
    hipLaunchKernelGGL(k_mp_layer_bwd, dim3(lb), dim3(kThreads), 0, s, n->g, h[l], S[l], h[l + 1],
                       gout, wr + l * kKZ * kH, gin, dV, dF, lpart, R);
    hipLaunchKernelGGL(k_mp_reduce, dim3(blocks_for(kLayerPart)), dim3(kThreads), 0, s, lpart,
                       int32_t(lb), kLayerPart, kLayerPart, grads + kOffL + l * kLayerSize);
    hipLaunchKernelGGL(k_mp_varsum<float>, item_batch_grid(n->g.NV * 8, B), dim3(kThreads), 0, s,
                       n->g, static_cast<const float*>(dV), B, T);
    hipLaunchKernelGGL(k_mp_combine, dim3(blocks_for(R * kH)), dim3(kThreads), 0, s, n->g, gin,
                       static_cast<const float*>(T), static_cast<const float*>(dV),
                       static_cast<const float*>(dF), R);
    std::swap(gout, gin);
  }
  hipLaunchKernelGGL(k_mp_embed_bwd, dim3(hb_blocks), dim3(kThreads), 0, s, n->g, lp, lp_stride,
                     ev, ev_stride, h[0], gout, R, spart);
  hipLaunchKernelGGL(k_mp_reduce, dim3(blocks_for(kEmbedPart)), dim3(kThreads), 0, s, spart,
                     int32_t(hb_blocks), kEmbedPart, kEmbedPart, grads + kOffW0);
  PGX_HIP(hipGetLastError());
  return PGX_OK;
}

int pgx_check_numerics(int32_t which, uint32_t first, uint32_t count,
                       unsigned long long* mismatches, void* stream) {
  if ((which != 0 && which != 1) || !mismatches)
    return set_error(PGX_EINVAL, "pgx_check_numerics: bad argument");
  if (count == 0) return PGX_OK;
  const unsigned blocks = unsigned((uint64_t(count) + kThreads - 1) / kThreads);
  hipLaunchKernelGGL(k_check_numerics, dim3(blocks), dim3(kThreads), 0,
                     static_cast<hipStream_t>(stream), which, first, count, mismatches);
  PGX_HIP(hipGetLastError());
  return PGX_OK;
}

int pgx_mpnn_adam(float* params, const float* grads, float* m, float* v, int64_t n, int32_t step,
                  float lr, float beta1, float beta2, float eps, void* stream) {
  if (!params || !grads || !m || !v || n <= 0 || step <= 0)
    return set_error(PGX_EINVAL, "pgx_mpnn_adam: bad argument");
  const float bc1 = float(1.0 - std::pow(double(beta1), step));
  const float bc2 = float(1.0 - std::pow(double(beta2), step));
  hipLaunchKernelGGL(k_mp_adam, dim3(blocks_for(n)), dim3(kThreads), 0,
                     static_cast<hipStream_t>(stream), params, grads, m, v, n, lr, beta1, beta2,
                     eps, bc1, bc2);
  PGX_HIP(hipGetLastError());
  return PGX_OK;
}

}  // extern "C"
