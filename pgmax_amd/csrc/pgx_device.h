// Device-side math and per-factor message updates of the loopy-BP engine.
//
// Every function restates one piece of the reference's FP32 arithmetic with
// the same operation order (pgmax/infer/bp.py, pgmax/factor/enum.py,
// pgmax/factor/logical.py, pgmax/factor/pool.py, pgmax/factor/update_utils.py),
// one thread per factor. The file is compiled with -ffp-contract=off so no
// multiply-add is fused where the reference rounds twice.
#ifndef PGX_DEVICE_H_
#define PGX_DEVICE_H_

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pgx_math.h"
#include "pgx_plan.h"

namespace pgx {

constexpr float kMsgNegInf = -1e32f;    // pgmax/utils/__init__.py:26
constexpr float kLogPotMaxAbs = 1e6f;   // pgmax/utils/__init__.py:32
constexpr float kNegInf = -__builtin_inff();
constexpr float kStabilityThre = 0.5f;  // pgmax/factor/logical.py:33
constexpr float kLn2 = 0.693147182464599609375f;  // float32(log(2))

struct DevGraph {
  const int32_t* var_vs0;
  const int32_t* var_eptr;
  const int32_t* var_adj_es;
  const int32_t* edge_es0;
  const int32_t* edge_vs0;
  const int32_t* fac_e0;
  const int32_t* tables;
  const Block* blocks;
  const Task* tasks;
  const int32_t* heavy_vars;
  const int32_t* heavy_cptr;
  const int32_t* heavy_pofs;
  const int32_t* heavy_chunk_var;
  const int32_t* heavy_chunk_k;
  const int32_t* heavy_gpofs;  // group-sum slots of very heavy variables, or -1
  const int32_t* heavy_chunk_stride;  // evenly spaced full chunks: position step, else 0
  const int32_t* heavy_col;    // chunks summed one per thread (null: all, in order)
  const int32_t* heavy_row;    // chunks of consecutive binary edges, one per 8-lane group
  int32_t num_heavy_col, num_heavy_row;
  const int4* chunk_desc;      // per chunk (chunk order) -- see Plan::chunk_desc
  const int4* col_desc;        // per column chunk (null: chunk_desc, all chunks)
  const int4* row_desc;        // per row chunk
  const int4* heavy_desc;      // per heavy variable -- see Plan::heavy_desc
  const int32_t* vh_group_h;
  const int32_t* vh_group_j;
  int32_t num_vh_groups;
  const int32_t* perm;            // engine position of each reference edge state, or null
  const int32_t* var_adj_es_ref;  // var_adj_es in the reference layout
  const int32_t* edge_var;
  int32_t num_vars, num_edges, num_tasks, num_blocks, num_heavy, num_heavy_chunks;
  int32_t fusable;  // all factors pairwise and all variables of degree <= kFuseMaxDeg
  // complete bipartite binary graph (Plan::bip_*): rows / columns as
  // (first variable state, first chunk slot); bip = 1 runs k_bipartite
  const int2* bip_row;
  const int2* bip_col;
  int32_t bip_na, bip_nb, bip;
  int32_t bip_unary, bip_pb, bip_uc0;  // Plan::bip_unary, bip_pb, bip_uc0
  int32_t bip_affine;   // bip_row / bip_col entries are affine: see Plan::bip_aff
  int32_t bip_aff[8];
  const int4* var_pos4;   // Plan::var_pos4 (null: none)
  int32_t packed;         // engine messages as one float per binary edge (pair_encode)
  const int4* edge_pos4;  // Plan::edge_pos4 (null: none)
  int32_t even2;    // 2-state edges and variables start at even offsets in every buffer
  int32_t lp_nt;    // lane-group pairwise kernels read log-potentials non-temporally
  int32_t lp_al4;   // 1: log-potentials of every batch element are 16-byte aligned;
                    // 2: also read them non-temporally
  int64_t partial_size;  // partial-sum slots per batch element
  int64_t num_es, num_vs, num_lp;
};

// NaN-propagating max / min: XLA's max and min (and jnp.maximum/minimum)
// return NaN when either operand is NaN. CDNA4's v_maximum3_f32 /
// v_minimum3_f32 (IEEE 754-2019 maximum / minimum) do so in one
// instruction; they order -0 below +0, which changes no value.
__device__ __forceinline__ float nmax(float a, float b) { return __builtin_elementwise_maximum(a, b); }
__device__ __forceinline__ float nmin(float a, float b) { return __builtin_elementwise_minimum(a, b); }

// jnp.clip(lp, -1e6, 1e6) (pgmax/infer/bp.py:85-87).
__device__ __forceinline__ float clip_lp(float x) { return nmin(nmax(x, -kLogPotMaxAbs), kLogPotMaxAbs); }

// 8- and 16-byte vector accesses for aligned binary edges / pairwise factors.
__device__ __forceinline__ float2 ld2(const float* p) { return *reinterpret_cast<const float2*>(p); }
__device__ __forceinline__ float4 ld4(const float* p) { return *reinterpret_cast<const float4*>(p); }
__device__ __forceinline__ void st2(float* p, float a, float b) {
  *reinterpret_cast<float2*>(p) = make_float2(a, b);
}

// Exactly N consecutive floats at a 4-byte aligned p, as 16-, 8- and 4-byte
// accesses (multi-dword global accesses need only dword alignment): the
// same addresses as N scalar accesses in fewer instructions and cache-line
// lookups.
template <int N>
__device__ __forceinline__ void ldv(const float* p, float* x) {
  constexpr int n4 = N / 4, r = N % 4;
#pragma unroll
  for (int j = 0; j < n4; ++j) {
    float4 v;
    __builtin_memcpy(&v, p + 4 * j, 16);
    x[4 * j] = v.x;
    x[4 * j + 1] = v.y;
    x[4 * j + 2] = v.z;
    x[4 * j + 3] = v.w;
  }
  if constexpr (r >= 2) {
    float2 v;
    __builtin_memcpy(&v, p + 4 * n4, 8);
    x[4 * n4] = v.x;
    x[4 * n4 + 1] = v.y;
  }
  if constexpr (r & 1) x[N - 1] = p[N - 1];
}
// ldv with non-temporal loads: streamed data (read once per iteration)
// that should not displace reused lines from the L2.
template <int N>
__device__ __forceinline__ void ldv_nt(const float* p, float* x) {
  typedef float f4u __attribute__((ext_vector_type(4), aligned(4)));
  typedef float f2u __attribute__((ext_vector_type(2), aligned(4)));
  constexpr int n4 = N / 4, r = N % 4;
#pragma unroll
  for (int j = 0; j < n4; ++j) {
    const f4u v = __builtin_nontemporal_load(reinterpret_cast<const f4u*>(p + 4 * j));
    x[4 * j] = v.x;
    x[4 * j + 1] = v.y;
    x[4 * j + 2] = v.z;
    x[4 * j + 3] = v.w;
  }
  if constexpr (r >= 2) {
    const f2u v = __builtin_nontemporal_load(reinterpret_cast<const f2u*>(p + 4 * n4));
    x[4 * n4] = v.x;
    x[4 * n4 + 1] = v.y;
  }
  if constexpr (r & 1) x[N - 1] = __builtin_nontemporal_load(p + N - 1);
}

template <int N>
__device__ __forceinline__ void stv(float* p, const float* x) {
  constexpr int n4 = N / 4, r = N % 4;
#pragma unroll
  for (int j = 0; j < n4; ++j) {
    const float4 v = make_float4(x[4 * j], x[4 * j + 1], x[4 * j + 2], x[4 * j + 3]);
    __builtin_memcpy(p + 4 * j, &v, 16);
  }
  if constexpr (r >= 2) {
    const float2 v = make_float2(x[4 * n4], x[4 * n4 + 1]);
    __builtin_memcpy(p + 4 * n4, &v, 8);
  }
  if constexpr (r & 1) p[N - 1] = x[N - 1];
}

// x / T and T * x. Both are exact identities at T = 1 (sum-product), where
// they are skipped; the results are bit-identical either way.
__device__ __forceinline__ float div_t(float x, float t) { return t == 1.f ? x : x / t; }
__device__ __forceinline__ float mul_t(float x, float t) { return t == 1.f ? x : t * x; }

// exp(d / T) for d = x - max.
__device__ __forceinline__ float exp_rel(float d, float t) { return pgx_exp(div_t(d, t)); }

// Log-sum-exp over values x_i with their NaN-propagating maximum m
// (pgmax/factor/update_utils.py:88-98): T * log(sum exp((x_i - m) / T)) + m.
// When m is finite no x_i is NaN and every x_i - m is <= 0 or -inf, so each
// term is pgx_exp_shift(x_i - m) -- pgx_exp on such arguments, without the
// special-case selects -- and the sum is at least 1 (the maximal term).
// When m is NaN or infinite, some difference is NaN (a NaN value, or
// inf - inf) and the reference's sum is NaN: lse_finish returns NaN.
// Only for log-sum-exps over at least one value (full pairwise tables: every
// edge state has a config per state of the other edge); over no values the
// reference gives T * log(0) + (-inf) = -inf, which exp_le0 + the plain
// finish below keep.
__device__ __forceinline__ float exp_term(float d, float t) { return pgx_exp_shift(div_t(d, t)); }
__device__ __forceinline__ float lse_finish(float u, float m, float t) {
  const float r = mul_t(pgx_log(u), t) + m;
  return (m - m == 0.f) ? r : __builtin_nanf("");  // m - m == 0 iff m is finite
}
// exp_rel where d <= 0 or NaN (pgx_exp_nonpos gives the same value for such
// arguments), for log-sum-exps that may run over no values.
__device__ __forceinline__ float exp_le0(float d, float t) { return pgx_exp_nonpos(div_t(d, t)); }

// Temperature log-sum-exp of two config summaries sharing an edge state,
// as pgmax/factor/update_utils.py:88-98 computes it: with m = max(x, y),
// T * log((0 + exp((x - m)/T)) + exp((y - m)/T)) + m. One of the two
// exponentials is exp(0) = 1 exactly and IEEE addition commutes, so the sum
// is 1 + exp((min - max)/T) bit for bit, with one exponential per lane and
// no divergence. (Ties give exp(0) = 1 for both; NaN and -inf inputs give
// NaN differences as in the reference.)
__device__ __forceinline__ float lse2(float x, float y, float t) {
  const float m = nmax(x, y);
  const float d = nmin(x, y) - m;
  const float r = mul_t(pgx_log(1.f + pgx_exp_nonpos(div_t(d, t))), t) + m;  // d <= 0 or NaN
  // the reference's (+inf - +inf) inside exp makes a +inf maximum NaN
  return m == __builtin_inff() ? __builtin_nanf("") : r;
}

// T * log(exp(a/T) + exp(b/T)) (pgmax/factor/update_utils.py:150-166).
__device__ __forceinline__ float logaddexp_t(float a, float b, float t) {
  const float mx = nmax(a, b), mn = nmin(a, b);
  return mul_t(pgx_log1p(pgx_exp_nonpos(div_t(mn - mx, t))), t) + mx;
}

// log(1 - exp(-x)) (pgmax/factor/update_utils.py:134-147).
__device__ __forceinline__ float log1mexp(float x) {
  return (x <= kLn2) ? pgx_log(-pgx_expm1(-x)) : pgx_log1p(-pgx_exp(-x));
}

// T * log(exp(a/T) - exp(b/T)) if a >= b + eps else -inf
// (pgmax/factor/update_utils.py:169-190).
__device__ __forceinline__ float logminusexp_t(float a, float b, float t, float eps) {
  return (a >= b + eps) ? mul_t(log1mexp(div_t(a - b, t)), t) + a : kNegInf;
}

// ---------------------------------------------------------------------------
// I/O policies. A factor update asks its policy for incoming var->factor
// messages and hands it the outgoing factor->var messages of each edge.
// ---------------------------------------------------------------------------

// One-float form of a normalised binary message (r0, r1). Normalisation
// subtracts the larger state, so with finite values one state is exactly 0
// and the pair is stored as v = r1 (r0 == 0, v <= 0) or v = -r0 (r1 == 0,
// v > 0). Non-finite inputs leave either both states NaN or one NaN and the
// other the clip value -1e32 (the other state was +inf); those three pairs
// get reserved NaN bit patterns. Decoding returns the same values (a -0
// state comes back as +0).
constexpr uint32_t kPairNaN = 0x7fc00000u;   // (NaN, NaN)
constexpr uint32_t kPairNaN0 = 0x7fc00001u;  // (NaN, -1e32)
constexpr uint32_t kPairNaN1 = 0x7fc00002u;  // (-1e32, NaN)
__device__ __forceinline__ float pair_encode(float r0, float r1) {
  float v = (r0 == 0.f) ? r1 : -r0;
  if (r0 != r0 || r1 != r1) {  // NaN states (non-finite inputs only)
    const bool n0 = r0 != r0, n1 = r1 != r1;
    v = __uint_as_float((n0 && n1) ? kPairNaN : (n0 ? kPairNaN0 : kPairNaN1));
  }
  return v;
}
__device__ __forceinline__ float2 pair_decode(float v) {
  float2 r = (v <= 0.f) ? make_float2(0.f, v) : make_float2(-v, 0.f);
  if (v != v) {
    const uint32_t u = __float_as_uint(v);
    const float nan = __builtin_nanf("");
    r = (u == kPairNaN0) ? make_float2(nan, kMsgNegInf)
                         : ((u == kPairNaN1) ? make_float2(kMsgNegInf, nan) : make_float2(nan, nan));
  }
  return r;
}

// Damped, normalised and clipped new message of one binary edge from its
// old message (o0, o1) and the factor's output (f0, f1)
// (pgmax/infer/bp.py:127-133, :249-259); folds max |new - old| into dl.
__device__ __forceinline__ void damp_norm2(float d, float omd, float o0, float o1, float f0,
                                           float f1, float& r0, float& r1, float& dl) {
  const float n0 = d * o0 + omd * f0;
  const float n1 = d * o1 + omd * f1;
  const float mx = nmax(n0, n1);
  r0 = nmax(n0 - mx, kMsgNegInf);
  r1 = nmax(n1 - mx, kMsgNegInf);
  dl = nmax(dl, nmax(fabsf(r0 - o0), fabsf(r1 - o1)));
}

// BP iteration: vtof = var_sum - msg (pgmax/infer/bp.py:217-218); outgoing
// messages are damped, normalised per edge and clipped (bp.py:127-133,
// :249-259), and the max-abs change is tracked (bp.py:136).
struct EngineIO {
  static constexpr bool kPlanar = true;  // messages in the engine layout
  const float* vsum;
  const float* mo;
  float* mn;
  const float* lp;
  float d, omd;
  const int32_t* edge_vs0;
  bool even2;  // 2-state edges / variables start at even indices
  // messages as one float per binary edge at index es / 2 (pair_encode);
  // used for graphs of 2x2 factors only (edge_vt2 / out2)
  bool packed = false;
  __device__ __forceinline__ float vt(int32_t vs, int32_t es) const { return vsum[vs] - mo[es]; }
  // vtof of the ns <= MAXS states of edge e (first edge state es)
  // vtof of a binary edge whose first variable state vs0 the caller loaded
  __device__ __forceinline__ void edge_vt2(int32_t, int32_t vs0, int32_t es, float* x) const {
    if (packed) {
      const float2 v = ld2(vsum + vs0), o = pair_decode(mo[es >> 1]);
      x[0] = v.x - o.x;
      x[1] = v.y - o.y;
    } else if (even2) {
      const float2 v = ld2(vsum + vs0), o = ld2(mo + es);
      x[0] = v.x - o.x;
      x[1] = v.y - o.y;
    } else {
      x[0] = vsum[vs0] - mo[es];
      x[1] = vsum[vs0 + 1] - mo[es + 1];
    }
  }
  // the old message (o[0], o[1]) of a binary edge, as out2 reads it
  __device__ __forceinline__ void old2(int32_t es, float* o) const {
    const float2 v = packed ? pair_decode(mo[es >> 1])
                            : (even2 ? ld2(mo + es) : make_float2(mo[es], mo[es + 1]));
    o[0] = v.x;
    o[1] = v.y;
  }
  // edge_vt2 that also hands back the old message, for out2o
  __device__ __forceinline__ void edge_vt2o(int32_t, int32_t vs0, int32_t es, float* x,
                                            float* o) const {
    old2(es, o);
    const float2 v = (packed || even2) ? ld2(vsum + vs0) : make_float2(vsum[vs0], vsum[vs0 + 1]);
    x[0] = v.x - o[0];
    x[1] = v.y - o[1];
  }
  template <int MAXS>
  __device__ __forceinline__ void edge_vt(int32_t e, int32_t es, int32_t ns, float* x) const {
    const int32_t vs0 = edge_vs0[e];
    if constexpr (MAXS == 2) {
      if (even2) {
        const float2 v = ld2(vsum + vs0), o = ld2(mo + es);
        x[0] = v.x - o.x;
        x[1] = v.y - o.y;
        return;
      }
    }
#pragma unroll
    for (int s = 0; s < MAXS; ++s)
      if (s < ns) x[s] = vsum[vs0 + s] - mo[es + s];
  }
  // exactly S states (pairwise tables of a known size): vector accesses
  // o: the edge's old messages, handed back to outx (mo and mn are distinct
  // buffers, so no reload is needed after the first edge's store)
  template <int S>
  __device__ __forceinline__ void edge_vtx(int32_t e, int32_t es, float* x, float* o) const {
    float v[S];
    ldv<S>(vsum + edge_vs0[e], v);
    ldv<S>(mo + es, o);
#pragma unroll
    for (int s = 0; s < S; ++s) x[s] = v[s] - o[s];
  }
  template <int N>
  __device__ __forceinline__ void potv(int32_t c, float* x) const {
    ldv<N>(lp + c, x);
#pragma unroll
    for (int k = 0; k < N; ++k) x[k] = clip_lp(x[k]);
  }
  __device__ __forceinline__ float pot(int32_t c) const { return clip_lp(lp[c]); }
  // four consecutive log-potentials, 16-byte aligned when al
  // mode 0: unaligned, 1: one 16-byte load, 2: one 16-byte non-temporal load
  __device__ __forceinline__ float4 pot4(int32_t c, int32_t mode) const {
    float4 v;
    if (mode == 2) {
      typedef float f32x4v __attribute__((ext_vector_type(4)));
      const f32x4v t = __builtin_nontemporal_load(reinterpret_cast<const f32x4v*>(lp + c));
      v = make_float4(t.x, t.y, t.z, t.w);
    } else if (mode == 1) {
      v = ld4(lp + c);
    } else {
      v = make_float4(lp[c], lp[c + 1], lp[c + 2], lp[c + 3]);
    }
    return make_float4(clip_lp(v.x), clip_lp(v.y), clip_lp(v.z), clip_lp(v.w));
  }
  // out2 with the edge's old message already loaded (edge_vt2o)
  __device__ __forceinline__ void out2o(int32_t es, float f0, float f1, const float* o,
                                        float& dl) const {
    float r0, r1;
    damp_norm2(d, omd, o[0], o[1], f0, f1, r0, r1, dl);
    if (packed) {
      mn[es >> 1] = pair_encode(r0, r1);
    } else if (even2) {
      st2(mn + es, r0, r1);
    } else {
      mn[es] = r0;
      mn[es + 1] = r1;
    }
  }
  // one binary edge: ftov (f0, f1)
  __device__ __forceinline__ void out2(int32_t es, float f0, float f1, float& dl) const {
    float o0, o1;
    if (packed) {
      const float2 o = pair_decode(mo[es >> 1]);
      float r0, r1;
      damp_norm2(d, omd, o.x, o.y, f0, f1, r0, r1, dl);
      mn[es >> 1] = pair_encode(r0, r1);
      return;
    }
    if (even2) {
      const float2 o = ld2(mo + es);
      o0 = o.x;
      o1 = o.y;
    } else {
      o0 = mo[es];
      o1 = mo[es + 1];
    }
    float r0, r1;
    damp_norm2(d, omd, o0, o1, f0, f1, r0, r1, dl);
    if (even2) {
      st2(mn + es, r0, r1);
    } else {
      mn[es] = r0;
      mn[es + 1] = r1;
    }
  }
  // an edge of ns <= MAXS states, ftov in registers
  template <int MAXS>
  __device__ __forceinline__ void outn(int32_t es, int32_t ns, const float* f, float& dl) const {
    float n[MAXS], o[MAXS];
    float mx = kNegInf;
#pragma unroll
    for (int s = 0; s < MAXS; ++s)
      if (s < ns) {
        o[s] = mo[es + s];
        n[s] = d * o[s] + omd * f[s];
        mx = nmax(mx, n[s]);
      }
#pragma unroll
    for (int s = 0; s < MAXS; ++s)
      if (s < ns) {
        const float r = nmax(n[s] - mx, kMsgNegInf);
        mn[es + s] = r;
        dl = nmax(dl, fabsf(r - o[s]));
      }
  }
  template <int S>
  __device__ __forceinline__ void outx(int32_t es, const float* f, const float* o,
                                      float& dl) const {
    float n[S], r[S];
    float mx = kNegInf;
#pragma unroll
    for (int s = 0; s < S; ++s) {
      n[s] = d * o[s] + omd * f[s];
      mx = nmax(mx, n[s]);
    }
#pragma unroll
    for (int s = 0; s < S; ++s) {
      r[s] = nmax(n[s] - mx, kMsgNegInf);
      dl = nmax(dl, fabsf(r[s] - o[s]));
    }
    stv<S>(mn + es, r);
  }
  // general path: raw ftov staged in the output, then finished per edge
  __device__ __forceinline__ void raw(int32_t es, float f) const { mn[es] = f; }
  __device__ __forceinline__ void finish(int32_t es, int32_t ns, float& dl) const {
    float mx = kNegInf;
    for (int32_t s = 0; s < ns; ++s) {
      const float v = d * mo[es + s] + omd * mn[es + s];
      mn[es + s] = v;
      mx = nmax(mx, v);
    }
    for (int32_t s = 0; s < ns; ++s) {
      const float r = nmax(mn[es + s] - mx, kMsgNegInf);
      dl = nmax(dl, fabsf(r - mo[es + s]));
      mn[es + s] = r;
    }
  }
};

// One factor-type update on given vtof (FAC_TO_VAR_UPDATES[type], the
// plugin API): raw ftov out, no damping or normalisation.
struct OpIO {
  static constexpr bool kPlanar = false;  // caller's arrays in the reference layout
  const float* vtof;
  float* out;
  const float* lp;
  __device__ __forceinline__ float vt(int32_t, int32_t es) const { return vtof[es]; }
  template <int MAXS>
  __device__ __forceinline__ void edge_vt(int32_t, int32_t es, int32_t ns, float* x) const {
#pragma unroll
    for (int s = 0; s < MAXS; ++s)
      if (s < ns) x[s] = vtof[es + s];
  }
  template <int S>
  __device__ __forceinline__ void edge_vtx(int32_t, int32_t es, float* x, float*) const {
    ldv<S>(vtof + es, x);
  }
  __device__ __forceinline__ float pot(int32_t c) const { return lp[c]; }
  template <int N>
  __device__ __forceinline__ void potv(int32_t c, float* x) const {
    ldv<N>(lp + c, x);
  }
  __device__ __forceinline__ void edge_vt2(int32_t, int32_t, int32_t es, float* x) const {
    x[0] = vtof[es];
    x[1] = vtof[es + 1];
  }
  __device__ __forceinline__ void edge_vt2o(int32_t e, int32_t vs0, int32_t es, float* x,
                                            float* o) const {
    edge_vt2(e, vs0, es, x);
    o[0] = o[1] = 0.f;  // no damping here
  }
  __device__ __forceinline__ void out2o(int32_t es, float f0, float f1, const float*,
                                        float& dl) const {
    out2(es, f0, f1, dl);
  }
  __device__ __forceinline__ float4 pot4(int32_t c, int32_t) const {
    return make_float4(lp[c], lp[c + 1], lp[c + 2], lp[c + 3]);
  }
  __device__ __forceinline__ void out2(int32_t es, float f0, float f1, float&) const {
    out[es] = f0;
    out[es + 1] = f1;
  }
  template <int MAXS>
  __device__ __forceinline__ void outn(int32_t es, int32_t ns, const float* f, float&) const {
#pragma unroll
    for (int s = 0; s < MAXS; ++s)
      if (s < ns) out[es + s] = f[s];
  }
  template <int S>
  __device__ __forceinline__ void outx(int32_t es, const float* f, const float*, float&) const {
    stv<S>(out + es, f);
  }
  __device__ __forceinline__ void raw(int32_t es, float f) const { out[es] = f; }
  __device__ __forceinline__ void finish(int32_t, int32_t, float&) const {}
};

// ---------------------------------------------------------------------------
// Enum factors (pgmax/factor/enum.py:451-475).
//   summary[c] = (sum over the config's rows of vtof, from 0) + lp[c]
//   out[es]    = max over rows with edge state es (T = 0), or
//                T * log(sum exp((summary - max) / T)) + max (T > 0)
//   ftov[es]   = out[es] - vtof[es]
// ---------------------------------------------------------------------------

// ftov of a 2x2 factor, config (p, q) at lp4 component 2p + q, from the
// vtof (a0, a1) of its first edge and (b0, b1) of its second
// (pgmax/factor/enum.py:451-474 for two variables of two states).
template <bool kT0>
__device__ __forceinline__ void bin2_ftov(float a0, float a1, float b0, float b1, float4 lp4,
                                          float t, float& fa0, float& fa1, float& fb0,
                                          float& fb1) {
  const float s00 = (a0 + b0) + lp4.x;
  const float s01 = (a0 + b1) + lp4.y;
  const float s10 = (a1 + b0) + lp4.z;
  const float s11 = (a1 + b1) + lp4.w;
  float oa0, oa1, ob0, ob1;
  if (kT0) {
    oa0 = nmax(s00, s01);
    oa1 = nmax(s10, s11);
    ob0 = nmax(s00, s10);
    ob1 = nmax(s01, s11);
  } else {
    oa0 = lse2(s00, s01, t);
    oa1 = lse2(s10, s11, t);
    ob0 = lse2(s00, s10, t);
    ob1 = lse2(s01, s11, t);
  }
  fa0 = oa0 - a0;
  fa1 = oa1 - a1;
  fb0 = ob0 - b0;
  fb1 = ob1 - b1;
}

// 2 x 2 pairwise factor with all four configs in row-major order.
template <class IO, bool kT0>
__device__ __forceinline__ void enum_bin2(const DevGraph& g, const Block& b, int32_t i,
                                          const IO& io, float t, float& dl) {
  const bool planar = IO::kPlanar && (b.pad[0] & 2) != 0;
  const int32_t es = planar ? b.es0 + 2 * i : b.es0 + 4 * i;             // first edge
  const int32_t es_b = planar ? b.es0 + 2 * (b.count + i) : es + 2;       // second edge
  const int32_t e = b.e0 + 2 * i;
  const int32_t c = b.c0 + 4 * i;
  // both edges' first variable states with one 8-byte load when aligned
  int32_t va, vb;
  if ((e & 1) == 0) {
    const int2 p = *reinterpret_cast<const int2*>(g.edge_vs0 + e);
    va = p.x;
    vb = p.y;
  } else {
    va = g.edge_vs0[e];
    vb = g.edge_vs0[e + 1];
  }
  float xa[2], xb[2];
  io.edge_vt2(e, va, es, xa);
  io.edge_vt2(e + 1, vb, es_b, xb);
  const float a0 = xa[0], a1 = xa[1], b0 = xb[0], b1 = xb[1];
  const float4 lp4 = io.pot4(c, (b.pad[0] & 1) != 0 ? g.lp_al4 : 0);
  float fa0, fa1, fb0, fb1;
  bin2_ftov<kT0>(a0, a1, b0, b1, lp4, t, fa0, fa1, fb0, fb1);
  io.out2(es, fa0, fa1, dl);
  io.out2(es_b, fb0, fb1, dl);
}

// N 2x2 factors per thread (indices idx[k], -1: none) for the per-phase
// kernel on even2 graphs: every factor's offsets are loaded first, then
// every operand (variable sums, old messages, raw log-potentials), then the
// updates run -- two rounds of memory latency for N factors instead of
// three per factor. Same arithmetic as enum_bin2 + EngineIO::out2.
template <bool kT0, int N>
__device__ __forceinline__ void enum_bin2_n(const DevGraph& g, const Block& b, const int32_t* idx,
                                            const EngineIO& io, float t, float& dl) {
  const bool planar = (b.pad[0] & 2) != 0;
  int32_t va[N], vb[N];
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const int32_t e = b.e0 + 2 * max(idx[k], 0);
    if ((e & 1) == 0) {
      const int2 p = *reinterpret_cast<const int2*>(g.edge_vs0 + e);
      va[k] = p.x;
      vb[k] = p.y;
    } else {
      va[k] = g.edge_vs0[e];
      vb[k] = g.edge_vs0[e + 1];
    }
  }
  float2 sa[N], sb[N], oa[N], ob[N];
  float4 lq[N];
  const bool al4 = (b.pad[0] & 1) && g.lp_al4;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const int32_t i = max(idx[k], 0);
    const int32_t es = planar ? b.es0 + 2 * i : b.es0 + 4 * i;
    const int32_t es_b = planar ? b.es0 + 2 * (b.count + i) : es + 2;
    sa[k] = ld2(io.vsum + va[k]);
    sb[k] = ld2(io.vsum + vb[k]);
    if (io.packed) {
      oa[k] = make_float2(io.mo[es >> 1], 0.f);
      ob[k] = make_float2(io.mo[es_b >> 1], 0.f);
    } else {
      oa[k] = ld2(io.mo + es);
      ob[k] = ld2(io.mo + es_b);
    }
    const float* q = io.lp + b.c0 + 4 * i;
    lq[k] = al4 ? ld4(q) : make_float4(q[0], q[1], q[2], q[3]);
  }
#pragma unroll
  for (int k = 0; k < N; ++k) {
    if (idx[k] < 0) continue;
    const int32_t i = idx[k];
    const int32_t es = planar ? b.es0 + 2 * i : b.es0 + 4 * i;
    const int32_t es_b = planar ? b.es0 + 2 * (b.count + i) : es + 2;
    const float2 o_a = io.packed ? pair_decode(oa[k].x) : oa[k];
    const float2 o_b = io.packed ? pair_decode(ob[k].x) : ob[k];
    const float4 lc = make_float4(clip_lp(lq[k].x), clip_lp(lq[k].y), clip_lp(lq[k].z),
                                  clip_lp(lq[k].w));
    float fa0, fa1, fb0, fb1;
    bin2_ftov<kT0>(sa[k].x - o_a.x, sa[k].y - o_a.y, sb[k].x - o_b.x, sb[k].y - o_b.y, lc, t, fa0,
                   fa1, fb0, fb1);
    float na0, na1, nb0, nb1;
    damp_norm2(io.d, io.omd, o_a.x, o_a.y, fa0, fa1, na0, na1, dl);
    damp_norm2(io.d, io.omd, o_b.x, o_b.y, fb0, fb1, nb0, nb1, dl);
    if (io.packed) {
      io.mn[es >> 1] = pair_encode(na0, na1);
      io.mn[es_b >> 1] = pair_encode(nb0, nb1);
    } else {
      st2(io.mn + es, na0, na1);
      st2(io.mn + es_b, nb0, nb1);
    }
  }
}

// S0 x S1 pairwise factor (S <= 8) with all configs in row-major order:
// config a * S1 + b is (a, b) (pgmax/fgroup/enum.py:303-310).
template <class IO, bool kT0, int M0, int M1, bool kExact>
__device__ __forceinline__ void enum_pair_t(const DevGraph& g, const Block& b, int32_t i,
                                            const IO& io, float t, float& dl) {
  // kExact: the block's tables are exactly M0 x M1, so no lane runs masked
  // iterations; otherwise M0, M1 bound the runtime sizes.
  const int32_t s0 = kExact ? M0 : b.s0, s1 = kExact ? M1 : b.s1;
  const int32_t es = b.es0 + b.esf * i;
  const int32_t e = b.e0 + 2 * i;
  const int32_t c = b.c0 + b.nc * i;
  float xa[M0], xb[M1], oa[M0], ob[M1];
  // exact sizes: whole edges and whole table rows in vector loads (a row at
  // a time, which keeps register use near the scalar form's); the old
  // messages stay in registers for the output step
  float mra[kExact ? M0 : 1], mrb[kExact ? M1 : 1];
  if constexpr (kExact) {
    io.template edge_vtx<M0>(e, es, xa, mra);
    io.template edge_vtx<M1>(e + 1, es + M0, xb, mrb);
  } else {
    io.template edge_vt<M0>(e, es, s0, xa);
    io.template edge_vt<M1>(e + 1, es + s0, s1, xb);
  }
  auto pot_row = [&](int p, float* row) {
    if constexpr (kExact) {
      io.template potv<M1>(c + p * M1, row);
    } else {
#pragma unroll
      for (int q = 0; q < M1; ++q)
        if (q < s1) row[q] = io.pot(c + p * s1 + q);
    }
  };
#pragma unroll
  for (int k = 0; k < M0; ++k) oa[k] = kNegInf;
#pragma unroll
  for (int k = 0; k < M1; ++k) ob[k] = kNegInf;
#pragma unroll
  for (int p = 0; p < M0; ++p) {
    if (p >= s0) continue;
    float row[M1];
    pot_row(p, row);
#pragma unroll
    for (int q = 0; q < M1; ++q)
      if (q < s1) {
        const float s = (xa[p] + xb[q]) + row[q];
        oa[p] = nmax(oa[p], s);
        ob[q] = nmax(ob[q], s);
      }
  }
  if (!kT0) {
    float ua[M0], ub[M1];
#pragma unroll
    for (int k = 0; k < M0; ++k) ua[k] = 0.f;
#pragma unroll
    for (int k = 0; k < M1; ++k) ub[k] = 0.f;
#pragma unroll
    for (int p = 0; p < M0; ++p) {
      if (p >= s0) continue;
      float row[M1];
      pot_row(p, row);
#pragma unroll
      for (int q = 0; q < M1; ++q)
        if (q < s1) {
          const float s = (xa[p] + xb[q]) + row[q];
          ua[p] += exp_term(s - oa[p], t);
          ub[q] += exp_term(s - ob[q], t);
        }
    }
#pragma unroll
    for (int k = 0; k < M0; ++k)
      if (k < s0) oa[k] = lse_finish(ua[k], oa[k], t);
#pragma unroll
    for (int k = 0; k < M1; ++k)
      if (k < s1) ob[k] = lse_finish(ub[k], ob[k], t);
  }
#pragma unroll
  for (int k = 0; k < M0; ++k)
    if (k < s0) oa[k] = oa[k] - xa[k];
#pragma unroll
  for (int k = 0; k < M1; ++k)
    if (k < s1) ob[k] = ob[k] - xb[k];
  if constexpr (kExact) {
    io.template outx<M0>(es, oa, mra, dl);
    io.template outx<M1>(es + M0, ob, mrb, dl);
  } else {
    io.template outn<M0>(es, s0, oa, dl);
    io.template outn<M1>(es + s0, s1, ob, dl);
  }
}

// Pairwise Enum factor with full table (s0, s1 <= 8), config p * s1 + q.
// Square tables of 3..8 states (uniform within a block) run exact-size
// code; other shapes run the 8 x 8 bounded form.
template <class IO, bool kT0>
__device__ __forceinline__ void enum_pair(const DevGraph& g, const Block& b, int32_t i,
                                          const IO& io, float t, float& dl) {
  if (b.s0 == b.s1) {
    switch (b.s0) {
      case 3: return enum_pair_t<IO, kT0, 3, 3, true>(g, b, i, io, t, dl);
      case 4: return enum_pair_t<IO, kT0, 4, 4, true>(g, b, i, io, t, dl);
      case 5: return enum_pair_t<IO, kT0, 5, 5, true>(g, b, i, io, t, dl);
      case 6: return enum_pair_t<IO, kT0, 6, 6, true>(g, b, i, io, t, dl);
      case 7: return enum_pair_t<IO, kT0, 7, 7, true>(g, b, i, io, t, dl);
      case 8: return enum_pair_t<IO, kT0, 8, 8, true>(g, b, i, io, t, dl);
      default: break;
    }
  }
  enum_pair_t<IO, kT0, 8, 8, false>(g, b, i, io, t, dl);
}

// Square pairwise Enum factors (S x S states, 3 <= S <= 8, full table in
// row-major order) on S lanes each, for the per-phase factor kernel. Lane p
// of a factor's lane group owns row p of the table (state p of the first
// edge) and column p (state p of the second edge). A group's loads are
// contiguous runs -- its table's S rows are one S * S-float block, each
// edge's old messages and variable sums S consecutive floats -- so a
// wavefront reads its factors' data in whole cache lines. The lanes of a
// group exchange values through the wavefront's own LDS slots (no
// workgroup barrier): the edges' vtof (xs, 2 S floats), the config
// summaries row by row (sm, S rows padded to S + 1), then the damped
// messages for the per-edge maxima. Every max and sum runs over the same
// operands in the same order as enum_pair_t + EngineIO::outx, so the
// messages are bit-identical to the thread-per-factor form.
//
// Ordering inside one wavefront: its LDS instructions execute in program
// order, so a lane reads what another lane of the same wavefront wrote by an
// earlier instruction; wave_lds_fence keeps the compiler from moving LDS
// accesses across that point.
__device__ __forceinline__ void wave_lds_fence() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One lane's operands of one pairwise factor (pair_lanes).
template <int S>
struct PairOps {
  int32_t es_a, es_b;    // first edge states of the factor's two edges
  float sa, sb, oa, ob;  // variable sums and old messages of state p of both edges
  float row[S];          // raw log-potentials of row p
};

// Edge states of factor i of a square pairwise block: its two edges' states
// in order, or for a 2x2 block stored planar (Plan::perm) the first edges'
// pairs in one plane and the second edges' in the next.
template <int S>
__device__ __forceinline__ int2 pair_edge_states(const Block& b, int32_t i) {
  if (S == 2 && (b.pad[0] & 2)) return make_int2(b.es0 + 2 * i, b.es0 + 2 * (b.count + i));
  return make_int2(b.es0 + 2 * S * i, b.es0 + 2 * S * i + S);
}

// Two consecutive int32 (8 bytes when aligned).
__device__ __forceinline__ int2 ld_pair_i32(const int32_t* a, int32_t e) {
  if ((e & 1) == 0) return *reinterpret_cast<const int2*>(a + e);
  return make_int2(a[e], a[e + 1]);
}

// The factor's edges' first variable states (8 bytes when aligned).
__device__ __forceinline__ int2 pair_edge_vs(const DevGraph& g, int32_t e) {
  return ld_pair_i32(g.edge_vs0, e);
}

// Operands with variable sums from the variable phase (io.vsum).
struct PairVsumLoader {
  const DevGraph& g;
  const EngineIO& io;
  const Block& b;
  int32_t lp_fix;    // >= 0: every factor reads this factor's (identical) table
  using Id = int2;   // the edges' first variable states
  using Pre = int2;  // the same (no second level)
  __device__ __forceinline__ Id id(int32_t i) const { return pair_edge_vs(g, b.e0 + 2 * i); }
  __device__ __forceinline__ Pre pre(Id v) const { return v; }
  template <int S, bool kFix = false>
  __device__ __forceinline__ void load(int32_t i, Pre vs, int32_t p, PairOps<S>& d) const {
    const int2 es = pair_edge_states<S>(b, i);
    d.es_a = es.x;
    d.es_b = es.y;
    d.sa = io.vsum[vs.x + p];
    d.sb = io.vsum[vs.y + p];
    d.oa = io.mo[es.x + p];
    d.ob = io.mo[es.y + p];
    if (kFix) return;  // the row is the caller's, loaded once
    const int32_t li = lp_fix >= 0 ? lp_fix : i;
    if (g.lp_nt)
      ldv_nt<S>(io.lp + b.c0 + S * S * li + p * S, d.row);
    else
      ldv<S>(io.lp + b.c0 + S * S * li + p * S, d.row);
  }
};

// Operands with each variable's sum formed here from its edges' messages
// (Plan::var_pos4: every variable of the block has at most 4 edges), in the
// order of var_sums: the messages in edge order from 0, then + evidence. No
// variable phase runs; a variable's other edges' messages are read by each
// of its factors (mostly from L2).
struct PairFusedLoader {
  const DevGraph& g;
  const EngineIO& io;  // vsum unused
  const float* ev;
  const Block& b;
  int32_t lp_fix;  // >= 0: every factor reads this factor's (identical) table
  struct Id {
    int2 var;  // the edges' variables
    int2 vs;   // and their first variable states
  };
  struct Pre {
    int4 qa, qb;  // the variables' edge positions (Plan::var_pos4)
    int2 vs;
  };
  __device__ __forceinline__ Id id(int32_t i) const {
    const int32_t e = b.e0 + 2 * i;
    return Id{ld_pair_i32(g.edge_var, e), ld_pair_i32(g.edge_vs0, e)};
  }
  __device__ __forceinline__ Pre pre(Id v) const {
    return Pre{g.var_pos4[v.var.x], g.var_pos4[v.var.y], v.vs};
  }
  __device__ __forceinline__ float lane_sum(int4 q, int32_t p) const {
    const int32_t pos[4] = {q.x, q.y, q.z, q.w};
    float x[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) x[u] = io.mo[max(pos[u], 0) + p];
    float acc = 0.f;
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (pos[u] >= 0) acc += x[u];
    return acc;
  }
  template <int S, bool kFix = false>
  __device__ __forceinline__ void load(int32_t i, Pre pr, int32_t p, PairOps<S>& d) const {
    const int2 es = pair_edge_states<S>(b, i);
    d.es_a = es.x;
    d.es_b = es.y;
    const float ea = ev[pr.vs.x + p], eb = ev[pr.vs.y + p];
    d.oa = io.mo[es.x + p];
    d.ob = io.mo[es.y + p];
    if (!kFix) {  // else the row is the caller's, loaded once
      const int32_t li = lp_fix >= 0 ? lp_fix : i;
      if (g.lp_nt)
        ldv_nt<S>(io.lp + b.c0 + S * S * li + p * S, d.row);
      else
        ldv<S>(io.lp + b.c0 + S * S * li + p * S, d.row);
    }
    d.sa = lane_sum(pr.qa, p) + ea;
    d.sb = lane_sum(pr.qb, p) + eb;
  }
};

// The update of one factor from its loaded operands; act = false: no stores.
// kFix: rowc holds the clipped row p of the table every factor shares.
template <bool kT0, int S, bool kFix = false>
__device__ __forceinline__ void pair_update(const PairOps<S>& d, bool act, int32_t p,
                                            const EngineIO& io, float t, float& dl, float* xs,
                                            float* sm, const float* rowc) {
  const float xa = d.sa - d.oa, xb = d.sb - d.ob;
  xs[p] = xa;
  xs[S + p] = xb;
  wave_lds_fence();
  // row p: summaries (xa[p] + xb[q]) + lp[p][q], max, T-logsumexp over q
  float sr[S];
  float m = kNegInf;
#pragma unroll
  for (int q = 0; q < S; ++q) {
    sr[q] = (xa + xs[S + q]) + (kFix ? rowc[q] : clip_lp(d.row[q]));
    m = nmax(m, sr[q]);
  }
  float fa = m;
  if (!kT0) {
    float u = 0.f;
#pragma unroll
    for (int q = 0; q < S; ++q) u += exp_term(sr[q] - m, t);
    fa = lse_finish(u, m, t);
  }
  fa = fa - xa;
#pragma unroll
  for (int q = 0; q < S; ++q) sm[p * (S + 1) + q] = sr[q];
  wave_lds_fence();
  // column p: the same summaries over the rows k
  float sc[S];
  float mb = kNegInf;
#pragma unroll
  for (int k = 0; k < S; ++k) {
    sc[k] = sm[k * (S + 1) + p];
    mb = nmax(mb, sc[k]);
  }
  float fb = mb;
  if (!kT0) {
    float u = 0.f;
#pragma unroll
    for (int k = 0; k < S; ++k) u += exp_term(sc[k] - mb, t);
    fb = lse_finish(u, mb, t);
  }
  fb = fb - xb;
  // damped messages; every lane has read the vtof in xs by now
  const float na = io.d * d.oa + io.omd * fa;
  const float nb = io.d * d.ob + io.omd * fb;
  wave_lds_fence();
  xs[p] = na;
  xs[S + p] = nb;
  wave_lds_fence();
  float mxa = kNegInf, mxb = kNegInf;
#pragma unroll
  for (int k = 0; k < S; ++k) mxa = nmax(mxa, xs[k]);
#pragma unroll
  for (int k = 0; k < S; ++k) mxb = nmax(mxb, xs[S + k]);
  wave_lds_fence();  // the next factor's writes to xs come after these reads
  const float ra = nmax(na - mxa, kMsgNegInf);
  const float rb = nmax(nb - mxb, kMsgNegInf);
  if (act) {
    dl = nmax(dl, fabsf(ra - d.oa));
    dl = nmax(dl, fabsf(rb - d.ob));
    io.mn[d.es_a + p] = ra;
    io.mn[d.es_b + p] = rb;
  }
}

// One wavefront's factors of a square pairwise block: lane group fl takes
// factors i0 + k * step for passes k < passes (those < i_end), its operands
// from the loader L in three dependent levels: id(i) (the first loads of
// factor i), pre(id) (loads that need them) and load(i, pre, p, ops) (the
// rest). The loop is software-pipelined over passes and unrolled by two, so
// that every level runs a pass ahead of the next and no register copy waits
// for a load: while pass k is computed, the operands of pass k + 1, the
// second-level loads of pass k + 2 and the first of pass k + 3 are in flight.
// Loads past the last pass read factor 0 again (always valid).
template <bool kT0, int S, class L, bool kFix = false>
__device__ __forceinline__ void pair_lanes(const L& ld, int32_t i0, int32_t step, int32_t passes,
                                           int32_t i_end, bool grp, int32_t p, const EngineIO& io,
                                           float t, float& dl, float* xs, float* sm,
                                           const float* rowc = nullptr) {
  auto fac = [&](int32_t k) {  // factor of pass k, or -1
    const int32_t i = i0 + k * step;
    return (grp && k < passes && i < i_end) ? i : -1;
  };
  auto idk = [&](int32_t k) { return ld.id(max(fac(k), 0)); };
  // register sets by pass parity: ops of pass k in O[k & 1], second-level
  // loads in P[k & 1], first-level loads in I[k & 1]
  PairOps<S> O0, O1;
  typename L::Id I0 = idk(0), I1 = idk(1);
  typename L::Pre P0 = ld.pre(I0), P1;
  I0 = idk(2);
  ld.template load<S, kFix>(max(fac(0), 0), P0, p, O0);
  P1 = ld.pre(I1);
  I1 = idk(3);
  for (int32_t k = 0; k < passes; k += 2) {
    // invariants: O0 = pass k, P1 = pass k + 1, I0 = pass k + 2, I1 = pass k + 3
    P0 = ld.pre(I0);
    I0 = idk(k + 4);
    ld.template load<S, kFix>(max(fac(k + 1), 0), P1, p, O1);
    pair_update<kT0, S, kFix>(O0, fac(k) >= 0, p, io, t, dl, xs, sm, rowc);
    if (k + 1 >= passes) break;
    P1 = ld.pre(I1);
    I1 = idk(k + 5);
    ld.template load<S, kFix>(max(fac(k + 2), 0), P0, p, O0);
    pair_update<kT0, S, kFix>(O1, fac(k + 1) >= 0, p, io, t, dl, xs, sm, rowc);
  }
}

// LDS floats per wavefront of pair_lanes: (64 / S + 1) lane groups (the
// last for the lanes past 64 / S whole groups), each 2 S vtof + S (S + 1)
// summaries.
__host__ __device__ constexpr int pair_lanes_lds(int s) { return (64 / s + 1) * (2 * s + s * (s + 1)); }
constexpr int kPairLanesLds = pair_lanes_lds(8);  // the largest over S = 2..8
static_assert(pair_lanes_lds(2) <= kPairLanesLds && pair_lanes_lds(3) <= kPairLanesLds &&
                  pair_lanes_lds(7) <= kPairLanesLds,
              "S = 8 needs the most LDS");

// Any Enum factor: config table in `tables` as [nv edge offsets][esf owning
// edge][nc * nv edge-state offsets]; summaries are recomputed per edge state
// instead of being stored, so no per-thread arrays are needed.
template <class IO, bool kT0>
__device__ void enum_general(const DevGraph& g, const Block& b, int32_t i, const IO& io,
                             float t, float& dl) {
  const int32_t nv = b.nv, nc = b.nc, esf = b.esf;
  const int32_t es0 = b.es0 + esf * i;
  const int32_t e0 = b.e0 + nv * i;
  const int32_t c0 = b.c0 + nc * i;
  const int32_t* eoff = g.tables + b.tbl;
  const int32_t* owner = eoff + nv;
  const int32_t* rows = owner + esf;
  auto row_vt = [&](int32_t k) {
    const int32_t j = owner[k];
    return io.vt(g.edge_vs0[e0 + j] + (k - eoff[j]), es0 + k);
  };
  auto summary = [&](int32_t c) {
    float s = 0.f;
    for (int32_t j = 0; j < nv; ++j) s += row_vt(rows[c * nv + j]);
    return s + io.pot(c0 + c);
  };
  for (int32_t k = 0; k < esf; ++k) {
    const int32_t j = owner[k];
    float mx = kNegInf;
    for (int32_t c = 0; c < nc; ++c)
      if (rows[c * nv + j] == k) mx = nmax(mx, summary(c));
    float out = mx;
    if (!kT0) {
      float u = 0.f;
      for (int32_t c = 0; c < nc; ++c)
        if (rows[c * nv + j] == k) u += exp_le0(summary(c) - mx, t);
      out = mul_t(pgx_log(u), t) + mx;
    }
    io.raw(es0 + k, out - row_vt(k));
  }
  for (int32_t j = 0; j < nv; ++j) {
    const int32_t a = eoff[j], z = (j + 1 < nv) ? eoff[j + 1] : esf;
    io.finish(es0 + a, z - a, dl);
  }
}

// One general Enum factor per wavefront (block flag 8): the wavefront stages
// the factor's edge-state vtof (wv, esf floats) and config summaries (ws,
// nc floats) in LDS -- each summary computed once -- then lane k reduces
// edge state k over its own configs from the table's per-edge-state lists,
// in increasing config order, the order enum_general uses. Every thread of
// the workgroup must call it (it synchronises the workgroup); active = false
// marks a wavefront without a factor.
template <class IO, bool kT0>
__device__ void enum_general_wave(const DevGraph& g, const Block& b, int32_t i, bool active,
                                  const IO& io, float t, float& dl, int32_t lane, float* ws,
                                  float* wv) {
  const int32_t nv = b.nv, nc = b.nc, esf = b.esf;
  const int32_t es0 = b.es0 + esf * i;
  const int32_t e0 = b.e0 + nv * i;
  const int32_t c0 = b.c0 + nc * i;
  const int32_t* eoff = g.tables + b.tbl;
  const int32_t* owner = eoff + nv;
  const int32_t* rows = owner + esf;
  const int32_t* lptr = rows + nc * nv;
  const int32_t* lcfg = lptr + esf + 1;
  if (active)
    for (int32_t k = lane; k < esf; k += 64) {
      const int32_t j = owner[k];
      wv[k] = io.vt(g.edge_vs0[e0 + j] + (k - eoff[j]), es0 + k);
    }
  __syncthreads();
  if (active)
    for (int32_t c = lane; c < nc; c += 64) {
      float s = 0.f;
      for (int32_t j = 0; j < nv; ++j) s += wv[rows[c * nv + j]];
      ws[c] = s + io.pot(c0 + c);
    }
  __syncthreads();
  if (active)
    for (int32_t k = lane; k < esf; k += 64) {
      const int32_t a = lptr[k], z = lptr[k + 1];
      float mx = kNegInf;
      for (int32_t p = a; p < z; ++p) mx = nmax(mx, ws[lcfg[p]]);
      float out = mx;
      if (!kT0) {
        float u = 0.f;
        for (int32_t p = a; p < z; ++p) u += exp_le0(ws[lcfg[p]] - mx, t);
        out = mul_t(pgx_log(u), t) + mx;
      }
      io.raw(es0 + k, out - wv[k]);
    }
  __syncthreads();  // the raw outputs of all lanes are visible to finish
  if (active)
    for (int32_t j = lane; j < nv; j += 64) {
      const int32_t a = eoff[j], z = (j + 1 < nv) ? eoff[j + 1] : esf;
      io.finish(es0 + a, z - a, dl);
    }
}

// ---------------------------------------------------------------------------
// OR / AND factors (pgmax/factor/logical.py:561-779). For OR the relevant
// state of every variable is 1, for AND it is 0 (edge_states_offset -1).
// Parents are the factor's edges 0..n-1, the child its last edge.
// ---------------------------------------------------------------------------
template <class IO>
__device__ __forceinline__ void logical_outputs(const DevGraph& g, int32_t e0, int32_t np,
                                                int32_t sr, const IO& io, float t,
                                                bool normalize, float s_oth, float s_acc,
                                                float m1, float m2, int32_t a1, int32_t k0,
                                                int32_t k1, int32_t kstep, bool write_child,
                                                float& dl);

// Sequential accumulation in the engine's fixed order for long sums: runs of
// kVarChunk values (each from 0), groups of kVarChunk run sums (each from
// 0), group sums added from 0. Up to kVarChunk values it is the plain
// sequential sum (oracle/lbp.py chunked_segment_sum states the same order).
struct ChunkedSum {
  float run = 0.f, grp = 0.f, tot = 0.f;
  int32_t n = 0;
  __device__ __forceinline__ void add(float x) {
    run += x;
    if (++n % kVarChunk == 0) {
      grp += run;
      run = 0.f;
      if (n % (kVarChunk * kVarChunk) == 0) {
        tot += grp;
        grp = 0.f;
      }
    }
  }
  __device__ __forceinline__ float value() const {
    float g2 = grp, t2 = tot;
    if (n % kVarChunk) g2 += run;
    if (n % (kVarChunk * kVarChunk)) t2 += g2;
    return t2;
  }
};

// Outgoing (relevant, other) pair of one parent of a logical factor from the
// factor's reductions (pgmax/factor/logical.py:598-757).
__device__ __forceinline__ void logical_parent_msg(bool t0, float t, int32_t np, bool is_a1,
                                                   float rel, float oth, float c_rel,
                                                   float c_oth, float s_oth, float s_acc,
                                                   float m1, float m2, float& par_rel,
                                                   float& par_oth) {
  if (t0) {
    par_rel = (s_acc + c_rel) - nmax(oth, rel);
    const float opt1 = (c_oth + s_oth) - oth;
    const float opt2 = par_rel + nmin(0.f, is_a1 ? m2 : m1);
    par_oth = nmax(opt1, opt2);
  } else {
    const float lse = logaddexp_t(rel, oth, t);
    const float s_lse_wo = s_acc - lse;
    const float s_oth_wo = s_oth - oth;
    par_rel = c_rel + s_lse_wo;
    const float opt1 = c_oth + s_oth_wo;
    const float opt2 = c_rel + s_lse_wo;
    const float opt3 = c_rel + s_oth_wo;
    par_oth = logminusexp_t(logaddexp_t(opt1, opt2, t), opt3, t, 1e-4f);
    if (t < kStabilityThre) {
      const float x = s_oth_wo + (is_a1 ? m2 : m1);
      par_oth = nmax(par_oth, logaddexp_t(opt1, c_rel + x, t));
    }
  }
  if (np == 1) {  // single parent (pgmax/factor/logical.py:739-757)
    par_rel = c_rel;
    par_oth = c_oth;
  }
}

// Relevant-state message to the child (pgmax/factor/logical.py:598-737);
// its other-state message is s_oth.
__device__ __forceinline__ float logical_child_rel(bool t0, float t, float s_oth, float s_acc,
                                                   float m1, float m2) {
  float child_rel;
  if (t0) {
    child_rel = s_acc + nmin(0.f, m1);
  } else {
    child_rel = logminusexp_t(s_acc, s_oth, t, 1e-4f);
    if (t < kStabilityThre)
      child_rel = nmax(child_rel, logaddexp_t(s_oth + m1, s_oth + m2, t));
  }
  return child_rel;
}

// A logical factor of at most kSmallParents parents (AND / OR of a few
// variables): every edge's indices and vtof are loaded before any is used,
// then the same reductions, order and outputs as logical_factor.
constexpr int kSmallParents = 3;
template <class IO>
__device__ __forceinline__ void logical_small(const DevGraph& g, int32_t e0, int32_t np,
                                              int32_t off, const IO& io, float t,
                                              bool normalize, float& dl) {
  const int32_t sr = off == 1 ? 1 : 0;
  const int32_t so = 1 - sr;
  const bool t0 = (t == 0.f);
  // slots 0 .. kSmallParents-1: parents (slots past np load nothing and are
  // not used); slot kSmallParents: the child
  // (old[k]: the slot's old message, kept for the damping of its output)
  int32_t es[kSmallParents + 1];
  float rel[kSmallParents + 1], oth[kSmallParents + 1], old[kSmallParents + 1][2];
#pragma unroll
  for (int k = 0; k <= kSmallParents; ++k) {
    es[k] = 0;
    rel[k] = oth[k] = 0.f;
    old[k][0] = old[k][1] = 0.f;
    if (k < np || k == kSmallParents) {
      const int32_t e = e0 + (k < kSmallParents ? k : np);
      es[k] = g.edge_es0[e];
      float x[2];
      io.edge_vt2o(e, g.edge_vs0[e], es[k], x, old[k]);
      rel[k] = x[sr];
      oth[k] = x[so];
    }
  }
  ChunkedSum c_oth_sum, c_acc_sum;
  float m1 = kNegInf, m2 = kNegInf;
  int32_t a1 = -1;
#pragma unroll
  for (int k = 0; k < kSmallParents; ++k) {
    if (k < np) {
      const float diff = rel[k] - oth[k];
      c_oth_sum.add(oth[k]);
      c_acc_sum.add(t0 ? nmax(oth[k], rel[k]) : logaddexp_t(rel[k], oth[k], t));
      if (diff >= m1) {
        m2 = m1;
        m1 = diff;
        a1 = k;
      } else {
        m2 = nmax(m2, diff);
      }
    }
  }
  const float s_oth = c_oth_sum.value(), s_acc = c_acc_sum.value();
  const float c_rel = rel[kSmallParents], c_oth = oth[kSmallParents];
#pragma unroll
  for (int k = 0; k < kSmallParents; ++k) {
    if (k < np) {
      float par_rel, par_oth;
      logical_parent_msg(t0, t, np, k == a1, rel[k], oth[k], c_rel, c_oth, s_oth, s_acc, m1, m2,
                         par_rel, par_oth);
      float pf[2];
      pf[sr] = normalize ? par_rel - par_oth : par_rel;
      pf[so] = normalize ? 0.f : par_oth;
      io.out2o(es[k], pf[0], pf[1], old[k], dl);
    }
  }
  const float child_rel = logical_child_rel(t0, t, s_oth, s_acc, m1, m2);
  float cf[2];
  cf[sr] = normalize ? child_rel - s_oth : child_rel;
  cf[so] = normalize ? 0.f : s_oth;
  io.out2o(es[kSmallParents], cf[0], cf[1], old[kSmallParents], dl);
}

template <class IO>
__device__ void logical_factor(const DevGraph& g, int32_t f, int32_t off, const IO& io,
                               float t, bool normalize, float& dl) {
  const int32_t e0 = g.fac_e0[f];
  const int32_t np = g.fac_e0[f + 1] - e0 - 1;
  if (np <= kSmallParents) {
    logical_small<IO>(g, e0, np, off, io, t, normalize, dl);
    return;
  }
  const int32_t sr = off == 1 ? 1 : 0;  // relevant state
  const int32_t so = 1 - sr;            // other state
  const bool t0 = (t == 0.f);
  auto rel_of = [&](int32_t k) { return io.vt(g.edge_vs0[e0 + k] + sr, g.edge_es0[e0 + k] + sr); };
  auto oth_of = [&](int32_t k) { return io.vt(g.edge_vs0[e0 + k] + so, g.edge_es0[e0 + k] + so); };

  // s_oth: sum of the other-state messages; s_acc: sum of max (T=0) or
  // logaddexp (T>0); both in the engine's chunked order
  ChunkedSum c_oth_sum, c_acc_sum;
  float m1 = kNegInf, m2 = kNegInf;
  int32_t a1 = -1;
  for (int32_t k = 0; k < np; ++k) {
    const float rel = rel_of(k), oth = oth_of(k);
    const float diff = rel - oth;
    c_oth_sum.add(oth);
    c_acc_sum.add(t0 ? nmax(oth, rel) : logaddexp_t(rel, oth, t));
    // first and second max of the diffs; ties pick the last index
    // (pgmax/factor/update_utils.py:25-64)
    if (diff >= m1) {
      m2 = m1;
      m1 = diff;
      a1 = k;
    } else {
      m2 = nmax(m2, diff);
    }
  }
  const float s_oth = c_oth_sum.value(), s_acc = c_acc_sum.value();
  logical_outputs(g, e0, np, sr, io, t, normalize, s_oth, s_acc, m1, m2, a1, 0, np, 1, true, dl);
}

// Outputs of a logical factor for parents k0, k0 + kstep, ... < k1 (and its
// child when write_child) from the factor's reductions
// (pgmax/factor/logical.py:598-768).
template <class IO>
__device__ __forceinline__ void logical_outputs(const DevGraph& g, int32_t e0, int32_t np,
                                                int32_t sr, const IO& io, float t,
                                                bool normalize, float s_oth, float s_acc,
                                                float m1, float m2, int32_t a1, int32_t k0,
                                                int32_t k1, int32_t kstep, bool write_child,
                                                float& dl) {
  const int32_t so = 1 - sr;
  const bool t0 = (t == 0.f);
  auto rel_of = [&](int32_t k) { return io.vt(g.edge_vs0[e0 + k] + sr, g.edge_es0[e0 + k] + sr); };
  auto oth_of = [&](int32_t k) { return io.vt(g.edge_vs0[e0 + k] + so, g.edge_es0[e0 + k] + so); };
  const float c_rel = rel_of(np), c_oth = oth_of(np);
  for (int32_t k = k0; k < k1; k += kstep) {
    const float rel = rel_of(k), oth = oth_of(k);
    float par_rel, par_oth;
    logical_parent_msg(t0, t, np, k == a1, rel, oth, c_rel, c_oth, s_oth, s_acc, m1, m2, par_rel,
                       par_oth);
    float pf[2];
    if (normalize) {
      pf[sr] = par_rel - par_oth;
      pf[so] = 0.f;
    } else {
      pf[sr] = par_rel;
      pf[so] = par_oth;
    }
    io.out2(g.edge_es0[e0 + k], pf[0], pf[1], dl);
  }
  if (!write_child) return;
  const float child_rel = logical_child_rel(t0, t, s_oth, s_acc, m1, m2);
  const float child_oth = s_oth;
  float cf[2];
  if (normalize) {
    cf[sr] = child_rel - child_oth;
    cf[so] = 0.f;
  } else {
    cf[sr] = child_rel;
    cf[so] = child_oth;
  }
  io.out2(g.edge_es0[e0 + np], cf[0], cf[1], dl);
}

// A logical factor with up to kVarChunk runs of kVarChunk parents on a group
// of kWideLanes lanes. The lanes take the parents in turn (lane q: parents
// q, q + 8, ...), so a group reads consecutive parents together; each run's
// sums are accumulated in parent order through shuffles, run sums are added
// in order, the top-2 diffs are merged (ties pick the larger index), and the
// lanes write their parents' outputs. Same arithmetic and order as
// logical_factor. Every lane of the group must call it.
constexpr int kWideLanes = 8;
template <class IO>
__device__ void logical_factor_wide(const DevGraph& g, int32_t f, int32_t off, const IO& io,
                                    float t, bool normalize, float& dl, int32_t q) {
  const int32_t e0 = g.fac_e0[f];
  const int32_t np = g.fac_e0[f + 1] - e0 - 1;
  const int32_t sr = off == 1 ? 1 : 0;
  const int32_t so = 1 - sr;
  const bool t0 = (t == 0.f);
  const int32_t base = (threadIdx.x & 63) & ~(kWideLanes - 1);
  float s_oth = 0.f, s_acc = 0.f, m1 = kNegInf, m2 = kNegInf;
  int32_t a1 = -1;
  for (int32_t r0 = 0; r0 < np; r0 += kVarChunk) {
    float rs_oth = 0.f, rs_acc = 0.f;  // this run's sums, in parent order
    for (int32_t j = 0; j < kVarChunk; j += kWideLanes) {
      const int32_t k = r0 + j + q;
      float oth = 0.f, acc = 0.f;
      if (k < np) {
        const float rel = io.vt(g.edge_vs0[e0 + k] + sr, g.edge_es0[e0 + k] + sr);
        oth = io.vt(g.edge_vs0[e0 + k] + so, g.edge_es0[e0 + k] + so);
        const float diff = rel - oth;
        acc = t0 ? nmax(oth, rel) : logaddexp_t(rel, oth, t);
        if (diff >= m1) {
          m2 = m1;
          m1 = diff;
          a1 = k;
        } else {
          m2 = nmax(m2, diff);
        }
      }
      for (int32_t u = 0; u < kWideLanes; ++u) {
        const float xo = __shfl(oth, base + u, 64);
        const float xa = __shfl(acc, base + u, 64);
        if (r0 + j + u < np) {
          rs_oth += xo;
          rs_acc += xa;
        }
      }
    }
    s_oth += rs_oth;
    s_acc += rs_acc;
  }
  // top-2 merge across the group
  for (int32_t o = 1; o < kWideLanes; o <<= 1) {
    const float om1 = __shfl_xor(m1, o, 64);
    const float om2 = __shfl_xor(m2, o, 64);
    const int32_t oa1 = __shfl_xor(a1, o, 64);
    if (om1 > m1 || (om1 == m1 && oa1 > a1)) {
      m2 = nmax(om2, m1);
      m1 = om1;
      a1 = oa1;
    } else {
      m2 = nmax(m2, om1);
    }
  }
  logical_outputs(g, e0, np, sr, io, t, normalize, s_oth, s_acc, m1, m2, a1, q, np, kWideLanes,
                  q == 0, dl);
}

// A logical factor of at most kWaveParents parents on one wavefront: lane L
// takes parents 4L .. 4L + 3, so a run of kVarChunk parents is eight
// consecutive lanes and every load of the wavefront is one contiguous
// span. Each lane loads all its parents before any reduction (the loads
// overlap) and keeps them for the outputs (one pass over variable sums and
// messages). A run's sums are formed in parent order by passing the partial
// sum along the run's eight lanes (each adds its four parents), the run sums
// are added in order, and the top-2 diffs are merged across the wavefront
// (ties pick the larger index): the same values, order and outputs as
// logical_factor. Every lane of the wavefront must call it.
constexpr int kWaveItems = 4;                         // parents per lane
constexpr int kWaveParents = 64 * kWaveItems;         // 256
constexpr int kRunLanes = kVarChunk / kWaveItems;     // 8 lanes per run
static_assert(kWaveParents == kWaveLogicalParents && kVarChunk % kWaveItems == 0,
              "the planner's wavefront logical blocks must fit logical_factor_wave");
template <class IO>
__device__ void logical_factor_wave(const DevGraph& g, int32_t f, int32_t off, const IO& io,
                                    float t, bool normalize, float& dl, int32_t lane) {
  const int32_t e0 = g.fac_e0[f];
  const int32_t np = g.fac_e0[f + 1] - e0 - 1;
  const int32_t sr = off == 1 ? 1 : 0;
  const int32_t so = 1 - sr;
  const bool t0 = (t == 0.f);
  const int32_t k0 = kWaveItems * lane;
  float rel[kWaveItems], oth[kWaveItems];
  int32_t es[kWaveItems];
#pragma unroll
  for (int i = 0; i < kWaveItems; ++i) {
    rel[i] = 0.f;
    oth[i] = 0.f;
    es[i] = 0;
    if (k0 + i < np) {
      const int32_t vs0 = g.edge_vs0[e0 + k0 + i];
      es[i] = g.edge_es0[e0 + k0 + i];
      rel[i] = io.vt(vs0 + sr, es[i] + sr);
      oth[i] = io.vt(vs0 + so, es[i] + so);
    }
  }
  const int32_t cvs = g.edge_vs0[e0 + np], ces = g.edge_es0[e0 + np];
  const float c_rel = io.vt(cvs + sr, ces + sr), c_oth = io.vt(cvs + so, ces + so);
  // per lane: the top-2 diffs of its parents in index order (ties: later)
  float m1 = kNegInf, m2 = kNegInf;
  int32_t a1 = -1;
  float acc[kWaveItems];
#pragma unroll
  for (int i = 0; i < kWaveItems; ++i) {
    acc[i] = 0.f;
    if (k0 + i < np) {
      const float diff = rel[i] - oth[i];
      acc[i] = t0 ? nmax(oth[i], rel[i]) : logaddexp_t(rel[i], oth[i], t);
      if (diff >= m1) {
        m2 = m1;
        m1 = diff;
        a1 = k0 + i;
      } else {
        m2 = nmax(m2, diff);
      }
    }
  }
  // run sums: the partial sum travels along the run's lanes in order
  const int32_t pos = lane % kRunLanes;
  float co = 0.f, ca = 0.f;
  for (int32_t step = 0; step < kRunLanes; ++step) {
    const float po = __shfl(co, lane > 0 ? lane - 1 : 0, 64);
    const float pa = __shfl(ca, lane > 0 ? lane - 1 : 0, 64);
    if (pos == step) {
      float so_ = step == 0 ? 0.f : po, sa_ = step == 0 ? 0.f : pa;
#pragma unroll
      for (int i = 0; i < kWaveItems; ++i)
        if (k0 + i < np) {
          so_ += oth[i];
          sa_ += acc[i];
        }
      co = so_;
      ca = sa_;
    }
  }
  // run sums added in order (the last lane of each run holds its run's sum)
  float s_oth = 0.f, s_acc = 0.f;
  for (int32_t r = 0; r < 64 / kRunLanes; ++r) {
    const float xo = __shfl(co, r * kRunLanes + kRunLanes - 1, 64);
    const float xa = __shfl(ca, r * kRunLanes + kRunLanes - 1, 64);
    if (r * kVarChunk < np) {
      s_oth += xo;
      s_acc += xa;
    }
  }
  // top-2 merge across the wavefront
  for (int32_t o = 1; o < 64; o <<= 1) {
    const float om1 = __shfl_xor(m1, o, 64);
    const float om2 = __shfl_xor(m2, o, 64);
    const int32_t oa1 = __shfl_xor(a1, o, 64);
    if (om1 > m1 || (om1 == m1 && oa1 > a1)) {
      m2 = nmax(om2, m1);
      m1 = om1;
      a1 = oa1;
    } else {
      m2 = nmax(m2, om1);
    }
  }
#pragma unroll
  for (int i = 0; i < kWaveItems; ++i) {
    if (k0 + i < np) {
      float par_rel, par_oth;
      logical_parent_msg(t0, t, np, k0 + i == a1, rel[i], oth[i], c_rel, c_oth, s_oth, s_acc, m1,
                         m2, par_rel, par_oth);
      float pf[2];
      if (normalize) {
        pf[sr] = par_rel - par_oth;
        pf[so] = 0.f;
      } else {
        pf[sr] = par_rel;
        pf[so] = par_oth;
      }
      io.out2(es[i], pf[0], pf[1], dl);
    }
  }
  if (lane != 0) return;
  const float child_rel = logical_child_rel(t0, t, s_oth, s_acc, m1, m2);
  float cf[2];
  if (normalize) {
    cf[sr] = child_rel - s_oth;
    cf[so] = 0.f;
  } else {
    cf[sr] = child_rel;
    cf[so] = s_oth;
  }
  io.out2(ces, cf[0], cf[1], dl);
}

// ---------------------------------------------------------------------------
// Pool factors (pgmax/factor/pool.py:328-474): choices are the factor's
// edges 0..n-1, the indicator its last edge.
// ---------------------------------------------------------------------------
template <class IO>
__device__ void pool_factor(const DevGraph& g, int32_t f, const IO& io, float t,
                            bool normalize, float& dl) {
  const int32_t e0 = g.fac_e0[f];
  const int32_t nch = g.fac_e0[f + 1] - e0 - 1;
  const bool t0 = (t == 0.f);
  auto zero_of = [&](int32_t k) { return io.vt(g.edge_vs0[e0 + k], g.edge_es0[e0 + k]); };
  auto one_of = [&](int32_t k) { return io.vt(g.edge_vs0[e0 + k] + 1, g.edge_es0[e0 + k] + 1); };

  float s_zero = 0.f, m1 = kNegInf, m2 = kNegInf;
  int32_t a1 = -1;
  for (int32_t k = 0; k < nch; ++k) {
    const float z = zero_of(k);
    const float diff = one_of(k) - z;
    s_zero += z;
    if (diff >= m1) {
      m2 = m1;
      m1 = diff;
      a1 = k;
    } else {
      m2 = nmax(m2, diff);
    }
  }
  const float i_zero = zero_of(nch), i_one = one_of(nch);
  const float pid = i_one - i_zero;
  float pi_diff, f_lse = 0.f, am_diff = 0.f;
  if (t0) {
    pi_diff = m1;
  } else {
    const float mam = nmax(m2, -pid);  // max of the diffs with the argmax replaced
    float u1 = 0.f, u2 = 0.f;
    for (int32_t k = 0; k < nch; ++k) {
      const float diff = one_of(k) - zero_of(k);
      u1 += exp_rel(diff - m1, t);
      u2 += exp_rel((k == a1 ? -pid : diff) - mam, t);
    }
    pi_diff = mul_t(pgx_log(u1), t) + m1;
    f_lse = logaddexp_t(pi_diff, -pid, t);
    am_diff = -(mul_t(pgx_log(u2), t) + mam);
  }
  for (int32_t k = 0; k < nch; ++k) {
    const float z = zero_of(k);
    const float diff = one_of(k) - z;
    float pc_diff;
    if (t0) {
      pc_diff = nmin(pid, -(k == a1 ? m2 : m1));
    } else {
      pc_diff = (k == a1) ? am_diff : -logminusexp_t(f_lse, diff, t, 1e-30f);
    }
    float pc_one = (s_zero + i_one) - z;
    if (nch == 1) {  // single choice (pgmax/factor/pool.py:430-450)
      pc_diff = pid;
      pc_one = i_one;
    }
    const int32_t es = g.edge_es0[e0 + k];
    if (normalize)
      io.out2(es, 0.f, pc_diff, dl);
    else
      io.out2(es, pc_one - pc_diff, pc_one, dl);
  }
  const int32_t ies = g.edge_es0[e0 + nch];
  if (normalize)
    io.out2(ies, 0.f, pi_diff, dl);
  else
    io.out2(ies, s_zero, s_zero + pi_diff, dl);
}

// Sets of block kinds a kernel is compiled for (bit k = BlockKind k). A
// kernel only holds the code of its kinds, which keeps the lean pairwise
// paths free of the register pressure of the general ones.
constexpr int kKindsBin2 = 1 << kEnumBin2;
constexpr int kKindsPairwise = (1 << kEnumBin2) | (1 << kEnumPair);
constexpr int kKindsAll = 0x3f;

// Dispatches factor i of block b (b.kind must be in KINDS).
template <class IO, bool kT0, int KINDS = kKindsAll>
__device__ __forceinline__ void factor_update(const DevGraph& g, const Block& b, int32_t i,
                                              const IO& io, float t, bool normalize, float& dl) {
  const int kind = b.kind;
  if constexpr ((KINDS & (1 << kEnumBin2)) != 0) {
    if (kind == kEnumBin2) {
      enum_bin2<IO, kT0>(g, b, i, io, t, dl);
      return;
    }
  }
  if constexpr ((KINDS & (1 << kEnumPair)) != 0) {
    if (kind == kEnumPair) {
      enum_pair<IO, kT0>(g, b, i, io, t, dl);
      return;
    }
  }
  if constexpr ((KINDS & (1 << kEnumGeneral)) != 0) {
    if (kind == kEnumGeneral) {
      enum_general<IO, kT0>(g, b, i, io, t, dl);
      return;
    }
  }
  if constexpr ((KINDS & ((1 << kOr) | (1 << kAnd))) != 0) {
    if (kind == kOr || kind == kAnd) {
      logical_factor<IO>(g, b.f0 + i, b.off, io, t, normalize, dl);
      return;
    }
  }
  if constexpr ((KINDS & (1 << kPool)) != 0) {
    if (kind == kPool) pool_factor<IO>(g, b.f0 + i, io, t, normalize, dl);
  }
}

// ---------------------------------------------------------------------------
// Variable side (pgmax/infer/bp.py:217): per variable state, the incoming
// ftov of its edges summed in edge order from 0, then + evidence. Variables
// with more than kVarChunk edges sum runs ("chunks") of kVarChunk
// consecutive edges first and add the chunk sums in order; the CPU oracle
// uses the same order. Loads are issued eight edges at a time so each thread
// keeps 8-16 memory requests in flight.
// ---------------------------------------------------------------------------

// acc[s] += m[adj[p] + s] for p in [a, z), in order, s < ns <= MAXS.
// even2 (with MAXS == 2): every edge's two states are one aligned float2.
template <int MAXS>
__device__ __forceinline__ void edge_run_sum(const int32_t* adj, int32_t a, int32_t z,
                                             const float* m, int32_t ns, float* acc,
                                             bool even2 = false) {
  if constexpr (MAXS == 2) {
    if (even2) {
      int32_t p = a;
      for (; p + 8 <= z; p += 8) {
        int32_t es[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) es[u] = adj[p + u];
        float2 x[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) x[u] = ld2(m + es[u]);
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          acc[0] += x[u].x;
          acc[1] += x[u].y;
        }
      }
      if (p + 4 <= z) {
        int32_t es[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) es[u] = adj[p + u];
        float2 x[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) x[u] = ld2(m + es[u]);
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          acc[0] += x[u].x;
          acc[1] += x[u].y;
        }
        p += 4;
      }
      for (; p < z; ++p) {
        const float2 x = ld2(m + adj[p]);
        acc[0] += x.x;
        acc[1] += x.y;
      }
      return;
    }
  }
  int32_t p = a;
  for (; p + 8 <= z; p += 8) {
    int32_t es[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) es[u] = adj[p + u];
    float x[8][MAXS];
#pragma unroll
    for (int u = 0; u < 8; ++u)
#pragma unroll
      for (int s = 0; s < MAXS; ++s)
        if (s < ns) x[u][s] = m[es[u] + s];
#pragma unroll
    for (int u = 0; u < 8; ++u)
#pragma unroll
      for (int s = 0; s < MAXS; ++s)
        if (s < ns) acc[s] += x[u][s];
  }
  if (p + 4 <= z) {
    int32_t es[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) es[u] = adj[p + u];
    float x[4][MAXS];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int s = 0; s < MAXS; ++s)
        if (s < ns) x[u][s] = m[es[u] + s];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int s = 0; s < MAXS; ++s)
        if (s < ns) acc[s] += x[u][s];
    p += 4;
  }
  for (; p < z; ++p) {
    const int32_t es = adj[p];
#pragma unroll
    for (int s = 0; s < MAXS; ++s)
      if (s < ns) acc[s] += m[es + s];
  }
}

template <int MAXS>
__device__ __forceinline__ void light_var_t(const DevGraph& g, int32_t v, int32_t vs0,
                                            int32_t ns, const float* m, const float* ev,
                                            float* out) {
  float acc[MAXS];
#pragma unroll
  for (int s = 0; s < MAXS; ++s) acc[s] = 0.f;
  edge_run_sum<MAXS>(g.var_adj_es, g.var_eptr[v], g.var_eptr[v + 1], m, ns, acc, g.even2 != 0);
  if constexpr (MAXS == 2) {
    if (g.even2) {
      const float2 e2 = ld2(ev + vs0);
      st2(out + vs0, acc[0] + e2.x, acc[1] + e2.y);
      return;
    }
  }
#pragma unroll
  for (int s = 0; s < MAXS; ++s)
    if (s < ns) out[vs0 + s] = acc[s] + ev[vs0 + s];
}

// Sum of the binary messages at positions q (in order, from 0; -1 ends).
__device__ __forceinline__ float2 pos4_sum2(int4 q, const float* m) {
  float a0 = 0.f, a1 = 0.f;
  const int32_t p[4] = {q.x, q.y, q.z, q.w};
  float2 x[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) x[u] = p[u] >= 0 ? ld2(m + p[u]) : make_float2(0.f, 0.f);
#pragma unroll
  for (int u = 0; u < 4; ++u)
    if (p[u] >= 0) {
      a0 += x[u].x;
      a1 += x[u].y;
    }
  return make_float2(a0, a1);
}

// pos4_sum2 on one-float messages (pair_encode) at index position / 2.
__device__ __forceinline__ float2 pos4_sum2_packed(int4 q, const float* m) {
  float a0 = 0.f, a1 = 0.f;
  const int32_t p[4] = {q.x, q.y, q.z, q.w};
  float x[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) x[u] = p[u] >= 0 ? m[p[u] >> 1] : 0.f;
#pragma unroll
  for (int u = 0; u < 4; ++u)
    if (p[u] >= 0) {
      const float2 r = pair_decode(x[u]);
      a0 += r.x;
      a1 += r.y;
    }
  return make_float2(a0, a1);
}

// Per-state sums of the messages of edges at positions q (ns <= MAXS states
// each; in edge order, from 0; -1 ends), all loads issued first.
template <int MAXS>
__device__ __forceinline__ void pos4_sums(int4 q, int32_t ns, const float* m, float* acc) {
  const int32_t p[4] = {q.x, q.y, q.z, q.w};
  float x[4][MAXS];
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int st = 0; st < MAXS; ++st) x[u][st] = (p[u] >= 0 && st < ns) ? m[p[u] + st] : 0.f;
#pragma unroll
  for (int st = 0; st < MAXS; ++st) acc[st] = 0.f;
#pragma unroll
  for (int u = 0; u < 4; ++u)
    if (p[u] >= 0)
#pragma unroll
      for (int st = 0; st < MAXS; ++st)
        if (st < ns) acc[st] += x[u][st];
}

// pos4_sums for exactly NS states, each edge's states in vector loads; the
// evidence is added and the NS sums stored in the same way.
template <int NS>
__device__ __forceinline__ void pos4_var_x(int4 q, const float* m, const float* ev,
                                           float* out) {
  const int32_t p[4] = {q.x, q.y, q.z, q.w};
  float x[4][NS];
#pragma unroll
  for (int u = 0; u < 4; ++u)
    if (p[u] >= 0) ldv<NS>(m + p[u], x[u]);
  float acc[NS], e[NS];
  ldv<NS>(ev, e);
#pragma unroll
  for (int st = 0; st < NS; ++st) acc[st] = 0.f;
#pragma unroll
  for (int u = 0; u < 4; ++u)
    if (p[u] >= 0)
#pragma unroll
      for (int st = 0; st < NS; ++st) acc[st] += x[u][st];
#pragma unroll
  for (int st = 0; st < NS; ++st) acc[st] = acc[st] + e[st];
  stv<NS>(out, acc);
}

// One variable of degree <= kVarChunk (returns for heavier ones).
__device__ __forceinline__ void var_sums(const DevGraph& g, int32_t v, const float* m,
                                         const float* ev, float* out) {
  if (g.var_pos4) {  // variables of <= 4 edges: one index load
    const int4 q = g.var_pos4[v];
    const int32_t vs0 = g.var_vs0[v];
    const int32_t ns = g.var_vs0[v + 1] - vs0;
    if (q.x != -2 && ns == 2 && g.even2) {
      const float2 acc = g.packed ? pos4_sum2_packed(q, m) : pos4_sum2(q, m);
      const float2 e2 = ld2(ev + vs0);
      st2(out + vs0, acc.x + e2.x, acc.y + e2.y);
      return;
    }
    if (q.x != -2 && ns <= 8 && !g.packed) {
      switch (ns) {
        case 3: return pos4_var_x<3>(q, m, ev + vs0, out + vs0);
        case 4: return pos4_var_x<4>(q, m, ev + vs0, out + vs0);
        case 5: return pos4_var_x<5>(q, m, ev + vs0, out + vs0);
        case 6: return pos4_var_x<6>(q, m, ev + vs0, out + vs0);
        case 7: return pos4_var_x<7>(q, m, ev + vs0, out + vs0);
        case 8: return pos4_var_x<8>(q, m, ev + vs0, out + vs0);
        default: break;
      }
      float acc[8];
      pos4_sums<8>(q, ns, m, acc);
#pragma unroll
      for (int st = 0; st < 8; ++st)
        if (st < ns) out[vs0 + st] = acc[st] + ev[vs0 + st];
      return;
    }
  }
  const int32_t p0 = g.var_eptr[v], p1 = g.var_eptr[v + 1];
  if (p1 - p0 > kVarChunk) return;
  const int32_t vs0 = g.var_vs0[v], ns = g.var_vs0[v + 1] - vs0;
  if (ns == 2) {
    light_var_t<2>(g, v, vs0, ns, m, ev, out);
  } else if (ns <= 8) {
    light_var_t<8>(g, v, vs0, ns, m, ev, out);
  } else {
    for (int32_t s = 0; s < ns; ++s) {
      float a = 0.f;
      for (int32_t p = p0; p < p1; ++p) a += m[g.var_adj_es[p] + s];
      out[vs0 + s] = a + ev[vs0 + s];
    }
  }
}

template <int MAXS>
__device__ __forceinline__ void heavy_chunk_t(const DevGraph& g, int32_t a, int32_t z,
                                              int32_t ns, const float* m, float* part,
                                              int32_t stride) {
  float acc[MAXS];
#pragma unroll
  for (int s = 0; s < MAXS; ++s) acc[s] = 0.f;
  if constexpr (MAXS == 2) {
    if (g.even2 && z - a == kVarChunk) {
      // a full chunk of binary edges: all loads in flight at once, then the
      // same in-order sum; evenly spaced chunks need no position loads
      int32_t es[kVarChunk];
      if (stride > 0) {
        const int32_t p0 = g.var_adj_es[a];
#pragma unroll
        for (int u = 0; u < kVarChunk; ++u) es[u] = p0 + u * stride;
      } else {
#pragma unroll
        for (int u = 0; u < kVarChunk; ++u) es[u] = g.var_adj_es[a + u];
      }
      float2 x[kVarChunk];
#pragma unroll
      for (int u = 0; u < kVarChunk; ++u) x[u] = ld2(m + es[u]);
#pragma unroll
      for (int u = 0; u < kVarChunk; ++u) {
        acc[0] += x[u].x;
        acc[1] += x[u].y;
      }
      part[0] = acc[0];
      part[1] = acc[1];
      return;
    }
  }
  edge_run_sum<MAXS>(g.var_adj_es, a, z, m, ns, acc, g.even2 != 0);
#pragma unroll
  for (int s = 0; s < MAXS; ++s)
    if (s < ns) part[s] = acc[s];
}

// Sum of one chunk (descriptor d) of a heavy variable into `partial`.
__device__ __forceinline__ void heavy_chunk_d(const DevGraph& g, int4 d, const float* m,
                                             float* partial) {
  const int32_t a = d.x, len = d.y & 0xff, ns = d.y >> 8;
  float* part = partial + d.z;
  if (ns == 2) {
    heavy_chunk_t<2>(g, a, a + len, ns, m, part, d.w);
  } else if (ns <= 8) {
    heavy_chunk_t<8>(g, a, a + len, ns, m, part, 0);
  } else {
    for (int32_t s = 0; s < ns; ++s) {
      float c = 0.f;
      for (int32_t p = a; p < a + len; ++p) c += m[g.var_adj_es[p] + s];
      part[s] = c;
    }
  }
}

// Chunk sum number hc of the heavy variables into `partial`.
__device__ __forceinline__ void heavy_chunk(const DevGraph& g, int32_t hc, const float* m,
                                           float* partial) {
  heavy_chunk_d(g, g.chunk_desc[hc], m, partial);
}

// Sum from 0 of the pairs base[2k], base[2k + 1] for k in [k0, k1), both
// states at once with 8 loads in flight.
__device__ __forceinline__ void sum_range2(const float* base, int32_t k0, int32_t k1, float& o0,
                                           float& o1) {
  float a0 = 0.f, a1 = 0.f;
  int32_t k = k0;
  for (; k + 8 <= k1; k += 8) {
    float x0[8], x1[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      x0[u] = base[2 * (k + u)];
      x1[u] = base[2 * (k + u) + 1];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      a0 += x0[u];
      a1 += x1[u];
    }
  }
  for (; k < k1; ++k) {
    a0 += base[2 * k];
    a1 += base[2 * k + 1];
  }
  o0 = a0;
  o1 = a1;
}

// Message total of a binary heavy variable from its nch chunk sums (chunk k
// at part[2k], part[2k + 1]): groups of kVarChunk chunk sums, each from 0,
// then the group sums from 0 -- the engine's order for long sums.
__device__ __forceinline__ void chunk_total2(const float* part, int32_t nch, float& t0,
                                             float& t1) {
  const int32_t ng = (nch + kVarChunk - 1) / kVarChunk;
  t0 = 0.f;
  t1 = 0.f;
  for (int32_t j = 0; j < ng; ++j) {
    float g0, g1;
    sum_range2(part, j * kVarChunk, min(nch, (j + 1) * kVarChunk), g0, g1);
    t0 += g0;
    t1 += g1;
  }
}

// Heavy variable h: chunk sums in order, then + evidence.
__device__ __forceinline__ void heavy_combine(const DevGraph& g, int32_t h,
                                             const float* partial, const float* ev,
                                             float* out, bool groups_ready = false) {
  // chunk sums are added in groups of kVarChunk (each from 0), then the
  // group sums in order (from 0), then the evidence; with groups_ready the
  // group sums were written by k_var_groups
  const int4 hd = g.heavy_desc[h];
  const int32_t vs0 = hd.x, ns = hd.y >> 16;
  const int32_t nch = hd.y & 0xffff;
  const int32_t ng = (nch + kVarChunk - 1) / kVarChunk;
  const float* part = partial + hd.z;
  if (ns == 2) {
    float t0, t1;
    if (groups_ready && ng > 1) {
      sum_range2(partial + hd.w, 0, ng, t0, t1);
    } else {
      chunk_total2(part, nch, t0, t1);
    }
    out[vs0] = t0 + ev[vs0];
    out[vs0 + 1] = t1 + ev[vs0 + 1];
    return;
  }
  for (int32_t s = 0; s < ns; ++s) {
    float tot = 0.f;
    if (groups_ready && ng > 1) {
      const float* gp = partial + hd.w;
      for (int32_t j = 0; j < ng; ++j) tot += gp[j * ns + s];
    } else {
      for (int32_t j = 0; j < ng; ++j) {
        float gs = 0.f;
        const int32_t k1 = min(nch, (j + 1) * kVarChunk);
        for (int32_t k = j * kVarChunk; k < k1; ++k) gs += part[k * ns + s];
        tot += gs;
      }
    }
    out[vs0 + s] = tot + ev[vs0 + s];
  }
}

// Row chunk i (edges at consecutive 2-state positions), summed by the 8
// lanes of a group: lane q loads edges 4q..4q+3 (the group reads the chunk's
// 256 bytes together) into LDS, then lane 0 of the group adds the 32 values
// in edge order, so the sum equals heavy_chunk's. Every thread of the
// workgroup must call it (it synchronises the workgroup); i < 0 marks a
// lane group without a chunk. stage holds kRowStride floats per group.
constexpr int kRowLanes = 8;
constexpr int kRowStride = 2 * kVarChunk + 2;  // floats per group in LDS (padded)
__device__ __forceinline__ void heavy_chunk_row(const DevGraph& g, int32_t i, const float* m,
                                               float* partial, int32_t q, float* stage) {
  int32_t len = 0, slot = 0;
  if (i >= 0) {
    const int4 d = g.row_desc[i];
    len = d.y & 0xff;
    slot = d.z;
    const float* src = m + d.x;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int32_t idx = 4 * q + u;
      if (idx < len) {
        if (g.even2) {
          const float2 t = ld2(src + 2 * idx);
          stage[2 * idx] = t.x;
          stage[2 * idx + 1] = t.y;
        } else {
          stage[2 * idx] = src[2 * idx];
          stage[2 * idx + 1] = src[2 * idx + 1];
        }
      }
    }
  }
  __syncthreads();
  if (i >= 0 && q == 0) {
    float s0 = 0.f, s1 = 0.f;
    for (int32_t idx = 0; idx < len; ++idx) {
      s0 += stage[2 * idx];
      s1 += stage[2 * idx + 1];
    }
    partial[slot] = s0;
    partial[slot + 1] = s1;
  }
}

// Group sum number gi of a variable with more than kVarChunk chunks.
__device__ __forceinline__ void heavy_group(const DevGraph& g, int32_t gi, float* partial) {
  const int32_t h = g.vh_group_h[gi], j = g.vh_group_j[gi];
  const int32_t v = g.heavy_vars[h];
  const int32_t ns = g.var_vs0[v + 1] - g.var_vs0[v];
  const int32_t nch = g.heavy_cptr[h + 1] - g.heavy_cptr[h];
  const float* part = partial + g.heavy_pofs[h];
  const int32_t k1 = min(nch, (j + 1) * kVarChunk);
  for (int32_t s = 0; s < ns; ++s) {
    float gs = 0.f;
    for (int32_t k = j * kVarChunk; k < k1; ++k) gs += part[k * ns + s];
    partial[g.heavy_gpofs[h] + j * ns + s] = gs;
  }
}

// Low-degree pairwise graphs: the factor side recomputes the variable sums
// it needs from the previous messages (edge_run_sum, the same order as the
// variable phase, so vtof is bit-identical) and no variable phase runs.
constexpr int32_t kFuseMaxDeg = 8;

struct FusedIO : EngineIO {
  const int32_t* var_eptr;
  const int32_t* var_adj_es;
  const int32_t* edge_var;
  const float* ev;
  __device__ __forceinline__ void edge_vt2(int32_t e, int32_t, int32_t es, float* x) const {
    edge_vt<2>(e, es, 2, x);
  }
  // shadows EngineIO::edge_vt2o, which reads stored variable sums
  __device__ __forceinline__ void edge_vt2o(int32_t e, int32_t vs0, int32_t es, float* x,
                                            float* o) const {
    edge_vt2(e, vs0, es, x);
    old2(es, o);
  }
  const int4* edge_pos4;  // null: the CSR path only
  // shadows EngineIO::edge_vtx, which reads stored variable sums
  template <int S>
  __device__ __forceinline__ void edge_vtx(int32_t e, int32_t es, float* x, float* o) const {
    edge_vt<S>(e, es, S, x);
    ldv<S>(mo + es, o);
  }
  template <int MAXS>
  __device__ __forceinline__ void edge_vt(int32_t e, int32_t es, int32_t ns, float* x) const {
    if constexpr (MAXS == 2) {
      if (edge_pos4 && even2) {  // binary variable of <= 4 edges: one index load
        const int4 q = edge_pos4[e];
        if (q.x != -2) {
          const float2 acc = packed ? pos4_sum2_packed(q, mo) : pos4_sum2(q, mo);
          const float2 e2 = ld2(ev + edge_vs0[e]);
          const float2 o = packed ? pair_decode(mo[es >> 1]) : ld2(mo + es);
          x[0] = (acc.x + e2.x) - o.x;
          x[1] = (acc.y + e2.y) - o.y;
          return;
        }
      }
    }
    const int32_t v = edge_var[e];
    const int32_t vs0 = edge_vs0[e];
    float acc[MAXS];
#pragma unroll
    for (int s = 0; s < MAXS; ++s) acc[s] = 0.f;
    edge_run_sum<MAXS>(var_adj_es, var_eptr[v], var_eptr[v + 1], mo, ns, acc, even2);
    if constexpr (MAXS == 2) {
      if (even2) {
        const float2 e2 = ld2(ev + vs0), o = ld2(mo + es);
        x[0] = (acc[0] + e2.x) - o.x;
        x[1] = (acc[1] + e2.y) - o.y;
        return;
      }
    }
#pragma unroll
    for (int s = 0; s < MAXS; ++s)
      if (s < ns) x[s] = (acc[s] + ev[vs0 + s]) - mo[es + s];
  }
};

template <bool kT0, int KINDS>
__device__ __forceinline__ void fused_factor_update(const DevGraph& g, const Block& b, int32_t i,
                                                    const FusedIO& io, float t, float& dl) {
  if constexpr (KINDS == kKindsBin2) {
    enum_bin2<FusedIO, kT0>(g, b, i, io, t, dl);
  } else {
    if (b.kind == kEnumBin2)
      enum_bin2<FusedIO, kT0>(g, b, i, io, t, dl);
    else
      enum_pair<FusedIO, kT0>(g, b, i, io, t, dl);
  }
}

}  // namespace pgx

#endif  // PGX_DEVICE_H_
