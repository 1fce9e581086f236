// Reverse mode of BP (pgx_run_vjp): the product of a cotangent of the final
// messages of pgx_run with the run's Jacobian, for graphs whose factors are
// all Enum factors -- what jax.grad through BP.run computes in the
// reference's learning example (examples/grid_mrf.ipynb cells 15-16).
//
// The forward run is replayed one iteration at a time by pgx_run and kept
// (m_0 .. m_n, reference layout); the reverse pass then walks the
// iterations backwards, recomputing each one's variable sums
// (pgx_get_beliefs) and factor-to-variable messages (pgx_fac_to_var) with
// the engine's own arithmetic, and pulls the cotangent back through
//   normalise + clip     y = max(x - max_edge(x), -1e32)    (bp.py:249-259)
//   damping              x = d m + (1 - d) f                (bp.py:127)
//   Enum update          f_k = LSE_{c ∋ k}(s_c) - vt_k      (enum.py:451-474)
//   summaries            s_c = clip(lp_c) + sum_rows vt     (enum.py:451-460)
//   variable sums        vt = (ev + sum m)[var state] - m   (bp.py:217-218)
// A maximum attained by several entries shares its derivative equally among
// them; max(a, b) and the clip pass half at equality.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/pgx.h"
#include "pgx_internal.h"
#include "pgx_launch.h"
#include "pgx_math.h"
#include "pgx_plan.h"

struct pgx_vjp {
  pgx_graph* graph = nullptr;
  int64_t num_es = 0, num_vs = 0, num_lp = 0, num_edges = 0, num_rows = 0;
  int64_t num_enum_es = 0;  // the Enum edge states come first
  // reference-layout tables (device)
  int32_t* vsi = nullptr;       // [E_s] variable state of each edge state
  int32_t* edge_ptr = nullptr;  // [E + 1] first edge state of each edge
  int32_t* cfg_ptr = nullptr;   // [C + 1] first row of each config
  int32_t* row_es = nullptr;    // [R] edge state of each row
  int32_t* es_ptr = nullptr;    // [E_s + 1] first entry of each edge state in es_cfg
  int32_t* es_cfg = nullptr;    // [R] configs of each edge state, in row order
  int32_t* vs_ptr = nullptr;    // [V_s + 1] first entry of each variable state in vs_es
  int32_t* vs_es = nullptr;     // [E_s] edge states of each variable state, ascending
  int4* vs_pos4 = nullptr;      // [V_s] the same as four positions (-1 past the last), or
                                // null when a variable state has more than four
  // graphs of square pairwise factors (every factor two edges of S <= 8
  // states, canonical rows): the register path, one list of (first edge
  // state, first config) per S; pair = false otherwise
  bool pair = false;
  int64_t pair_count[9] = {0};
  int2* pair_desc[9] = {nullptr};
  // unary factors of such graphs (one edge of S states, config i on state
  // i), listed as (first edge state, first config) per S
  int64_t unary_count[9] = {0};
  int2* unary_desc[9] = {nullptr};
  // OR (index 0) and AND (1) factors (pgx_vjp_set_logical): parents of each
  // factor [F + 1], each listed parent's message index (absolute), each
  // factor's child edge state (absolute)
  bool logical_needed[2] = {false, false};
  int64_t logical_factors[2] = {0, 0};
  int32_t logical_off[2] = {1, -1};
  int32_t* logical_ptr[2] = {nullptr, nullptr};
  int32_t* logical_pes[2] = {nullptr, nullptr};
  int32_t* logical_ces[2] = {nullptr, nullptr};
  // Pool factors (pgx_vjp_set_pool): choices of each factor [F + 1], each
  // listed choice's message index (absolute), each indicator edge state
  bool pool_needed = false;
  int64_t pool_factors = 0;
  int32_t* pool_ptr = nullptr;
  int32_t* pool_pes = nullptr;
  int32_t* pool_ies = nullptr;
  float* scratch = nullptr;
  size_t scratch_bytes = 0;
};

namespace pgx {
namespace {

#define PGX_HIP(call)                                                              \
  do {                                                                             \
    hipError_t _e = (call);                                                        \
    if (_e != hipSuccess)                                                          \
      return set_error(PGX_EHIP, std::string(#call) + ": " + hipGetErrorString(_e)); \
  } while (0)

constexpr float kMsgFloor = -1e32f;
constexpr float kLpMax = 1e6f;

__device__ __forceinline__ float vmax(float a, float b) { return (a > b || a != a) ? a : b; }

dim3 grid_for(int64_t n, int32_t batch) {
  return dim3(unsigned(std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 8192))), batch);
}

// y = max(x - max_edge(x), -1e32), x = the caller's messages (stride per
// batch element): the entry normalisation of pgx_run (bp.py:92-96).
__global__ void k_norm_fwd(const int32_t* edge_ptr, int64_t num_edges, int64_t num_es,
                           const float* x, int64_t x_stride, float* y) {
  const int32_t b = blockIdx.y;
  const float* xb = x + b * x_stride;
  float* yb = y + b * num_es;
  for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < num_edges;
       e += int64_t(gridDim.x) * blockDim.x) {
    const int32_t a = edge_ptr[e], z = edge_ptr[e + 1];
    float mx = -__builtin_inff();
    for (int32_t i = a; i < z; ++i) mx = vmax(mx, xb[i]);
    for (int32_t i = a; i < z; ++i) yb[i] = vmax(xb[i] - mx, kMsgFloor);
  }
}

// Pull-back through y = max(x - max_edge(x), -1e32) for one edge; x is the
// damped message (kDamped: x = d m + (1 - d) f, writes gm = d gx, gf =
// (1 - d) gx) or the caller's initial message (writes gx to gm).
template <bool kDamped>
__global__ void k_norm_bwd(const int32_t* edge_ptr, int64_t num_edges, int64_t num_es,
                           const float* m, int64_t m_stride, const float* f, float d, float omd,
                           float* gm, float* gf) {
  const int32_t b = blockIdx.y;
  const float* mb = m + b * m_stride;
  const float* fb = kDamped ? f + b * num_es : nullptr;
  float* gmb = gm + b * num_es;
  float* gfb = kDamped ? gf + b * num_es : nullptr;
  for (int64_t e = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; e < num_edges;
       e += int64_t(gridDim.x) * blockDim.x) {
    const int32_t a = edge_ptr[e], z = edge_ptr[e + 1];
    auto xv = [&](int32_t i) { return kDamped ? d * mb[i] + omd * fb[i] : mb[i]; };
    float mx = -__builtin_inff();
    for (int32_t i = a; i < z; ++i) mx = vmax(mx, xv(i));
    float tot = 0.f, cnt = 0.f;
    for (int32_t i = a; i < z; ++i) {
      const float x = xv(i), y = x - mx;
      const float keep = y > kMsgFloor ? 1.f : (y == kMsgFloor ? 0.5f : 0.f);
      tot += gmb[i] * keep;
      cnt += (x == mx) ? 1.f : 0.f;
    }
    for (int32_t i = a; i < z; ++i) {
      const float x = xv(i), y = x - mx;
      const float keep = y > kMsgFloor ? 1.f : (y == kMsgFloor ? 0.5f : 0.f);
      const float gx = gmb[i] * keep - ((x == mx && cnt > 0.f) ? tot / cnt : 0.f);
      if (kDamped) {
        gmb[i] = d * gx;
        gfb[i] = omd * gx;
      } else {
        gmb[i] = gx;
      }
    }
  }
}

// vt = vs[var state] - m
__global__ void k_vtof(const int32_t* vsi, int64_t num_es, int64_t num_vs, const float* vs,
                       const float* m, float* vt) {
  const int32_t b = blockIdx.y;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < num_es;
       i += int64_t(gridDim.x) * blockDim.x)
    vt[b * num_es + i] = vs[b * num_vs + vsi[i]] - m[b * num_es + i];
}

// s_c = (sum of vt over the config's rows, from 0) + clip(lp_c): the order of
// the engine's Enum kernels.
__global__ void k_summaries(const int32_t* cfg_ptr, const int32_t* row_es, int64_t num_cfg,
                            int64_t num_es, const float* vt, const float* lp, int64_t lp_stride,
                            float* s) {
  const int32_t b = blockIdx.y;
  const float* vb = vt + b * num_es;
  const float* lb = lp + b * lp_stride;
  for (int64_t c = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; c < num_cfg;
       c += int64_t(gridDim.x) * blockDim.x) {
    float acc = 0.f;
    for (int32_t r = cfg_ptr[c]; r < cfg_ptr[c + 1]; ++r) acc += vb[row_es[r]];
    const float l = fminf(fmaxf(lb[c], -kLpMax), kLpMax);
    s[b * num_cfg + c] = acc + l;
  }
}

// Per edge state, over its configs: the max summary mx and, at T > 0,
// z = sum of exp((s - mx) / T) (the softmax weights are exp((s - mx) / T) /
// z); at T = 0 z counts the configs attaining mx.
template <bool kT0>
__global__ void k_es_stats(const int32_t* es_ptr, const int32_t* es_cfg, int64_t num_es,
                           int64_t num_cfg, const float* s, float t, float* mx_out,
                           float* z_out) {
  const int32_t b = blockIdx.y;
  const float* sb = s + b * num_cfg;
  for (int64_t k = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; k < num_es;
       k += int64_t(gridDim.x) * blockDim.x) {
    const int32_t a = es_ptr[k], z = es_ptr[k + 1];
    float mx = -__builtin_inff();
    for (int32_t p = a; p < z; ++p) mx = vmax(mx, sb[es_cfg[p]]);
    float acc = 0.f;
    if (mx > -__builtin_inff() && mx < __builtin_inff()) {
      for (int32_t p = a; p < z; ++p) {
        const float v = sb[es_cfg[p]];
        acc += kT0 ? (v == mx ? 1.f : 0.f) : pgx_exp((v - mx) / t);
      }
    }
    mx_out[b * num_es + k] = mx;
    z_out[b * num_es + k] = acc;
  }
}

// gs_c = sum over the config's rows r of gf[es_r] * w(c, es_r); glp_c +=
// gs_c * (derivative of the clip).
template <bool kT0>
__global__ void k_gsum(const int32_t* cfg_ptr, const int32_t* row_es, int64_t num_cfg,
                       int64_t num_es, const float* s, const float* mx, const float* zz,
                       const float* gf, const float* lp, int64_t lp_stride, float t, float* gs,
                       float* glp) {
  const int32_t b = blockIdx.y;
  const float* lb = lp + b * lp_stride;
  for (int64_t c = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; c < num_cfg;
       c += int64_t(gridDim.x) * blockDim.x) {
    const float sc = s[b * num_cfg + c];
    float acc = 0.f;
    for (int32_t r = cfg_ptr[c]; r < cfg_ptr[c + 1]; ++r) {
      const int64_t k = b * num_es + row_es[r];
      const float z = zz[k], m = mx[k];
      float w = 0.f;
      if (z > 0.f) w = kT0 ? (sc == m ? 1.f / z : 0.f) : pgx_exp((sc - m) / t) / z;
      acc += gf[k] * w;
    }
    gs[b * num_cfg + c] = acc;
    const float l = lb[c], al = fabsf(l);
    glp[b * num_cfg + c] += acc * (al < kLpMax ? 1.f : (al == kLpMax ? 0.5f : 0.f));
  }
}

// gvt_k = -gf_k + sum of gs over the configs of Enum edge state k (k <
// num_enum; the edge states of OR / AND factors follow them)
__global__ void k_gvtof(const int32_t* es_ptr, const int32_t* es_cfg, int64_t num_es,
                        int64_t num_enum, int64_t num_cfg, const float* gs, const float* gf,
                        float* gvt) {
  const int32_t b = blockIdx.y;
  for (int64_t k = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; k < num_enum;
       k += int64_t(gridDim.x) * blockDim.x) {
    float acc = -gf[b * num_es + k];
    for (int32_t p = es_ptr[k]; p < es_ptr[k + 1]; ++p) acc += gs[b * num_cfg + es_cfg[p]];
    gvt[b * num_es + k] = acc;
  }
}

// gvs_v = sum of gvt over the edge states of variable state v; gev += gvs
// Sum of x[es] over the edge states es of variable state v, in increasing
// order: from pos4 (up to four positions, -1 past the last) when the graph's
// variable states have at most four edge states each, else from the CSR.
__device__ __forceinline__ float vs_gather(const int4* pos4, const int32_t* vs_ptr,
                                           const int32_t* vs_es, int64_t v, const float* x) {
  float acc = 0.f;
  if (pos4) {
    const int4 q = pos4[v];
    const int32_t p[4] = {q.x, q.y, q.z, q.w};
    float y[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) y[k] = p[k] >= 0 ? x[p[k]] : 0.f;
#pragma unroll
    for (int k = 0; k < 4; ++k)
      if (p[k] >= 0) acc += y[k];
    return acc;
  }
  for (int32_t p = vs_ptr[v]; p < vs_ptr[v + 1]; ++p) acc += x[vs_es[p]];
  return acc;
}

__global__ void k_gvs(const int4* pos4, const int32_t* vs_ptr, const int32_t* vs_es,
                      int64_t num_vs, int64_t num_es, const float* gvt, float* gvs, float* gev) {
  const int32_t b = blockIdx.y;
  for (int64_t v = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; v < num_vs;
       v += int64_t(gridDim.x) * blockDim.x) {
    const float acc = vs_gather(pos4, vs_ptr, vs_es, v, gvt + b * num_es);
    gvs[b * num_vs + v] = acc;
    gev[b * num_vs + v] += acc;
  }
}

// gm += -gvt + gvs[var state]
__global__ void k_gmsg(const int32_t* vsi, int64_t num_es, int64_t num_vs, const float* gvt,
                       const float* gvs, float* gm) {
  const int32_t b = blockIdx.y;
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < num_es;
       i += int64_t(gridDim.x) * blockDim.x)
    gm[b * num_es + i] += gvs[b * num_vs + vsi[i]] - gvt[b * num_es + i];
}

// ---- OR / AND factors: the reverse of one factor update ---------------------
// The update's outputs (pgmax/factor/logical.py:561-779; the relevant slots,
// the other slots being the constant 0) as functions of the factor's vtof.
// One thread per (factor, batch element) recomputes the forward quantities
// in the engine's order (chunked sums, top-2 with ties to the last index),
// so every branch -- maxima, the eps test of a log-minus-exp, the stability
// maximum below T = 0.5 -- is decided as in the forward run, then pulls the
// cotangents of the outputs back onto the vtof of its parents and child.
// The derivative of a maximum is shared equally by the entries that attain
// it; max and min pass half at equality.

__device__ __forceinline__ float dmax_a(float a, float b) {
  return a > b ? 1.f : (a == b ? 0.5f : 0.f);
}
__device__ __forceinline__ float dmin0(float m) { return m < 0.f ? 1.f : (m == 0.f ? 0.5f : 0.f); }
// d logaddexp_t(a, b) / da (the partial in b is one minus it)
__device__ __forceinline__ float lae_da(float a, float b, float t) {
  return 1.f / (1.f + pgx_exp((b - a) / t));
}
// partials of logminusexp_t(a, b) (0 in its -inf branch)
__device__ __forceinline__ void lme_d(float a, float b, float t, float& da, float& db,
                                      float eps = 1e-4f) {
  if (a >= b + eps) {
    const float q = pgx_exp((b - a) / t);
    da = 1.f / (1.f - q);
    db = -q / (1.f - q);
  } else {
    da = 0.f;
    db = 0.f;
  }
}

// cotangents of the factor-wide quantities
struct LogicalG {
  float crel = 0.f, coth = 0.f, soth = 0.f, sacc = 0.f, m1 = 0.f, m2 = 0.f;
};

// Parent k of a factor of more than one parent: adds what its output's
// cotangent gu gives the factor-wide quantities to G; l_acc / l_oth get the
// parts that reach its own acc (max or log-add-exp of its two states) and
// other-state vtof directly.
template <bool kT0>
__device__ __forceinline__ void logical_parent_bwd(float t, bool is_a1, float rel, float oth,
                                                   float c_rel, float c_oth, float s_oth,
                                                   float s_acc, float m1, float m2, float gu,
                                                   LogicalG& G, float& l_acc, float& l_oth) {
  if (kT0) {
    const float par_rel = (s_acc + c_rel) - nmax(oth, rel);
    const float opt1 = (c_oth + s_oth) - oth;
    const float opt2 = par_rel + nmin(0.f, is_a1 ? m2 : m1);
    const float g_opt1 = -gu * dmax_a(opt1, opt2);
    const float g_opt2 = -gu * dmax_a(opt2, opt1);
    const float g_par_rel = gu + g_opt2;
    if (is_a1)
      G.m2 += g_opt2 * dmin0(m2);
    else
      G.m1 += g_opt2 * dmin0(m1);
    G.coth += g_opt1;
    G.soth += g_opt1;
    G.sacc += g_par_rel;
    G.crel += g_par_rel;
    l_oth = -g_opt1;
    l_acc = -g_par_rel;
  } else {
    const float lse = logaddexp_t(rel, oth, t);
    const float s_lse_wo = s_acc - lse, s_oth_wo = s_oth - oth;
    const float opt1 = c_oth + s_oth_wo, opt2 = c_rel + s_lse_wo, opt3 = c_rel + s_oth_wo;
    const float la12 = logaddexp_t(opt1, opt2, t);
    const float po0 = logminusexp_t(la12, opt3, t, 1e-4f);
    float pa = 1.f, pb = 0.f, lb_da = 0.f;
    if (t < kStabilityThre) {
      const float y = c_rel + (s_oth_wo + (is_a1 ? m2 : m1));
      const float lb = logaddexp_t(opt1, y, t);
      pa = dmax_a(po0, lb);
      pb = dmax_a(lb, po0);
      lb_da = lae_da(opt1, y, t);
    }
    const float g_po0 = -gu * pa, g_lb = -gu * pb;
    float me_da, me_db;
    lme_d(la12, opt3, t, me_da, me_db);
    const float g_la12 = g_po0 * me_da, g_opt3 = g_po0 * me_db;
    const float a12 = lae_da(opt1, opt2, t);
    const float g_opt1 = g_la12 * a12 + g_lb * lb_da;
    const float g_opt2 = g_la12 * (1.f - a12);
    const float g_x = g_lb * (1.f - lb_da);
    G.crel += g_x + g_opt2 + g_opt3 + gu;
    G.coth += g_opt1;
    const float g_soth_wo = g_x + g_opt1 + g_opt3;
    const float g_slse_wo = g_opt2 + gu;
    G.soth += g_soth_wo;
    G.sacc += g_slse_wo;
    if (is_a1)
      G.m2 += g_x;
    else
      G.m1 += g_x;
    l_oth = -g_soth_wo;
    l_acc = -g_slse_wo;
  }
}

// The child's output (relevant minus other) with cotangent gw.
template <bool kT0>
__device__ __forceinline__ void logical_child_bwd(float t, float s_oth, float s_acc, float m1,
                                                  float m2, float gw, LogicalG& G) {
  G.soth -= gw;
  if (kT0) {
    G.sacc += gw;
    G.m1 += gw * dmin0(m1);
  } else {
    const float cr0 = logminusexp_t(s_acc, s_oth, t, 1e-4f);
    float pa = 1.f, pb = 0.f, c_da = 0.f;
    if (t < kStabilityThre) {
      const float crs = logaddexp_t(s_oth + m1, s_oth + m2, t);
      pa = dmax_a(cr0, crs);
      pb = dmax_a(crs, cr0);
      c_da = lae_da(s_oth + m1, s_oth + m2, t);
    }
    const float g_cr0 = gw * pa, g_crs = gw * pb;
    float da, db;
    lme_d(s_acc, s_oth, t, da, db);
    G.sacc += g_cr0 * da;
    G.soth += g_cr0 * db + g_crs;
    G.m1 += g_crs * c_da;
    G.m2 += g_crs * (1.f - c_da);
  }
}

template <bool kT0>
__global__ void __launch_bounds__(256)
    k_logical_bwd(const int32_t* ptr, const int32_t* pes, const int32_t* ces, int64_t num_f,
                  int32_t off, int64_t num_es, const float* vt, const float* gf, float t,
                  float* gvt) {
  const int64_t f = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (f >= num_f) return;
  const int32_t b = blockIdx.y;
  const float* v = vt + b * num_es;
  const float* g = gf + b * num_es;
  float* gv = gvt + b * num_es;
  const int32_t a = ptr[f], z = ptr[f + 1], np = z - a;
  const int32_t c = ces[f];
  const float c_rel = v[c + off], c_oth = v[c];
  // forward reductions, as the engine forms them
  ChunkedSum so, sa;
  float m1 = kNegInf, m2 = kNegInf;
  int32_t a1 = -1;
  for (int32_t k = 0; k < np; ++k) {
    const int32_t e = pes[a + k];
    const float rel = v[e + off], oth = v[e];
    const float diff = rel - oth;
    so.add(oth);
    sa.add(kT0 ? nmax(oth, rel) : logaddexp_t(rel, oth, t));
    if (diff >= m1) {
      m2 = m1;
      m1 = diff;
      a1 = k;
    } else {
      m2 = nmax(m2, diff);
    }
  }
  const float s_oth = so.value(), s_acc = sa.value();
  // factor-wide cotangents, and the ties of m1 and m2
  LogicalG G;
  float cnt1 = 0.f, cnt2 = 0.f;
  for (int32_t k = 0; k < np; ++k) {
    const int32_t e = pes[a + k];
    const float rel = v[e + off], oth = v[e];
    cnt1 += (rel - oth == m1) ? 1.f : 0.f;
    cnt2 += (k != a1 && rel - oth == m2) ? 1.f : 0.f;
    if (np > 1) {
      float l_acc, l_oth;
      logical_parent_bwd<kT0>(t, k == a1, rel, oth, c_rel, c_oth, s_oth, s_acc, m1, m2,
                              g[e + off], G, l_acc, l_oth);
    }
  }
  if (np == 1) {  // the single parent's output is the child's vtof
    const float gu = g[pes[a] + off];
    G.crel += gu;
    G.coth -= gu;
  }
  logical_child_bwd<kT0>(t, s_oth, s_acc, m1, m2, g[c + off], G);
  // each parent: its own parts, the sums' and the top-2's shares
  for (int32_t k = 0; k < np; ++k) {
    const int32_t e = pes[a + k];
    const float rel = v[e + off], oth = v[e];
    float l_acc = 0.f, l_oth = 0.f;
    if (np > 1) {
      LogicalG scratch;
      logical_parent_bwd<kT0>(t, k == a1, rel, oth, c_rel, c_oth, s_oth, s_acc, m1, m2,
                              g[e + off], scratch, l_acc, l_oth);
    }
    const float g_acc = l_acc + G.sacc;
    float g_oth = l_oth + G.soth;
    float g_rel;
    if (kT0) {
      g_rel = g_acc * dmax_a(rel, oth);
      g_oth += g_acc * dmax_a(oth, rel);
    } else {
      const float da = lae_da(rel, oth, t);
      g_rel = g_acc * da;
      g_oth += g_acc * (1.f - da);
    }
    const float d = rel - oth;
    const float g_d = (d == m1 ? G.m1 / cnt1 : 0.f) +
                      ((k != a1 && d == m2 && cnt2 > 0.f) ? G.m2 / cnt2 : 0.f);
    gv[e + off] = g_rel + g_d;
    gv[e] = g_oth - g_d;
  }
  gv[c + off] = G.crel;
  gv[c] = G.coth;
}

// ---- Pool factors: the reverse of one factor update --------------------------
// The normalised outputs of pgmax/factor/pool.py:328-474 (each choice's and
// the indicator's state-1 slot; state 0 is the constant 0) as functions of
// the diffs d_k = one_k - zero_k of the choices and pid of the indicator.
// Same scheme as k_logical_bwd: the forward quantities in pool_factor's
// order, then the pull-back.
template <bool kT0>
__global__ void __launch_bounds__(256)
    k_pool_bwd(const int32_t* ptr, const int32_t* pes, const int32_t* ies, int64_t num_f,
               int64_t num_es, const float* vt, const float* gf, float t, float* gvt) {
  const int64_t f = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (f >= num_f) return;
  const int32_t b = blockIdx.y;
  const float* v = vt + b * num_es;
  const float* g = gf + b * num_es;
  float* gv = gvt + b * num_es;
  const int32_t a = ptr[f], z = ptr[f + 1], n = z - a;
  const int32_t ie = ies[f];
  const float pid = v[ie + 1] - v[ie];
  float m1 = kNegInf, m2 = kNegInf;
  int32_t a1 = -1;
  for (int32_t k = 0; k < n; ++k) {
    const int32_t e = pes[a + k];
    const float diff = v[e + 1] - v[e];
    if (diff >= m1) {
      m2 = m1;
      m1 = diff;
      a1 = k;
    } else {
      m2 = nmax(m2, diff);
    }
  }
  float pi_diff = m1, f_lse = 0.f, lse_ind = 0.f, fa = 0.f;
  if (!kT0) {
    const float mam = nmax(m2, -pid);
    float u1 = 0.f, u2 = 0.f;
    for (int32_t k = 0; k < n; ++k) {
      const int32_t e = pes[a + k];
      const float diff = v[e + 1] - v[e];
      u1 += exp_rel(diff - m1, t);
      u2 += exp_rel((k == a1 ? -pid : diff) - mam, t);
    }
    pi_diff = mul_t(pgx_log(u1), t) + m1;
    f_lse = logaddexp_t(pi_diff, -pid, t);
    lse_ind = mul_t(pgx_log(u2), t) + mam;
    fa = lae_da(pi_diff, -pid, t);
  }
  // factor-wide cotangents
  float g_pid = 0.f, g_w = g[ie + 1], g_flse = 0.f, g_m1 = 0.f, g_m2 = 0.f, cnt1 = 0.f, cnt2 = 0.f;
  const float gu_a1 = a1 >= 0 ? g[pes[a + a1] + 1] : 0.f;
  if (kT0) g_m1 = g_w;
  for (int32_t k = 0; k < n; ++k) {
    const int32_t e = pes[a + k];
    const float diff = v[e + 1] - v[e];
    const float gu = g[e + 1];
    cnt1 += diff == m1 ? 1.f : 0.f;
    cnt2 += (k != a1 && diff == m2) ? 1.f : 0.f;
    if (n == 1) {
      g_pid += gu;
    } else if (kT0) {
      const float mk = k == a1 ? m2 : m1;
      g_pid += gu * dmax_a(-pid, mk);      // d min(pid, -m) / d pid
      const float gm = -gu * dmax_a(mk, -pid);  // ... / d m
      if (k == a1)
        g_m2 += gm;
      else
        g_m1 += gm;
    } else if (k == a1) {
      g_pid += gu * pgx_exp((-pid - lse_ind) / t);
    } else {
      float da, db;
      lme_d(f_lse, diff, t, da, db, 1e-30f);
      g_flse -= gu * da;
    }
  }
  if (!kT0 && n > 1) {
    g_w += g_flse * fa;
    g_pid -= g_flse * (1.f - fa);
  }
  for (int32_t k = 0; k < n; ++k) {
    const int32_t e = pes[a + k];
    const float diff = v[e + 1] - v[e];
    float gd;
    if (kT0) {
      gd = (diff == m1 ? g_m1 / cnt1 : 0.f) +
           ((k != a1 && diff == m2 && cnt2 > 0.f) ? g_m2 / cnt2 : 0.f);
    } else {
      gd = g_w * pgx_exp((diff - pi_diff) / t);
      if (n > 1 && k != a1) {
        float da, db;
        lme_d(f_lse, diff, t, da, db, 1e-30f);
        gd -= g[e + 1] * db;
        gd -= gu_a1 * pgx_exp((diff - lse_ind) / t);
      }
    }
    gv[e + 1] = gd;
    gv[e] = -gd;
  }
  gv[ie + 1] = g_pid;
  gv[ie] = -g_pid;
}

// out[i] = sum over batch elements of in[b * n + i] (shared inputs), or a copy
__global__ void k_batch_sum(const float* in, int64_t n, int32_t batch, float* out) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    float acc = 0.f;
    for (int32_t b = 0; b < batch; ++b) acc += in[b * n + i];
    out[i] = acc;
  }
}

__global__ void k_zero(float* p, int64_t n) {
  for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    p[i] = 0.f;
}

// ---- pairwise graphs: one thread per (factor, batch element) ----------------
// Every factor joins two edges of S states: it owns edge states [es0, es0 +
// 2S) (edge a, then edge b) and configs [c0, c0 + S^2), config (i, j) = c0 +
// S i + j with rows (a, i), (b, j): the layout of PairwiseFactorGroup
// (pgmax/fgroup/enum.py:303-327). The factors of each S are listed as
// (es0, c0). A factor's summaries, statistics and cotangents then stay in
// registers, and an iteration of the run is two passes over memory instead
// of a chain of gathers.

// vs = evidence + the sum of the messages of each variable state
// (pgmax/infer/bp.py:217)
__global__ void k_var_sums(const int4* pos4, const int32_t* vs_ptr, const int32_t* vs_es,
                           int64_t num_vs, int64_t num_es, const float* ev, int64_t ev_stride,
                           const float* m, float* vs) {
  const int32_t b = blockIdx.y;
  const float* mb = m + b * num_es;
  for (int64_t v = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; v < num_vs;
       v += int64_t(gridDim.x) * blockDim.x)
    vs[b * num_vs + v] = ev[b * ev_stride + v] + vs_gather(pos4, vs_ptr, vs_es, v, mb);
}

// N consecutive floats, in 16-byte vector accesses when N is a multiple of 4
// and the address is 16-byte aligned (a factor's edge states and configs
// start at multiples of 2S and S^2)
template <int N>
__device__ __forceinline__ void ldn(const float* p, float* x) {
  if constexpr (N % 4 == 0) {
    if ((reinterpret_cast<uintptr_t>(p) & 15) == 0) {
#pragma unroll
      for (int k = 0; k < N; k += 4) {
        const float4 v = *reinterpret_cast<const float4*>(p + k);
        x[k] = v.x;
        x[k + 1] = v.y;
        x[k + 2] = v.z;
        x[k + 3] = v.w;
      }
      return;
    }
  }
#pragma unroll
  for (int k = 0; k < N; ++k) x[k] = p[k];
}
template <int N>
__device__ __forceinline__ void stn(float* p, const float* x) {
  if constexpr (N % 4 == 0) {
    if ((reinterpret_cast<uintptr_t>(p) & 15) == 0) {
#pragma unroll
      for (int k = 0; k < N; k += 4)
        *reinterpret_cast<float4*>(p + k) = make_float4(x[k], x[k + 1], x[k + 2], x[k + 3]);
      return;
    }
  }
#pragma unroll
  for (int k = 0; k < N; ++k) p[k] = x[k];
}

// x / T and T * x, skipped at T = 1 (exact identities there, so the same
// bits); the kernels below call their bodies with a literal 1 at T = 1.
__device__ __forceinline__ float tdiv(float x, float t) { return t == 1.f ? x : x / t; }
__device__ __forceinline__ float tmul(float x, float t) { return t == 1.f ? x : t * x; }

// One pairwise factor's forward quantities: vt = vs[var state] - m, the
// summaries s[i][j] = (vt_a[i] + vt_b[j]) + clip(lp), per row i and column
// j the maximum and z (T > 0: the sum of exp((s - max) / T); T = 0: the
// number of entries attaining it), and the damped messages x = d m + (1 - d)
// (out - vt) with out = T log z + max (max at T = 0) -- enum.py:451-475,
// update_utils.py:67-98, bp.py:127.
template <int S, bool kT0>
struct PairFactor {
  // S <= 4 (T > 0): keep every summary's two exponentials for the weights
  static constexpr bool kKeep = !kT0 && S <= 4;
  static constexpr int KS = kKeep ? S : 1;
  float m[2 * S], vt[2 * S], x[2 * S];
  float s[S][S];
  float mx[2 * S], z[2 * S];
  float er[KS][KS], ec[KS][KS];  // exp((s - row max) / T), exp((s - column max) / T)

  // kVec: 16-byte loads of the messages and the table where aligned
  template <bool kVec>
  __device__ __forceinline__ void load(const int32_t* vsi, int64_t es0, const float* vsb,
                                       const float* mb, const float* lf, float d, float omd,
                                       float t) {
    if (kVec) {
      ldn<2 * S>(mb + es0, m);
      ldn<S * S>(lf, &s[0][0]);
    } else {
#pragma unroll
      for (int k = 0; k < 2 * S; ++k) m[k] = mb[es0 + k];
#pragma unroll
      for (int k = 0; k < S * S; ++k) (&s[0][0])[k] = lf[k];
    }
#pragma unroll
    for (int k = 0; k < 2 * S; ++k) {
      vt[k] = vsb[vsi[es0 + k]] - m[k];
      mx[k] = -__builtin_inff();
      z[k] = 0.f;
    }
#pragma unroll
    for (int i = 0; i < S; ++i)
#pragma unroll
      for (int j = 0; j < S; ++j) {
        s[i][j] = (vt[i] + vt[S + j]) + fminf(fmaxf(s[i][j], -kLpMax), kLpMax);
        mx[i] = vmax(mx[i], s[i][j]);
        mx[S + j] = vmax(mx[S + j], s[i][j]);
      }
#pragma unroll
    for (int i = 0; i < S; ++i)
#pragma unroll
      for (int j = 0; j < S; ++j) {
        const float v = s[i][j];
        if (kT0) {
          z[i] += v == mx[i] ? 1.f : 0.f;
          z[S + j] += v == mx[S + j] ? 1.f : 0.f;
        } else {
          const float a = pgx_exp(tdiv(v - mx[i], t)), c = pgx_exp(tdiv(v - mx[S + j], t));
          z[i] += a;
          z[S + j] += c;
          if constexpr (kKeep) {
            er[i][j] = a;
            ec[i][j] = c;
          }
        }
      }
#pragma unroll
    for (int k = 0; k < 2 * S; ++k) {
      const bool fin = mx[k] > -__builtin_inff() && mx[k] < __builtin_inff();
      if (!fin) z[k] = 0.f;
      const float out = kT0 ? mx[k] : tmul(pgx_log(z[k]), t) + mx[k];
      x[k] = d * m[k] + omd * (out - vt[k]);
    }
  }

  // softmax (T > 0) or tie-shared (T = 0) weight of summary v = s[i][j] in
  // row i (col = false) or column j (col = true)
  __device__ __forceinline__ float weight(bool col, int i, int j, float v, float t) const {
    const int k = col ? S + j : i;
    if (!(z[k] > 0.f)) return 0.f;
    if (kT0) return v == mx[k] ? 1.f / z[k] : 0.f;
    float e;
    if constexpr (kKeep)
      e = col ? ec[i][j] : er[i][j];
    else
      e = pgx_exp(tdiv(v - mx[k], t));
    return e / z[k];
  }
};

// One iteration of the run: m_out = normalise(x) per edge (bp.py:249-259).
template <int S, bool kT0>
__device__ __forceinline__ void pair_fwd_body(const int32_t* vsi, const int2* desc, int64_t num_f,
                                              int64_t num_es, int64_t num_vs, const float* vs,
                                              const float* m, const float* lp, int64_t lp_stride,
                                              float d, float omd, float t, float* m_out) {
  const int64_t f = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (f >= num_f) return;
  const int32_t b = blockIdx.y;
  const int2 fd = desc[f];  // (first edge state, first config)
  PairFactor<S, kT0> q;
  q.template load<false>(vsi, fd.x, vs + b * num_vs, m + b * num_es, lp + b * lp_stride + fd.y,
                         d, omd, t);
  float* ob = m_out + b * num_es + fd.x;
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    float mx = -__builtin_inff();
#pragma unroll
    for (int k = 0; k < S; ++k) mx = vmax(mx, q.x[h * S + k]);
#pragma unroll
    for (int k = 0; k < S; ++k) ob[h * S + k] = vmax(q.x[h * S + k] - mx, kMsgFloor);
  }
}

template <int S, bool kT0>
__global__ void __launch_bounds__(256)
    k_pair_fwd(const int32_t* vsi, const int2* desc, int64_t num_f, int64_t num_es,
               int64_t num_vs, const float* vs, const float* m, const float* lp,
               int64_t lp_stride, float d, float omd, float t, float* m_out) {
  if (!kT0 && t == 1.f)
    pair_fwd_body<S, kT0>(vsi, desc, num_f, num_es, num_vs, vs, m, lp, lp_stride, d, omd, 1.f,
                          m_out);
  else
    pair_fwd_body<S, kT0>(vsi, desc, num_f, num_es, num_vs, vs, m, lp, lp_stride, d, omd, t, m_out);
}

// The reverse of one iteration for one factor: gm (in: the cotangent of
// this iteration's output) is pulled back through the normalisation (out:
// d gx), the log-sum-exps and the summaries; writes gvt (the cotangent of
// vt) and adds the log-potentials' share to glp. Same formulas and orders
// as k_norm_bwd, k_gsum and k_gvtof. With `fold`, the incoming cotangent is
// first completed with the previous reverse step's variable side, gm +=
// gvs[var state] - gvt (k_gmsg), read where this thread reads gm anyway.
template <int S, bool kT0>
__device__ __forceinline__ void pair_bwd_body(const int32_t* vsi, const int2* desc, int64_t num_f,
                                              int64_t num_es, int64_t num_vs, int64_t num_cfg,
                                              const float* vs, const float* m, const float* lp,
                                              int64_t lp_stride, float d, float omd, float t,
                                              float* gm, float* gvt, const float* gvs, int32_t fold,
                                              float* glp) {
  const int64_t f = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (f >= num_f) return;
  const int32_t b = blockIdx.y;
  const int2 fd = desc[f];  // (first edge state, first config)
  const float* lf = lp + b * lp_stride + fd.y;
  PairFactor<S, kT0> q;
  q.template load<true>(vsi, fd.x, vs + b * num_vs, m + b * num_es, lf, d, omd, t);
  float* gmb = gm + b * num_es + fd.x;
  float* gvb = gvt + b * num_es + fd.x;
  float gin[2 * S], gmo[2 * S];
  ldn<2 * S>(gmb, gin);
  if (fold) {
    float gprev[2 * S];
    ldn<2 * S>(gvb, gprev);
#pragma unroll
    for (int k = 0; k < 2 * S; ++k) gin[k] += gvs[b * num_vs + vsi[fd.x + k]] - gprev[k];
  }
  float gf[2 * S], gv[2 * S];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    float mx = -__builtin_inff();
#pragma unroll
    for (int k = 0; k < S; ++k) mx = vmax(mx, q.x[h * S + k]);
    float tot = 0.f, cnt = 0.f, g[S];
#pragma unroll
    for (int k = 0; k < S; ++k) {
      const float x = q.x[h * S + k], y = x - mx;
      const float keep = y > kMsgFloor ? 1.f : (y == kMsgFloor ? 0.5f : 0.f);
      g[k] = gin[h * S + k] * keep;
      tot += g[k];
      cnt += (x == mx) ? 1.f : 0.f;
    }
#pragma unroll
    for (int k = 0; k < S; ++k) {
      const float gx = g[k] - ((q.x[h * S + k] == mx && cnt > 0.f) ? tot / cnt : 0.f);
      gmo[h * S + k] = d * gx;
      gf[h * S + k] = omd * gx;
      gv[h * S + k] = -gf[h * S + k];
    }
  }
  stn<2 * S>(gmb, gmo);
  float* gl = glp + b * num_cfg + fd.y;
  // the log-potential cotangents a row (R = S) or, for small tables, the
  // whole table (R = S^2) at a time: registers for R of them
  constexpr int R = S <= 4 ? S * S : S;
#pragma unroll
  for (int r0 = 0; r0 < S * S; r0 += R) {
    float gla[R], la[R];
    ldn<R>(gl + r0, gla);
    ldn<R>(lf + r0, la);
#pragma unroll
    for (int c = 0; c < R; ++c) {
      const int i = (r0 + c) / S, j = (r0 + c) % S;
      const float v = q.s[i][j];
      float gs = 0.f;
      gs += gf[i] * q.weight(false, i, j, v, t);
      gs += gf[S + j] * q.weight(true, i, j, v, t);
      const float al = fabsf(la[c]);
      gla[c] += gs * (al < kLpMax ? 1.f : (al == kLpMax ? 0.5f : 0.f));
      gv[i] += gs;
      gv[S + j] += gs;
    }
    stn<R>(gl + r0, gla);
  }
  stn<2 * S>(gvb, gv);
}

template <int S, bool kT0>
__global__ void __launch_bounds__(256)
    k_pair_bwd(const int32_t* vsi, const int2* desc, int64_t num_f, int64_t num_es,
               int64_t num_vs, int64_t num_cfg, const float* vs, const float* m,
               const float* lp, int64_t lp_stride, float d, float omd, float t, float* gm,
               float* gvt, const float* gvs, int32_t fold, float* glp) {
  if (!kT0 && t == 1.f)
    pair_bwd_body<S, kT0>(vsi, desc, num_f, num_es, num_vs, num_cfg, vs, m, lp, lp_stride, d, omd,
                          1.f, gm, gvt, gvs, fold, glp);
  else
    pair_bwd_body<S, kT0>(vsi, desc, num_f, num_es, num_vs, num_cfg, vs, m, lp, lp_stride, d, omd,
                          t, gm, gvt, gvs, fold, glp);
}

// Unary Enum factors (one edge of S states, config i on state i): the
// summary of state i is vt_i + clip(lp_i), the only config of its state, so
// the log-sum-exp is that summary itself (max, and z = exp(0) = 1 or one
// tie), with the same finiteness rule as PairFactor. Forward: one iteration
// of the factor's edge; reverse: as k_pair_bwd for one edge.
template <int S, bool kT0>
struct UnaryFactor {
  float m[S], vt[S], x[S], sm[S], z[S];
  __device__ __forceinline__ void load(const int32_t* vsi, int64_t es0, const float* vsb,
                                       const float* mb, const float* lf, float d, float omd,
                                       float t) {
#pragma unroll
    for (int k = 0; k < S; ++k) {
      m[k] = mb[es0 + k];
      vt[k] = vsb[vsi[es0 + k]] - m[k];
      sm[k] = vt[k] + fminf(fmaxf(lf[k], -kLpMax), kLpMax);
      const bool fin = sm[k] > -__builtin_inff() && sm[k] < __builtin_inff();
      z[k] = fin ? (kT0 ? 1.f : pgx_exp(tdiv(sm[k] - sm[k], t))) : 0.f;
      const float out = kT0 ? sm[k] : tmul(pgx_log(z[k]), t) + sm[k];
      x[k] = d * m[k] + omd * (out - vt[k]);
    }
  }
};

template <int S, bool kT0>
__device__ __forceinline__ void unary_fwd_body(const int32_t* vsi, const int2* desc, int64_t num_f,
                                               int64_t num_es, int64_t num_vs, const float* vs,
                                               const float* m, const float* lp, int64_t lp_stride,
                                               float d, float omd, float t, float* m_out) {
  const int64_t f = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (f >= num_f) return;
  const int32_t b = blockIdx.y;
  const int2 fd = desc[f];
  UnaryFactor<S, kT0> q;
  q.load(vsi, fd.x, vs + b * num_vs, m + b * num_es, lp + b * lp_stride + fd.y, d, omd, t);
  float mx = -__builtin_inff();
#pragma unroll
  for (int k = 0; k < S; ++k) mx = vmax(mx, q.x[k]);
  float* ob = m_out + b * num_es + fd.x;
#pragma unroll
  for (int k = 0; k < S; ++k) ob[k] = vmax(q.x[k] - mx, kMsgFloor);
}

template <int S, bool kT0>
__global__ void __launch_bounds__(256)
    k_unary_fwd(const int32_t* vsi, const int2* desc, int64_t num_f, int64_t num_es,
                int64_t num_vs, const float* vs, const float* m, const float* lp,
                int64_t lp_stride, float d, float omd, float t, float* m_out) {
  if (!kT0 && t == 1.f)
    unary_fwd_body<S, kT0>(vsi, desc, num_f, num_es, num_vs, vs, m, lp, lp_stride, d, omd, 1.f,
                           m_out);
  else
    unary_fwd_body<S, kT0>(vsi, desc, num_f, num_es, num_vs, vs, m, lp, lp_stride, d, omd, t,
                           m_out);
}

template <int S, bool kT0>
__device__ __forceinline__ void unary_bwd_body(const int32_t* vsi, const int2* desc, int64_t num_f,
                                               int64_t num_es, int64_t num_vs, int64_t num_cfg,
                                               const float* vs, const float* m, const float* lp,
                                               int64_t lp_stride, float d, float omd, float t,
                                               float* gm, float* gvt, const float* gvs,
                                               int32_t fold, float* glp) {
  const int64_t f = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (f >= num_f) return;
  const int32_t b = blockIdx.y;
  const int2 fd = desc[f];
  const float* lf = lp + b * lp_stride + fd.y;
  UnaryFactor<S, kT0> q;
  q.load(vsi, fd.x, vs + b * num_vs, m + b * num_es, lf, d, omd, t);
  float* gmb = gm + b * num_es + fd.x;
  float* gvb = gvt + b * num_es + fd.x;
  float* gl = glp + b * num_cfg + fd.y;
  float g[S];
  float mx = -__builtin_inff();
#pragma unroll
  for (int k = 0; k < S; ++k) {
    g[k] = gmb[k];
    if (fold) g[k] += gvs[b * num_vs + vsi[fd.x + k]] - gvb[k];
    mx = vmax(mx, q.x[k]);
  }
  float tot = 0.f, cnt = 0.f;
#pragma unroll
  for (int k = 0; k < S; ++k) {
    const float y = q.x[k] - mx;
    g[k] = g[k] * (y > kMsgFloor ? 1.f : (y == kMsgFloor ? 0.5f : 0.f));
    tot += g[k];
    cnt += (q.x[k] == mx) ? 1.f : 0.f;
  }
#pragma unroll
  for (int k = 0; k < S; ++k) {
    const float gx = g[k] - ((q.x[k] == mx && cnt > 0.f) ? tot / cnt : 0.f);
    gmb[k] = d * gx;
    const float gf = omd * gx;
    const float w =
        q.z[k] > 0.f ? (kT0 ? 1.f / q.z[k] : pgx_exp(tdiv(q.sm[k] - q.sm[k], t)) / q.z[k]) : 0.f;
    float gs = 0.f;
    gs += gf * w;
    const float l = lf[k], al = fabsf(l);
    gl[k] += gs * (al < kLpMax ? 1.f : (al == kLpMax ? 0.5f : 0.f));
    gvb[k] = -gf + gs;
  }
}

template <int S, bool kT0>
__global__ void __launch_bounds__(256)
    k_unary_bwd(const int32_t* vsi, const int2* desc, int64_t num_f, int64_t num_es,
                int64_t num_vs, int64_t num_cfg, const float* vs, const float* m,
                const float* lp, int64_t lp_stride, float d, float omd, float t, float* gm,
                float* gvt, const float* gvs, int32_t fold, float* glp) {
  if (!kT0 && t == 1.f)
    unary_bwd_body<S, kT0>(vsi, desc, num_f, num_es, num_vs, num_cfg, vs, m, lp, lp_stride, d, omd,
                           1.f, gm, gvt, gvs, fold, glp);
  else
    unary_bwd_body<S, kT0>(vsi, desc, num_f, num_es, num_vs, num_cfg, vs, m, lp, lp_stride, d, omd,
                           t, gm, gvt, gvs, fold, glp);
}

template <int S>
void launch_unary(bool fwd, bool t0, int32_t batch, hipStream_t s, const int32_t* vsi,
                  const int2* desc, int64_t num_f, int64_t num_es, int64_t num_vs,
                  int64_t num_cfg, const float* vs, const float* m, const float* lp,
                  int64_t lp_stride, float d, float omd, float t, float* out, float* gvt,
                  const float* gvs, int32_t fold, float* glp) {
  const dim3 grid(unsigned((num_f + 255) / 256), batch);
  if (fwd && t0)
    hipLaunchKernelGGL((k_unary_fwd<S, true>), grid, dim3(256), 0, s, vsi, desc, num_f, num_es,
                       num_vs, vs, m, lp, lp_stride, d, omd, t, out);
  else if (fwd)
    hipLaunchKernelGGL((k_unary_fwd<S, false>), grid, dim3(256), 0, s, vsi, desc, num_f,
                       num_es, num_vs, vs, m, lp, lp_stride, d, omd, t, out);
  else if (t0)
    hipLaunchKernelGGL((k_unary_bwd<S, true>), grid, dim3(256), 0, s, vsi, desc, num_f,
                       num_es, num_vs, num_cfg, vs, m, lp, lp_stride, d, omd, t, out, gvt, gvs,
                       fold, glp);
  else
    hipLaunchKernelGGL((k_unary_bwd<S, false>), grid, dim3(256), 0, s, vsi, desc, num_f,
                       num_es, num_vs, num_cfg, vs, m, lp, lp_stride, d, omd, t, out, gvt, gvs,
                       fold, glp);
}

template <int S>
void launch_pair(bool fwd, bool t0, int32_t batch, hipStream_t s, const int32_t* vsi,
                 const int2* desc, int64_t num_f, int64_t num_es, int64_t num_vs,
                 int64_t num_cfg, const float* vs, const float* m, const float* lp,
                 int64_t lp_stride, float d, float omd, float t, float* out, float* gvt,
                 const float* gvs, int32_t fold, float* glp) {
  const dim3 grid(unsigned((num_f + 255) / 256), batch);
  if (fwd && t0)
    hipLaunchKernelGGL((k_pair_fwd<S, true>), grid, dim3(256), 0, s, vsi, desc, num_f, num_es,
                       num_vs, vs, m, lp, lp_stride, d, omd, t, out);
  else if (fwd)
    hipLaunchKernelGGL((k_pair_fwd<S, false>), grid, dim3(256), 0, s, vsi, desc, num_f, num_es,
                       num_vs, vs, m, lp, lp_stride, d, omd, t, out);
  else if (t0)
    hipLaunchKernelGGL((k_pair_bwd<S, true>), grid, dim3(256), 0, s, vsi, desc, num_f, num_es,
                       num_vs, num_cfg, vs, m, lp, lp_stride, d, omd, t, out, gvt, gvs, fold,
                       glp);
  else
    hipLaunchKernelGGL((k_pair_bwd<S, false>), grid, dim3(256), 0, s, vsi, desc, num_f, num_es,
                       num_vs, num_cfg, vs, m, lp, lp_stride, d, omd, t, out, gvt, gvs, fold,
                       glp);
}

// CSR of `key` (values in [0, n)) over positions 0 .. m-1, positions ascending
void csr(const std::vector<int32_t>& key, int64_t n, std::vector<int32_t>* ptr,
         std::vector<int32_t>* idx) {
  ptr->assign(size_t(n) + 1, 0);
  for (int32_t k : key) (*ptr)[size_t(k) + 1]++;
  for (int64_t i = 0; i < n; ++i) (*ptr)[size_t(i) + 1] += (*ptr)[size_t(i)];
  std::vector<int32_t> fill(ptr->begin(), ptr->end() - 1);
  idx->assign(key.size(), 0);
  for (size_t p = 0; p < key.size(); ++p) (*idx)[size_t(fill[size_t(key[p])]++)] = int32_t(p);
}

}  // namespace
}  // namespace pgx

using namespace pgx;

extern "C" {

int pgx_vjp_create(pgx_graph* graph, const int32_t* var_states_for_edge_states,
                   const int32_t* edge_indices_for_edge_states, const int32_t* row_cfg,
                   const int32_t* row_es, int64_t num_rows, pgx_vjp** out) {
  if (!graph || !out || num_rows < 0 || (num_rows > 0 && (!row_cfg || !row_es)))
    return set_error(PGX_EINVAL, "pgx_vjp_create: bad argument");
  const Plan& P = graph_plan(graph);
  const int64_t E_s = P.num_es, V_s = P.num_vs, C = P.num_lp;

  if (E_s > 0 && (!var_states_for_edge_states || !edge_indices_for_edge_states))
    return set_error(PGX_EINVAL, "pgx_vjp_create: null wiring array");
  std::vector<int32_t> vsi(var_states_for_edge_states, var_states_for_edge_states + E_s);
  std::vector<int32_t> edge_ptr{0};
  for (int64_t i = 0; i < E_s; ++i) {
    const int32_t e = edge_indices_for_edge_states[i];
    if (vsi[size_t(i)] < 0 || vsi[size_t(i)] >= V_s)
      return set_error(PGX_EINVAL, "pgx_vjp_create: variable state out of range");
    if (e == int32_t(edge_ptr.size()) - 1) continue;
    if (e != int32_t(edge_ptr.size()))
      return set_error(PGX_EINVAL, "pgx_vjp_create: edge states of an edge must be consecutive");
    edge_ptr.push_back(int32_t(i));
  }
  edge_ptr.push_back(int32_t(E_s));
  if (E_s == 0) edge_ptr = {0};
  std::vector<int32_t> rc(row_cfg, row_cfg + num_rows), re(row_es, row_es + num_rows);
  std::vector<int32_t> cfg_ptr(size_t(C) + 1, 0);
  for (int64_t r = 0; r < num_rows; ++r) {
    if (rc[size_t(r)] < 0 || rc[size_t(r)] >= C || re[size_t(r)] < 0 || re[size_t(r)] >= E_s)
      return set_error(PGX_EINVAL, "pgx_vjp_create: row out of range");
    if (r > 0 && rc[size_t(r)] < rc[size_t(r - 1)])
      return set_error(PGX_EINVAL, "pgx_vjp_create: rows must be ordered by config");
    cfg_ptr[size_t(rc[size_t(r)]) + 1]++;
  }
  for (int64_t c = 0; c < C; ++c) cfg_ptr[size_t(c) + 1] += cfg_ptr[size_t(c)];
  std::vector<int32_t> es_ptr, es_rows, vs_ptr, vs_es;
  csr(re, E_s, &es_ptr, &es_rows);
  std::vector<int32_t> es_cfg(es_rows.size());
  for (size_t p = 0; p < es_rows.size(); ++p) es_cfg[p] = rc[size_t(es_rows[p])];
  csr(vsi, V_s, &vs_ptr, &vs_es);
  // the pairwise layout: factor by factor from its first config's rows,
  // every row checked, every edge state owned by exactly one factor
  std::vector<std::vector<int2>> lists(9), ulists(9);
  bool pair = C > 0 && E_s < (int64_t(1) << 31) && C < (int64_t(1) << 31);
  {
    const int64_t num_edges = int64_t(edge_ptr.size()) - 1;
    auto rows_of = [&](int64_t c) { return cfg_ptr[size_t(c) + 1] - cfg_ptr[size_t(c)]; };
    std::vector<char> owned(size_t(E_s), 0);
    for (int64_t c = 0; pair && c < C;) {
      const int64_t nr = rows_of(c), r0 = cfg_ptr[size_t(c)];
      const int32_t ea = nr >= 1 ? re[size_t(r0)] : 0;
      const int64_t ga = nr >= 1 ? edge_indices_for_edge_states[ea] : 0;
      const int64_t S = nr >= 1 ? edge_ptr[size_t(ga) + 1] - edge_ptr[size_t(ga)] : 0;
      pair = (nr == 1 || nr == 2) && S >= 1 && S <= 8 && edge_ptr[size_t(ga)] == ea;
      int64_t span = 0;
      if (pair && nr == 1) {  // unary: config c + i has one row, state ea + i
        pair = c + S <= C;
        for (int64_t i = 0; pair && i < S; ++i)
          pair = rows_of(c + i) == 1 && re[size_t(cfg_ptr[size_t(c + i)])] == ea + i;
        span = S;
      } else if (pair) {  // square pairwise
        const int32_t eb = re[size_t(r0) + 1];
        pair = S >= 2 && ga + 1 < num_edges && eb == ea + S &&
               edge_ptr[size_t(ga) + 2] - edge_ptr[size_t(ga) + 1] == S && c + S * S <= C;
        for (int64_t i = 0; pair && i < S; ++i)
          for (int64_t j = 0; pair && j < S; ++j) {
            const int64_t x = c + S * i + j, rx = cfg_ptr[size_t(x)];
            pair = rows_of(x) == 2 && re[size_t(rx)] == ea + i && re[size_t(rx) + 1] == eb + j;
          }
        span = S * S;
      }
      const int64_t nes = nr == 1 ? S : 2 * S;
      for (int64_t k = 0; pair && k < nes; ++k) pair = !owned[size_t(ea + k)]++;
      if (pair) (nr == 1 ? ulists : lists)[size_t(S)].push_back(make_int2(ea, int32_t(c)));
      c += pair ? span : 0;
    }
    for (int64_t k = 0; pair && k < E_s; ++k) pair = owned[size_t(k)] == 1;
  }
  pgx_vjp* v = new pgx_vjp();
  v->graph = graph;
  v->pair = pair;
  for (int t = 0; t < 2; ++t)
    v->logical_needed[t] = P.msgs_range[t + 1][1] > P.msgs_range[t + 1][0];
  v->pool_needed = P.msgs_range[PGX_POOL][1] > P.msgs_range[PGX_POOL][0];
  v->num_enum_es = P.msgs_range[PGX_ENUM][1];
  v->num_es = E_s;
  v->num_vs = V_s;
  v->num_lp = C;
  v->num_edges = int64_t(edge_ptr.size()) - 1;
  v->num_rows = num_rows;
  auto up = [&](const std::vector<int32_t>& h, int32_t** d) -> int {
    const size_t bytes = std::max<size_t>(h.size(), 1) * sizeof(int32_t);
    if (hipMalloc(reinterpret_cast<void**>(d), bytes) != hipSuccess) return PGX_EHIP;
    if (!h.empty() && hipMemcpy(*d, h.data(), h.size() * sizeof(int32_t),
                                hipMemcpyHostToDevice) != hipSuccess)
      return PGX_EHIP;
    return PGX_OK;
  };
  bool failed = up(vsi, &v->vsi) || up(edge_ptr, &v->edge_ptr) || up(cfg_ptr, &v->cfg_ptr) ||
                up(re, &v->row_es) || up(es_ptr, &v->es_ptr) || up(es_cfg, &v->es_cfg) ||
                up(vs_ptr, &v->vs_ptr) || up(vs_es, &v->vs_es);
  {  // variable states of at most four edge states: one 16-byte descriptor each
    bool small = true;
    for (int64_t k = 0; k < V_s && small; ++k)
      small = vs_ptr[size_t(k) + 1] - vs_ptr[size_t(k)] <= 4;
    if (small && V_s > 0 && !failed) {
      std::vector<int4> q(static_cast<size_t>(V_s));
      for (int64_t k = 0; k < V_s; ++k) {
        int32_t w[4] = {-1, -1, -1, -1};
        for (int32_t p = vs_ptr[size_t(k)]; p < vs_ptr[size_t(k) + 1]; ++p)
          w[p - vs_ptr[size_t(k)]] = vs_es[size_t(p)];
        q[size_t(k)] = make_int4(w[0], w[1], w[2], w[3]);
      }
      failed = hipMalloc(reinterpret_cast<void**>(&v->vs_pos4), q.size() * sizeof(int4)) !=
                   hipSuccess ||
               hipMemcpy(v->vs_pos4, q.data(), q.size() * sizeof(int4), hipMemcpyHostToDevice) !=
                   hipSuccess;
    }
  }
  auto up_desc = [&](const std::vector<int2>& l, int2** d) {
    return hipMalloc(reinterpret_cast<void**>(d), l.size() * sizeof(int2)) != hipSuccess ||
           hipMemcpy(*d, l.data(), l.size() * sizeof(int2), hipMemcpyHostToDevice) != hipSuccess;
  };
  for (int S = 1; pair && S <= 8 && !failed; ++S) {
    v->pair_count[S] = int64_t(lists[size_t(S)].size());
    v->unary_count[S] = int64_t(ulists[size_t(S)].size());
    if (!lists[size_t(S)].empty()) failed = up_desc(lists[size_t(S)], &v->pair_desc[S]);
    if (!failed && !ulists[size_t(S)].empty())
      failed = up_desc(ulists[size_t(S)], &v->unary_desc[S]);
  }
  if (failed) {
    pgx_vjp_destroy(v);
    return set_error(PGX_EHIP, "pgx_vjp_create: device allocation failed");
  }
  *out = v;
  return PGX_OK;
}

int pgx_vjp_destroy(pgx_vjp* v) {
  if (!v) return PGX_OK;
  for (int32_t* p : {v->vsi, v->edge_ptr, v->cfg_ptr, v->row_es, v->es_ptr, v->es_cfg,
                     v->vs_ptr, v->vs_es})
    if (p) (void)hipFree(p);
  for (int2* p : v->pair_desc)
    if (p) (void)hipFree(p);
  if (v->vs_pos4) (void)hipFree(v->vs_pos4);
  for (int2* p : v->unary_desc)
    if (p) (void)hipFree(p);
  for (int t = 0; t < 2; ++t)
    for (int32_t* p : {v->logical_ptr[t], v->logical_pes[t], v->logical_ces[t]})
      if (p) (void)hipFree(p);
  for (int32_t* p : {v->pool_ptr, v->pool_pes, v->pool_ies})
    if (p) (void)hipFree(p);
  if (v->scratch) (void)hipFree(v->scratch);
  delete v;
  return PGX_OK;
}

int pgx_vjp_set_logical(pgx_vjp* v, int32_t factor_type, const int32_t* parents_factor_indices,
                        const int32_t* parents_msg_indices, int64_t num_parents,
                        const int32_t* children_edge_states, int64_t num_factors,
                        int32_t edge_states_offset) {
  if (!v || (factor_type != PGX_OR && factor_type != PGX_AND) || num_parents < 0 ||
      num_factors < 0 || (num_parents > 0 && (!parents_factor_indices || !parents_msg_indices)) ||
      (num_factors > 0 && !children_edge_states))
    return set_error(PGX_EINVAL, "pgx_vjp_set_logical: bad argument");
  const int t = factor_type - 1;
  const Plan& P = graph_plan(v->graph);
  const int64_t m0 = P.msgs_range[factor_type][0], m1 = P.msgs_range[factor_type][1];
  if (num_factors > 0 && edge_states_offset != (factor_type == PGX_OR ? 1 : -1))
    return set_error(PGX_EINVAL, "pgx_vjp_set_logical: OR takes edge_states_offset 1, AND -1");
  const int32_t off = edge_states_offset;
  auto in_range = [&](int64_t e) { return e >= m0 && e < m1 && e + off >= m0 && e + off < m1; };
  std::vector<int32_t> ptr(size_t(num_factors) + 1, 0);
  std::vector<int32_t> pes(static_cast<size_t>(num_parents), 0);
  std::vector<int32_t> ces(static_cast<size_t>(num_factors), 0);
  for (int64_t j = 0; j < num_parents; ++j) {
    const int32_t f = parents_factor_indices[j];
    if (f < 0 || f >= num_factors || !in_range(m0 + parents_msg_indices[j]))
      return set_error(PGX_EINVAL, "pgx_vjp_set_logical: parent out of range");
    ptr[size_t(f) + 1]++;
  }
  for (int64_t f = 0; f < num_factors; ++f) {
    if (!in_range(m0 + children_edge_states[f]))
      return set_error(PGX_EINVAL, "pgx_vjp_set_logical: child out of range");
    ces[size_t(f)] = int32_t(m0 + children_edge_states[f]);
    ptr[size_t(f) + 1] += ptr[size_t(f)];
  }
  {
    std::vector<int32_t> fill(ptr.begin(), ptr.end() - 1);
    for (int64_t j = 0; j < num_parents; ++j)
      pes[size_t(fill[size_t(parents_factor_indices[j])]++)] =
          int32_t(m0 + parents_msg_indices[j]);
  }
  for (int32_t* q : {v->logical_ptr[t], v->logical_pes[t], v->logical_ces[t]})
    if (q) (void)hipFree(q);
  v->logical_ptr[t] = v->logical_pes[t] = v->logical_ces[t] = nullptr;
  v->logical_factors[t] = 0;
  auto up = [&](const std::vector<int32_t>& h, int32_t** d) {
    const size_t bytes = std::max<size_t>(h.size(), 1) * sizeof(int32_t);
    return hipMalloc(reinterpret_cast<void**>(d), bytes) != hipSuccess ||
           (!h.empty() && hipMemcpy(*d, h.data(), h.size() * sizeof(int32_t),
                                    hipMemcpyHostToDevice) != hipSuccess);
  };
  if (up(ptr, &v->logical_ptr[t]) || up(pes, &v->logical_pes[t]) || up(ces, &v->logical_ces[t]))
    return set_error(PGX_EHIP, "pgx_vjp_set_logical: device allocation failed");
  v->logical_factors[t] = num_factors;
  v->logical_off[t] = off;
  return PGX_OK;
}

int pgx_vjp_set_pool(pgx_vjp* v, const int32_t* choices_factor_indices,
                     const int32_t* choices_msg_indices, int64_t num_choices,
                     const int32_t* indicators_edge_states, int64_t num_factors) {
  if (!v || num_choices < 0 || num_factors < 0 ||
      (num_choices > 0 && (!choices_factor_indices || !choices_msg_indices)) ||
      (num_factors > 0 && !indicators_edge_states))
    return set_error(PGX_EINVAL, "pgx_vjp_set_pool: bad argument");
  const Plan& P = graph_plan(v->graph);
  const int64_t m0 = P.msgs_range[PGX_POOL][0], m1 = P.msgs_range[PGX_POOL][1];
  auto in_range = [&](int64_t e) { return e >= m0 && e + 1 < m1; };
  std::vector<int32_t> ptr(size_t(num_factors) + 1, 0);
  std::vector<int32_t> pes(static_cast<size_t>(num_choices), 0);
  std::vector<int32_t> ies(static_cast<size_t>(num_factors), 0);
  for (int64_t j = 0; j < num_choices; ++j) {
    const int32_t f = choices_factor_indices[j];
    if (f < 0 || f >= num_factors || !in_range(m0 + choices_msg_indices[j]))
      return set_error(PGX_EINVAL, "pgx_vjp_set_pool: choice out of range");
    ptr[size_t(f) + 1]++;
  }
  for (int64_t f = 0; f < num_factors; ++f) {
    if (!in_range(m0 + indicators_edge_states[f]))
      return set_error(PGX_EINVAL, "pgx_vjp_set_pool: indicator out of range");
    ies[size_t(f)] = int32_t(m0 + indicators_edge_states[f]);
    ptr[size_t(f) + 1] += ptr[size_t(f)];
  }
  {
    std::vector<int32_t> fill(ptr.begin(), ptr.end() - 1);
    for (int64_t j = 0; j < num_choices; ++j)
      pes[size_t(fill[size_t(choices_factor_indices[j])]++)] =
          int32_t(m0 + choices_msg_indices[j]);
  }
  for (int32_t* q : {v->pool_ptr, v->pool_pes, v->pool_ies})
    if (q) (void)hipFree(q);
  v->pool_ptr = v->pool_pes = v->pool_ies = nullptr;
  v->pool_factors = 0;
  auto up = [&](const std::vector<int32_t>& h, int32_t** d) {
    const size_t bytes = std::max<size_t>(h.size(), 1) * sizeof(int32_t);
    return hipMalloc(reinterpret_cast<void**>(d), bytes) != hipSuccess ||
           (!h.empty() && hipMemcpy(*d, h.data(), h.size() * sizeof(int32_t),
                                    hipMemcpyHostToDevice) != hipSuccess);
  };
  if (up(ptr, &v->pool_ptr) || up(pes, &v->pool_pes) || up(ies, &v->pool_ies))
    return set_error(PGX_EHIP, "pgx_vjp_set_pool: device allocation failed");
  v->pool_factors = num_factors;
  return PGX_OK;
}

int pgx_run_vjp(pgx_vjp* v, const float* lp, int64_t lp_stride, const float* ev,
                int64_t ev_stride, const float* ftov_in, int64_t ftov_in_stride,
                const float* g_out, float* g_lp, float* g_ev, float* g_ftov_in, int32_t batch,
                int32_t num_iters, float damping, float temperature, void* stream) {
  if (!v || batch <= 0 || (!lp && v->num_lp > 0) || !ev || !ftov_in || !g_out ||
      (!g_lp && v->num_lp > 0) || !g_ev || !g_ftov_in)
    return set_error(PGX_EINVAL, "pgx_run_vjp: bad argument");
  if (!(temperature >= 0.f)) return set_error(PGX_EINVAL, "pgx_run_vjp: temperature < 0");
  if (batch > 65535)  // the batch element is the grid's y index
    return set_error(PGX_EUNSUP, "pgx_run_vjp: batch above 65535; split it into calls");
  for (int t = 0; t < 2; ++t)
    if (v->logical_needed[t] && !v->logical_ptr[t])
      return set_error(PGX_EINVAL, "pgx_run_vjp: the graph has OR / AND factors; register "
                                   "them with pgx_vjp_set_logical first");
  if (v->pool_needed && !v->pool_ptr)
    return set_error(PGX_EINVAL, "pgx_run_vjp: the graph has Pool factors; register them "
                                 "with pgx_vjp_set_pool first");
  const int64_t E_s = v->num_es, V_s = v->num_vs, C = v->num_lp;
  if (E_s == 0) return PGX_OK;
  const int32_t n = std::max(num_iters, 1);
  hipStream_t s = static_cast<hipStream_t>(stream);
  // scratch: tape (n + 1) B E_s, then per batch element vs, gvs, gev (V_s),
  // vt, f, gm, gf, gvt, mx, z (E_s), s, gs, glp (C), deltas (1); the pairwise
  // path adds the variable sums of every iteration (n B V_s)
  const size_t B = size_t(batch);
  const size_t pair_tape = v->pair ? size_t(n) * B * size_t(V_s) : 0;
  const size_t floats = (size_t(n) + 1) * B * size_t(E_s) + pair_tape +
                        B * (3 * size_t(V_s) + 7 * size_t(E_s) + 3 * size_t(C) + 1);
  if (floats * 4 > (size_t(64) << 30))
    return set_error(PGX_EUNSUP, "pgx_run_vjp: the run's tape would pass 64 GiB; use fewer "
                                 "iterations or a smaller batch");
  if (v->scratch_bytes < floats * 4) {
    if (v->scratch) PGX_HIP(hipFree(v->scratch));
    v->scratch = nullptr;
    v->scratch_bytes = 0;
    PGX_HIP(hipMalloc(reinterpret_cast<void**>(&v->scratch), floats * 4));
    v->scratch_bytes = floats * 4;
  }
  float* tape = v->scratch;
  float* tape_vs = tape + (size_t(n) + 1) * B * size_t(E_s);
  float* p = tape_vs + pair_tape;
  auto take = [&](size_t k) {
    float* q = p;
    p += k;
    return q;
  };
  float *vs = take(B * V_s), *gvs = take(B * V_s), *gev = take(B * V_s);
  float *vt = take(B * E_s), *f = take(B * E_s), *gm = take(B * E_s), *gf = take(B * E_s);
  float *gvt = take(B * E_s), *mx = take(B * E_s), *zz = take(B * E_s);
  float *sm = take(B * C), *gs = take(B * C), *glp = take(B * C), *dl = take(B);
  const float d = damping, omd = float(1.0 - double(damping));
  const bool t0 = temperature == 0.f;
  // forward: m_0 = normalise(ftov_in), then one iteration at a time
  hipLaunchKernelGGL(k_norm_fwd, grid_for(v->num_edges, batch), dim3(256), 0, s, v->edge_ptr,
                     v->num_edges, E_s, ftov_in, ftov_in_stride, tape);
  if (v->pair) {
    KernelClock clk;
    {
      std::lock_guard<std::mutex> lock(g_opt_mu);
      clk.on = g_opt_profile == 2;
    }
    clk.s = s;
    // one launch per factor shape present (pairwise and unary, per S)
    auto pair = [&](bool fwd, const float* vsx, const float* m, float* out, int32_t fold) {
      for (int S = 1; S <= 8; ++S) {
        const int64_t F = v->pair_count[S], U = v->unary_count[S];
        switch (S) {
#define PGX_PAIR_CASE(K)                                                                        \
  case K:                                                                                       \
    if (F > 0)                                                                                  \
      launch_pair<K>(fwd, t0, batch, s, v->vsi, v->pair_desc[K], F, E_s, V_s, C, vsx, m, lp,    \
                     lp_stride, d, omd, temperature, out, gvt, gvs, fold, glp);                 \
    if (U > 0)                                                                                  \
      launch_unary<K>(fwd, t0, batch, s, v->vsi, v->unary_desc[K], U, E_s, V_s, C, vsx, m, lp,  \
                      lp_stride, d, omd, temperature, out, gvt, gvs, fold, glp);                \
    break;
          PGX_PAIR_CASE(2)
          PGX_PAIR_CASE(3)
          PGX_PAIR_CASE(4)
          PGX_PAIR_CASE(5)
          PGX_PAIR_CASE(6)
          PGX_PAIR_CASE(7)
          PGX_PAIR_CASE(8)
#undef PGX_PAIR_CASE
          case 1:
            if (U > 0)
              launch_unary<1>(fwd, t0, batch, s, v->vsi, v->unary_desc[1], U, E_s, V_s, C, vsx,
                              m, lp, lp_stride, d, omd, temperature, out, gvt, gvs, fold, glp);
            break;
          default:
            break;
        }
      }
    };
    for (int32_t it = 0; it < n; ++it) {
      const float* m = tape + size_t(it) * B * E_s;
      float* vsx = tape_vs + size_t(it) * B * V_s;
      clk.run(PGX_KERNEL_VJP_VAR_SUMS, [&] {
        hipLaunchKernelGGL(k_var_sums, grid_for(V_s, batch), dim3(256), 0, s, v->vs_pos4,
                           v->vs_ptr, v->vs_es, V_s, E_s, ev, ev_stride, m, vsx);
      });
      clk.run(PGX_KERNEL_VJP_PAIR_FWD,
              [&] { pair(true, vsx, m, tape + size_t(it + 1) * B * E_s, 0); });
    }
    hipLaunchKernelGGL(k_zero, grid_for(B * C, 1), dim3(256), 0, s, glp, int64_t(B * C));
    hipLaunchKernelGGL(k_zero, grid_for(B * V_s, 1), dim3(256), 0, s, gev, int64_t(B * V_s));
    PGX_HIP(hipMemcpyAsync(gm, g_out, B * E_s * 4, hipMemcpyDeviceToDevice, s));
    // each step leaves its variable side (gm += gvs - gvt) to the next one
    for (int32_t it = n - 1; it >= 0; --it) {
      clk.run(PGX_KERNEL_VJP_PAIR_BWD, [&] {
        pair(false, tape_vs + size_t(it) * B * V_s, tape + size_t(it) * B * E_s, gm, it < n - 1);
      });
      clk.run(PGX_KERNEL_VJP_VAR_BWD, [&] {
        hipLaunchKernelGGL(k_gvs, grid_for(V_s, batch), dim3(256), 0, s, v->vs_pos4, v->vs_ptr,
                           v->vs_es,
                           V_s, E_s, gvt, gvs, gev);
      });
    }
    hipLaunchKernelGGL(k_gmsg, grid_for(E_s, batch), dim3(256), 0, s, v->vsi, E_s, V_s, gvt, gvs,
                       gm);
  }
  for (int32_t it = 0; !v->pair && it < n; ++it) {
    const int rc = pgx_run(v->graph, lp, lp_stride, ev, ev_stride, tape + size_t(it) * B * E_s,
                           E_s, tape + size_t(it + 1) * B * E_s, dl, batch, 1, damping,
                           temperature, stream);
    if (rc) return rc;
  }
  if (!v->pair) {
    hipLaunchKernelGGL(k_zero, grid_for(B * C, 1), dim3(256), 0, s, glp, int64_t(B * C));
    hipLaunchKernelGGL(k_zero, grid_for(B * V_s, 1), dim3(256), 0, s, gev, int64_t(B * V_s));
    PGX_HIP(hipMemcpyAsync(gm, g_out, B * E_s * 4, hipMemcpyDeviceToDevice, s));
  }
  // reverse
  for (int32_t it = n - 1; !v->pair && it >= 0; --it) {
    const float* m = tape + size_t(it) * B * E_s;
    int rc = pgx_get_beliefs(v->graph, ev, ev_stride, m, E_s, vs, batch, stream);
    if (rc) return rc;
    hipLaunchKernelGGL(k_vtof, grid_for(E_s, batch), dim3(256), 0, s, v->vsi, E_s, V_s, vs, m, vt);
    for (int32_t type : {PGX_ENUM, PGX_OR, PGX_AND, PGX_POOL}) {
      rc = fac_to_var_batched(v->graph, type, vt, lp, lp_stride, f, temperature, 1, batch, s);
      if (rc) return rc;
    }
    hipLaunchKernelGGL(k_norm_bwd<true>, grid_for(v->num_edges, batch), dim3(256), 0, s,
                       v->edge_ptr, v->num_edges, E_s, m, E_s, f, d, omd, gm, gf);
    if (C > 0) {
      hipLaunchKernelGGL(k_summaries, grid_for(C, batch), dim3(256), 0, s, v->cfg_ptr, v->row_es,
                         C, E_s, vt, lp, lp_stride, sm);
      if (t0) {
        hipLaunchKernelGGL(k_es_stats<true>, grid_for(E_s, batch), dim3(256), 0, s, v->es_ptr,
                           v->es_cfg, E_s, C, sm, temperature, mx, zz);
        hipLaunchKernelGGL(k_gsum<true>, grid_for(C, batch), dim3(256), 0, s, v->cfg_ptr,
                           v->row_es, C, E_s, sm, mx, zz, gf, lp, lp_stride, temperature, gs, glp);
      } else {
        hipLaunchKernelGGL(k_es_stats<false>, grid_for(E_s, batch), dim3(256), 0, s, v->es_ptr,
                           v->es_cfg, E_s, C, sm, temperature, mx, zz);
        hipLaunchKernelGGL(k_gsum<false>, grid_for(C, batch), dim3(256), 0, s, v->cfg_ptr,
                           v->row_es, C, E_s, sm, mx, zz, gf, lp, lp_stride, temperature, gs, glp);
      }
    }
    if (v->num_enum_es > 0)
      hipLaunchKernelGGL(k_gvtof, grid_for(v->num_enum_es, batch), dim3(256), 0, s, v->es_ptr,
                         v->es_cfg, E_s, v->num_enum_es, C, gs, gf, gvt);
    for (int lt = 0; lt < 2; ++lt) {
      const int64_t F = v->logical_factors[lt];
      if (F == 0) continue;
      const dim3 lg(unsigned((F + 255) / 256), batch);
      if (t0)
        hipLaunchKernelGGL(k_logical_bwd<true>, lg, dim3(256), 0, s, v->logical_ptr[lt],
                           v->logical_pes[lt], v->logical_ces[lt], F, v->logical_off[lt], E_s, vt,
                           gf, temperature, gvt);
      else
        hipLaunchKernelGGL(k_logical_bwd<false>, lg, dim3(256), 0, s, v->logical_ptr[lt],
                           v->logical_pes[lt], v->logical_ces[lt], F, v->logical_off[lt], E_s, vt,
                           gf, temperature, gvt);
    }
    if (v->pool_factors > 0) {
      const dim3 pg(unsigned((v->pool_factors + 255) / 256), batch);
      if (t0)
        hipLaunchKernelGGL(k_pool_bwd<true>, pg, dim3(256), 0, s, v->pool_ptr, v->pool_pes,
                           v->pool_ies, v->pool_factors, E_s, vt, gf, temperature, gvt);
      else
        hipLaunchKernelGGL(k_pool_bwd<false>, pg, dim3(256), 0, s, v->pool_ptr, v->pool_pes,
                           v->pool_ies, v->pool_factors, E_s, vt, gf, temperature, gvt);
    }
    hipLaunchKernelGGL(k_gvs, grid_for(V_s, batch), dim3(256), 0, s, v->vs_pos4, v->vs_ptr,
                       v->vs_es, V_s,
                       E_s, gvt, gvs, gev);
    hipLaunchKernelGGL(k_gmsg, grid_for(E_s, batch), dim3(256), 0, s, v->vsi, E_s, V_s, gvt, gvs,
                       gm);
  }
  hipLaunchKernelGGL(k_norm_bwd<false>, grid_for(v->num_edges, batch), dim3(256), 0, s,
                     v->edge_ptr, v->num_edges, E_s, ftov_in, ftov_in_stride, nullptr, 1.f, 0.f,
                     gm, nullptr);
  // per batch element, or summed over the batch for a shared input
  auto emit = [&](const float* src, int64_t len, int64_t stride, float* dst) -> int {
    if (len == 0) return PGX_OK;
    if (stride)
      PGX_HIP(hipMemcpyAsync(dst, src, B * size_t(len) * 4, hipMemcpyDeviceToDevice, s));
    else
      hipLaunchKernelGGL(k_batch_sum, grid_for(len, 1), dim3(256), 0, s, src, len, batch, dst);
    return PGX_OK;
  };
  int rc = emit(glp, C, lp_stride, g_lp);
  if (!rc) rc = emit(gev, V_s, ev_stride, g_ev);
  if (!rc) rc = emit(gm, E_s, ftov_in_stride, g_ftov_in);
  if (rc) return rc;
  PGX_HIP(hipGetLastError());
  return PGX_OK;
}

}  // extern "C"
