/*
 * C/OpenMP restatement of the reference's BP iteration for Enum factors
 * (and, at the end of this file, for OR / AND factors at temperature 0) --
 * TEST INFRASTRUCTURE ONLY (the timed CPU baseline of bench.py and a
 * cross-check in tests/). The product path never links it.
 *
 * Follows the reference's scatter formulation operation by operation:
 *   pgmax/infer/bp.py:85-96     clip log-potentials, normalise + clip msgs
 *   pgmax/infer/bp.py:191-219   variable sums (in the oracle's chunked order for
 *                               states with more than 32 messages), then
 *                               vtof = vsum[var state] - msg
 *   pgmax/factor/enum.py:451-474 summary[c] = lp[c] + sum of its rows' vtof;
 *                                out[es] = max (T = 0) or T-logsumexp
 *                                (update_utils.py:67-98) over the rows of es;
 *                                ftov = out - vtof
 *   pgmax/infer/bp.py:127-136   damping, normalise + clip, max |change|
 * Batch elements run in parallel (OpenMP); within one element the loops are
 * sequential, so results do not depend on the thread count.
 *
 * exp and log are the oracle's own FP32 constructions (oracle/f32math.py:
 * Cody-Waite reduction and a degree-6 fitted polynomial in fused
 * multiply-adds for exp, the atanh series for log), with every fused
 * multiply-add written as fmaf and every other operation rounding once
 * (built with -ffp-contract=off). The results are therefore bit for bit
 * those of the NumPy oracle, which tests/test_oracle.py checks, and this
 * file serves as a fast bitwise oracle for full-size runs.
 */
#include <float.h>
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define MSG_NEG_INF (-1e32f)
#define LP_MAX_ABS (1e6f)
#define CHUNK 32 /* VAR_SUM_CHUNK of oracle/lbp.py */

static float nmaxf(float a, float b) { return (a > b || a != a) ? a : b; }

/* oracle/f32math.py exp: in-range arguments use x = k ln2 + r,
 * k = rint(x / ln2), ln2 split in a 9-bit head and a tail. */
static float ox_exp(float x) {
  if (x != x) return x;
  if (x > 88.7228317f) return INFINITY;
  if (x < -103.972084f) return 0.f;
  const float k = rintf(x * 1.44269504088896341f);
  float r = fmaf(-k, 0.693359375f, x);
  r = fmaf(-k, -2.12194440054690583e-4f, r);
  float p = 0x1.6c07b2p-10f;
  p = fmaf(p, r, 0x1.12609cp-7f);
  p = fmaf(p, r, 0x1.555748p-5f);
  p = fmaf(p, r, 0x1.555452p-3f);
  p = fmaf(p, r, 0x1.fffffcp-2f);
  p = fmaf(p, r, 1.f);
  p = fmaf(p, r, 1.f);
  return ldexpf(p, (int)k);
}

/* oracle/f32math.py log: x = m 2^e, m in [sqrt(1/2), sqrt(2)),
 * log m = f - s (f - R) with f = m - 1, s = f / (2 + f). */
static float ox_log(float x) {
  if (x != x) return x;
  if (x < 0.f) return NAN;
  if (x == 0.f) return -INFINITY;
  if (x == INFINITY) return x;
  int e;
  float m = frexpf(x, &e);
  if (m < 0.707106781186547524f) {
    m = m * 2.f;
    e = e - 1;
  }
  const float f = m - 1.f;
  const float s = f / (2.f + f);
  const float z = s * s;
  float q = 2.f / 11.f;
  q = q * z + 2.f / 9.f;
  q = q * z + 2.f / 7.f;
  q = q * z + 2.f / 5.f;
  q = q * z + 2.f / 3.f;
  const float logm = f - s * (f - q * z);
  const float ef = (float)e;
  return (ef * -2.12194440054690583e-4f + logm) + ef * 0.693359375f;
}

static void normalize(const int32_t* edge_of_es, int64_t num_es, int64_t num_edges,
                      float* msgs, float* emax) {
  for (int64_t e = 0; e < num_edges; ++e) emax[e] = -INFINITY;
  for (int64_t i = 0; i < num_es; ++i) emax[edge_of_es[i]] = nmaxf(emax[edge_of_es[i]], msgs[i]);
  for (int64_t i = 0; i < num_es; ++i) {
    const float v = msgs[i] - emax[edge_of_es[i]];
    msgs[i] = nmaxf(v, MSG_NEG_INF);
  }
}

/* Elementwise exp (op 0) or log (op 1) of n floats: lets tests compare these
 * two functions with oracle/f32math.py directly. */
void lbp_cpu_math(int op, const float* x, float* y, int64_t n) {
  for (int64_t i = 0; i < n; ++i) y[i] = op == 0 ? ox_exp(x[i]) : ox_log(x[i]);
}

/* Threads used by the next runs (0 keeps the OpenMP default); returns the
 * number a parallel region will use. */
int lbp_cpu_threads(int n) {
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
}

/* One BP run of one graph (one batch element): iters iterations from
 * ftov_in into ftov_out, deltas[iters]. Returns 0, or -1 if memory runs out. */
static int run_one(int64_t num_es, int64_t num_vs, int64_t num_edges, int64_t num_configs,
                   int64_t num_rows, const int32_t* var_state_of_es, const int32_t* edge_of_es,
                   const int32_t* cfg_of_row, const int32_t* es_of_row, const float* L,
                   const float* V, const float* ftov_in, float* ftov_out, float* deltas,
                   int32_t iters, float damping, float temperature) {
  int failed = 0;
  float* msgs = (float*)malloc(sizeof(float) * (size_t)(num_es > 0 ? num_es : 1));
  float* nw = (float*)malloc(sizeof(float) * (size_t)(num_es > 0 ? num_es : 1));
  float* vtof = (float*)malloc(sizeof(float) * (size_t)(num_es > 0 ? num_es : 1));
  float* out = (float*)malloc(sizeof(float) * (size_t)(num_es > 0 ? num_es : 1));
  float* vsum = (float*)malloc(sizeof(float) * (size_t)(num_vs > 0 ? num_vs : 1));
  float* emax = (float*)malloc(sizeof(float) * (size_t)(num_edges > 0 ? num_edges : 1));
  float* summary = (float*)malloc(sizeof(float) * (size_t)(num_configs > 0 ? num_configs : 1));
  float* plp = (float*)malloc(sizeof(float) * (size_t)(num_configs > 0 ? num_configs : 1));
  /* chunked variable-state sums (oracle/lbp.py chunked_segment_sum): each
   * state's messages, in edge-state order, are summed in runs of CHUNK from
   * 0, runs in groups of CHUNK from 0, and the group sums in order from 0 */
  int32_t* rank = (int32_t*)malloc(sizeof(int32_t) * (size_t)(num_es > 0 ? num_es : 1));
  int64_t* cnt = (int64_t*)calloc((size_t)(num_vs > 0 ? num_vs : 1), sizeof(int64_t));
  int64_t* c0 = (int64_t*)malloc(sizeof(int64_t) * (size_t)(num_vs + 1));
  int64_t* g0 = (int64_t*)malloc(sizeof(int64_t) * (size_t)(num_vs + 1));
  float* runs = NULL;
  float* groups = NULL;
  int chunked = 0;
  if (rank && cnt && c0 && g0) {
    for (int64_t i = 0; i < num_es; ++i) rank[i] = (int32_t)cnt[var_state_of_es[i]]++;
    c0[0] = g0[0] = 0;
    for (int64_t v = 0; v < num_vs; ++v) {
      const int64_t n1 = (cnt[v] + CHUNK - 1) / CHUNK;
      c0[v + 1] = c0[v] + n1;
      g0[v + 1] = g0[v] + (n1 + CHUNK - 1) / CHUNK;
      if (cnt[v] > CHUNK) chunked = 1;
    }
    runs = (float*)malloc(sizeof(float) * (size_t)(c0[num_vs] > 0 ? c0[num_vs] : 1));
    groups = (float*)malloc(sizeof(float) * (size_t)(g0[num_vs] > 0 ? g0[num_vs] : 1));
  }
  if (!msgs || !nw || !vtof || !out || !vsum || !emax || !summary || !plp || !rank || !cnt ||
      !c0 || !g0 || !runs || !groups) {
    failed = 1;
  } else {
    for (int64_t c = 0; c < num_configs; ++c) {
      const float x = L[c];
      plp[c] = x < -LP_MAX_ABS ? -LP_MAX_ABS : (x > LP_MAX_ABS ? LP_MAX_ABS : x);
    }
    memcpy(msgs, ftov_in, sizeof(float) * (size_t)num_es);
    normalize(edge_of_es, num_es, num_edges, msgs, emax);
    for (int32_t it = 0; it < iters; ++it) {
      /* variable-to-factor */
      if (!chunked) {
        for (int64_t s = 0; s < num_vs; ++s) vsum[s] = 0.f;
        for (int64_t i = 0; i < num_es; ++i) vsum[var_state_of_es[i]] += msgs[i];
      } else {
        for (int64_t r = 0; r < c0[num_vs]; ++r) runs[r] = 0.f;
        for (int64_t i = 0; i < num_es; ++i)
          runs[c0[var_state_of_es[i]] + rank[i] / CHUNK] += msgs[i];
        for (int64_t q = 0; q < g0[num_vs]; ++q) groups[q] = 0.f;
        for (int64_t s = 0; s < num_vs; ++s)
          for (int64_t r = 0; r < c0[s + 1] - c0[s]; ++r) groups[g0[s] + r / CHUNK] += runs[c0[s] + r];
        for (int64_t s = 0; s < num_vs; ++s) {
          float t = 0.f;
          for (int64_t q = g0[s]; q < g0[s + 1]; ++q) t += groups[q];
          vsum[s] = t;
        }
      }
      for (int64_t s = 0; s < num_vs; ++s) vsum[s] += V[s];
      for (int64_t i = 0; i < num_es; ++i) vtof[i] = vsum[var_state_of_es[i]] - msgs[i];
      /* Enum factor-to-variable */
      for (int64_t c = 0; c < num_configs; ++c) summary[c] = 0.f;
      for (int64_t r = 0; r < num_rows; ++r) summary[cfg_of_row[r]] += vtof[es_of_row[r]];
      for (int64_t c = 0; c < num_configs; ++c) summary[c] += plp[c];
      for (int64_t i = 0; i < num_es; ++i) out[i] = -INFINITY;
      for (int64_t r = 0; r < num_rows; ++r)
        out[es_of_row[r]] = nmaxf(out[es_of_row[r]], summary[cfg_of_row[r]]);
      if (temperature != 0.f) {
        for (int64_t i = 0; i < num_es; ++i) nw[i] = 0.f;
        for (int64_t r = 0; r < num_rows; ++r) {
          const int32_t es = es_of_row[r];
          nw[es] += ox_exp((summary[cfg_of_row[r]] - out[es]) / temperature);
        }
        for (int64_t i = 0; i < num_es; ++i) out[i] = temperature * ox_log(nw[i]) + out[i];
      }
      /* damping, normalisation, delta */
      for (int64_t i = 0; i < num_es; ++i)
        nw[i] = damping * msgs[i] + (1.f - damping) * (out[i] - vtof[i]);
      normalize(edge_of_es, num_es, num_edges, nw, emax);
      float d = 0.f;
      for (int64_t i = 0; i < num_es; ++i) d = nmaxf(d, fabsf(nw[i] - msgs[i]));
      deltas[it] = d;
      float* t = msgs;
      msgs = nw;
      nw = t;
    }
    memcpy(ftov_out, msgs, sizeof(float) * (size_t)num_es);
  }
  free(msgs);
  free(nw);
  free(vtof);
  free(out);
  free(vsum);
  free(emax);
  free(summary);
  free(plp);
  free(rank);
  free(cnt);
  free(c0);
  free(g0);
  free(runs);
  free(groups);
  return failed ? -1 : 0;
}

/* Runs `iters` iterations for `batch` elements. lp: batch x C (stride
 * lp_stride), ev: batch x V_s (ev_stride), ftov_in / ftov_out: batch x E_s,
 * deltas: batch x iters. Rows: cfg_of_row[r] (configs numbered in row order,
 * rows of one config consecutive) and es_of_row[r] (absolute edge state).
 * Returns 0, or -1 if memory runs out. */
int lbp_cpu_run_enum(int64_t num_es, int64_t num_vs, int64_t num_edges, int64_t num_configs,
                     int64_t num_rows, const int32_t* var_state_of_es,
                     const int32_t* edge_of_es, const int32_t* cfg_of_row,
                     const int32_t* es_of_row, int64_t lp_offset,
                     const float* lp, int64_t lp_stride, const float* ev, int64_t ev_stride,
                     const float* ftov_in, float* ftov_out, float* deltas, int32_t batch,
                     int32_t iters, float damping, float temperature) {
  int failed = 0;
  if (iters < 1) iters = 1;
#pragma omp parallel for schedule(dynamic, 1) reduction(| : failed)
  for (int32_t b = 0; b < batch; ++b)
    failed |= run_one(num_es, num_vs, num_edges, num_configs, num_rows, var_state_of_es,
                      edge_of_es, cfg_of_row, es_of_row, lp + (int64_t)b * lp_stride + lp_offset,
                      ev + (int64_t)b * ev_stride, ftov_in + (int64_t)b * num_es,
                      ftov_out + (int64_t)b * num_es, deltas + (int64_t)b * iters, iters,
                      damping, temperature) != 0;
  return failed ? -1 : 0;
}

/* Independent graphs of Enum factors (an uneven batch), in parallel over the
 * graphs. Graph g's arrays start at the offsets in sizes[6 g .. 6 g + 5] =
 * [edge states, variable states, edges, configs, rows, iterations] -- each
 * array of a kind is the graphs' arrays concatenated, in graph order, with
 * graph-local indices; sizes[6 g + 5] is unused. */
int lbp_cpu_run_enum_graphs(int32_t num_graphs, const int64_t* sizes,
                            const int32_t* var_state_of_es, const int32_t* edge_of_es,
                            const int32_t* cfg_of_row, const int32_t* es_of_row,
                            const float* lp, const float* ev, const float* ftov_in,
                            float* ftov_out, float* deltas, int32_t iters, float damping,
                            float temperature) {
  int failed = 0;
  if (iters < 1) iters = 1;
  int64_t* off = (int64_t*)calloc((size_t)(num_graphs + 1) * 5, sizeof(int64_t));
  if (!off) return -1;
  for (int32_t g = 0; g < num_graphs; ++g)
    for (int k = 0; k < 5; ++k) off[5 * (g + 1) + k] = off[5 * g + k] + sizes[6 * g + k];
#pragma omp parallel for schedule(dynamic, 1) reduction(| : failed)
  for (int32_t g = 0; g < num_graphs; ++g) {
    const int64_t* o = off + 5 * g;
    const int64_t* n = sizes + 6 * g;
    failed |= run_one(n[0], n[1], n[2], n[3], n[4], var_state_of_es + o[0], edge_of_es + o[0],
                      cfg_of_row + o[4], es_of_row + o[4], lp + o[3], ev + o[1], ftov_in + o[0],
                      ftov_out + o[0], deltas + (int64_t)g * iters, iters, damping,
                      temperature) != 0;
  }
  free(off);
  return failed ? -1 : 0;
}

/* ---- OR / AND factors, max-product ----------------------------------------
 * pgmax/factor/logical.py:561-779 at temperature 0, as oracle/lbp.py's
 * pass_logical_fac_to_var_messages states it (temperature-0 branch):
 *   s_oth = sum of the parents' other-state vtof, s_mx = sum of
 *   max(other, relevant), m1 / m2 the largest and second-largest diff
 *   (relevant - other; the last index wins ties, update_utils.py:25-64),
 *   parent relevant  = (s_mx + c_rel) - max(oth, rel),
 *   parent other     = max((c_oth + s_oth) - oth, parent relevant + min(0, m1)),
 *                      with m2 for the argmax parent,
 *   child relevant   = s_mx + min(0, m1), child other = s_oth,
 *   one-parent factors pass the child's vtof (logical.py:739-757);
 * normalised outputs (relevant - other, other state 0) and 0 in every entry
 * no update writes. Sums over parents and over a variable state's messages
 * use the chunked order (runs of CHUNK from 0, groups of CHUNK runs, group
 * sums in order). One graph; every loop over independent segments runs in
 * parallel, each segment's own reductions in array order, so the results do
 * not depend on the thread count. */

static float nminf(float a, float b) { return (a < b || a != a) ? a : b; }

static float chunked_sum(const float* v, const int64_t* idx, int64_t n) {
  float tot = 0.f, grp = 0.f, run = 0.f;
  for (int64_t i = 0; i < n; ++i) {
    run += v[idx[i]];
    if ((i + 1) % CHUNK == 0) {
      grp += run;
      run = 0.f;
      if ((i + 1) % (CHUNK * CHUNK) == 0) {
        tot += grp;
        grp = 0.f;
      }
    }
  }
  if (n % CHUNK) grp += run;
  if (n % (CHUNK * CHUNK)) tot += grp;
  return tot;
}

/* CSR of key[0 .. n) with values in [0, m): ptr[m + 1], idx[n] (positions in
 * increasing order within each key). */
static int csr_of(const int32_t* key, int64_t n, int64_t m, int64_t** ptr, int64_t** idx) {
  *ptr = (int64_t*)calloc((size_t)m + 1, sizeof(int64_t));
  *idx = (int64_t*)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
  int64_t* fill = (int64_t*)malloc(sizeof(int64_t) * (size_t)(m > 0 ? m : 1));
  if (!*ptr || !*idx || !fill) {
    free(fill);
    return -1;
  }
  for (int64_t i = 0; i < n; ++i) (*ptr)[key[i] + 1]++;
  for (int64_t k = 0; k < m; ++k) (*ptr)[k + 1] += (*ptr)[k];
  for (int64_t k = 0; k < m; ++k) fill[k] = (*ptr)[k];
  for (int64_t i = 0; i < n; ++i) (*idx)[fill[key[i]]++] = i;
  free(fill);
  return 0;
}

/* One logical factor type: msgs range start m0, parents (factor index,
 * message index relative to m0), children (edge state relative to m0),
 * edge_states_offset off (+1 OR, -1 AND). */
typedef struct {
  int64_t m0, num_parents, num_factors;
  const int32_t *pfi, *pmi, *ces;
  int32_t off;
} lbp_logical_type;

static void logical_t0(const lbp_logical_type* T, const int64_t* fptr, const int64_t* fpar,
                       const float* vt, float* out, float* oth_buf, float* mx_buf) {
  const float* v = vt + T->m0;
  float* o = out + T->m0;
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t f = 0; f < T->num_factors; ++f) {
    const int64_t a = fptr[f], z = fptr[f + 1];
    float m1 = -INFINITY;
    for (int64_t k = a; k < z; ++k) {
      const int64_t j = fpar[k];
      const float rel = v[T->pmi[j] + T->off], oth = v[T->pmi[j]];
      oth_buf[j] = oth;
      mx_buf[j] = nmaxf(oth, rel);
      m1 = nmaxf(m1, rel - oth);
    }
    /* the last parent whose diff equals the maximum (else the last parent) */
    int64_t a1 = z > a ? fpar[z - 1] : -1;
    for (int64_t k = z - 1; k >= a; --k) {
      const int64_t j = fpar[k];
      if (v[T->pmi[j] + T->off] - v[T->pmi[j]] == m1) {
        a1 = j;
        break;
      }
    }
    float m2 = -INFINITY;
    for (int64_t k = a; k < z; ++k) {
      const int64_t j = fpar[k];
      if (j != a1) m2 = nmaxf(m2, v[T->pmi[j] + T->off] - v[T->pmi[j]]);
    }
    const float s_oth = chunked_sum(oth_buf, fpar + a, z - a);
    const float s_mx = chunked_sum(mx_buf, fpar + a, z - a);
    const float c_rel = v[T->ces[f] + T->off], c_oth = v[T->ces[f]];
    const float mz = nminf(0.f, m1), mz2 = nminf(0.f, m2);
    for (int64_t k = a; k < z; ++k) {
      const int64_t j = fpar[k];
      float par_rel = (s_mx + c_rel) - mx_buf[j];
      const float opt1 = (c_oth + s_oth) - oth_buf[j];
      const float opt2 = par_rel + (j == a1 ? mz2 : mz);
      float par_oth = nmaxf(opt1, opt2);
      if (z - a == 1) {
        par_rel = c_rel;
        par_oth = c_oth;
      }
      o[T->pmi[j] + T->off] = par_rel - par_oth;
    }
    o[T->ces[f] + T->off] = (s_mx + mz) - s_oth;
  }
}

/* Runs `iters` max-product iterations on one graph of OR and AND factors
 * (types[0 .. num_types)), from ftov_in into ftov_out, deltas[iters].
 * Edge states of one edge must be consecutive. Returns 0, or -1 if memory
 * runs out or the edge states of an edge are not consecutive. */
int lbp_cpu_run_logical_t0(int64_t num_es, int64_t num_vs, int64_t num_edges,
                           const int32_t* var_state_of_es, const int32_t* edge_of_es,
                           int32_t num_types, const int64_t* type_sizes,
                           const int32_t* const* type_arrays, const int32_t* type_offs,
                           const float* V, const float* ftov_in, float* ftov_out,
                           float* deltas, int32_t iters, float damping) {
  if (iters < 1) iters = 1;
  int failed = 0;
  const size_t nes = (size_t)(num_es > 0 ? num_es : 1);
  float* msgs = (float*)malloc(sizeof(float) * nes);
  float* nw = (float*)malloc(sizeof(float) * nes);
  float* vt = (float*)malloc(sizeof(float) * nes);
  float* vsum = (float*)malloc(sizeof(float) * (size_t)(num_vs > 0 ? num_vs : 1));
  int64_t* eptr = (int64_t*)malloc(sizeof(int64_t) * (size_t)(num_edges + 1));
  int64_t *vptr = NULL, *vidx = NULL;
  lbp_logical_type T[2];
  int64_t* fptr[2] = {NULL, NULL};
  int64_t* fpar[2] = {NULL, NULL};
  float* pbuf[2] = {NULL, NULL};
  float* mbuf[2] = {NULL, NULL};
  if (!msgs || !nw || !vt || !vsum || !eptr || num_types > 2 ||
      csr_of(var_state_of_es, num_es, num_vs, &vptr, &vidx))
    failed = 1;
  for (int32_t t = 0; t < num_types && !failed; ++t) {
    T[t].m0 = type_sizes[3 * t];
    T[t].num_parents = type_sizes[3 * t + 1];
    T[t].num_factors = type_sizes[3 * t + 2];
    T[t].pfi = type_arrays[3 * t];
    T[t].pmi = type_arrays[3 * t + 1];
    T[t].ces = type_arrays[3 * t + 2];
    T[t].off = type_offs[t];
    const size_t np = (size_t)(T[t].num_parents > 0 ? T[t].num_parents : 1);
    pbuf[t] = (float*)malloc(sizeof(float) * np);
    mbuf[t] = (float*)malloc(sizeof(float) * np);
    if (!pbuf[t] || !mbuf[t] ||
        csr_of(T[t].pfi, T[t].num_parents, T[t].num_factors, &fptr[t], &fpar[t]))
      failed = 1;
  }
  if (!failed) {
    /* first edge state of each edge (edge states of an edge consecutive) */
    for (int64_t e = 0; e <= num_edges; ++e) eptr[e] = -1;
    for (int64_t i = num_es - 1; i >= 0; --i) eptr[edge_of_es[i]] = i;
    eptr[num_edges] = num_es;
    for (int64_t i = 0; i < num_es && !failed; ++i)
      if (i > 0 && edge_of_es[i] != edge_of_es[i - 1] && edge_of_es[i] != edge_of_es[i - 1] + 1)
        failed = 1;
  }
  if (!failed) {
    memcpy(msgs, ftov_in, sizeof(float) * (size_t)num_es);
    float* em = (float*)malloc(sizeof(float) * (size_t)(num_edges > 0 ? num_edges : 1));
    if (!em) failed = 1;
    if (!failed) normalize(edge_of_es, num_es, num_edges, msgs, em);
    free(em);
  }
  for (int32_t it = 0; it < iters && !failed; ++it) {
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < num_vs; ++s)
      vsum[s] = chunked_sum(msgs, vidx + vptr[s], vptr[s + 1] - vptr[s]) + V[s];
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < num_es; ++i) {
      vt[i] = vsum[var_state_of_es[i]] - msgs[i];
      nw[i] = 0.f;
    }
    for (int32_t t = 0; t < num_types; ++t) logical_t0(&T[t], fptr[t], fpar[t], vt, nw, pbuf[t], mbuf[t]);
    float d = 0.f;  /* max |change|, NaN if any change is NaN */
#pragma omp parallel
    {
      float dt = 0.f;
#pragma omp for schedule(static)
      for (int64_t e = 0; e < num_edges; ++e) {
        float mx = -INFINITY;
        for (int64_t i = eptr[e]; i < eptr[e + 1]; ++i) {
          nw[i] = damping * msgs[i] + (1.f - damping) * nw[i];
          mx = nmaxf(mx, nw[i]);
        }
        for (int64_t i = eptr[e]; i < eptr[e + 1]; ++i) {
          nw[i] = nmaxf(nw[i] - mx, MSG_NEG_INF);
          dt = nmaxf(fabsf(nw[i] - msgs[i]), dt);
        }
      }
#pragma omp critical
      d = nmaxf(dt, d);
    }
    deltas[it] = d;
    float* tmp = msgs;
    msgs = nw;
    nw = tmp;
  }
  if (!failed) memcpy(ftov_out, msgs, sizeof(float) * (size_t)num_es);
  for (int t = 0; t < 2; ++t) {
    free(fptr[t]);
    free(fpar[t]);
    free(pbuf[t]);
    free(mbuf[t]);
  }
  free(msgs);
  free(nw);
  free(vt);
  free(vsum);
  free(eptr);
  free(vptr);
  free(vidx);
  return failed ? -1 : 0;
}
