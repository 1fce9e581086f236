/*
 * C/OpenMP restatement of the reference's BP iteration for Enum factors --
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
 * Cody-Waite reduction and a degree-7 Taylor polynomial in fused
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
  float p = 1.f / 5040.f;
  p = fmaf(p, r, 1.f / 720.f);
  p = fmaf(p, r, 1.f / 120.f);
  p = fmaf(p, r, 1.f / 24.f);
  p = fmaf(p, r, 1.f / 6.f);
  p = fmaf(p, r, 0.5f);
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
