/*
 * pgx.h -- C ABI of the MI355X loopy-belief-propagation engine.
 *
 * The boundary is the reference's BP hot path: the compiled FactorGraph
 * wiring (flat int arrays) plus the three flat float arrays of BPArrays.
 * Everything is plain pointers and sizes; no C++ or torch types cross it.
 *
 * Conventions
 *   - Index arrays passed to pgx_graph_create are HOST pointers (int32).
 *   - Float arrays passed to pgx_run / pgx_get_beliefs / pgx_fac_to_var are
 *     DEVICE pointers (float32) in the current HIP device's memory.
 *   - Every call returns 0 on success or a negative PGX_E* code; the text of
 *     the last error on the calling thread is available from pgx_last_error.
 *   - Calls are ordered on the given stream (hipStream_t, NULL = default
 *     stream) and do not synchronise the host except pgx_graph_create.
 *   - A pgx_graph may be used from one host thread at a time.
 */
#ifndef PGX_H_
#define PGX_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PGX_OK 0
#define PGX_EINVAL -1   /* malformed arguments or wiring */
#define PGX_EUNSUP -2   /* wiring form this engine does not handle */
#define PGX_EHIP -3     /* HIP runtime error */
#define PGX_ENOMEM -4   /* device allocation failed */

/* Factor types in the reference's registry order (pgmax/factor/__init__.py:37-44). */
#define PGX_ENUM 0
#define PGX_OR 1
#define PGX_AND 2
#define PGX_POOL 3

/* enum_num_rows value meaning "no rows given": every Enum factor holds the
 * full table of its variables' states, configs in row-major order (last
 * variable fastest), which is the order pgmax/fgroup/enum.py:303-309 gives
 * pairwise factors. The row pointers are then ignored. */
#define PGX_ENUM_FULL_TABLES (-1)

/* Logical (OR/AND) or Pool wiring of one factor type; indices are relative
 * to the type's message range, as returned by get_inference_arguments
 * (pgmax/factor/logical.py:76-104, pgmax/factor/pool.py:62-83). */
typedef struct pgx_logical_wiring {
  int64_t num_parents;                 /* length of the two parent arrays */
  int64_t num_factors;                 /* length of children_edge_states */
  const int32_t* parents_factor_indices;
  const int32_t* parents_msg_indices;
  const int32_t* children_edge_states;
  int32_t edge_states_offset;          /* +1 OR and Pool, -1 AND */
} pgx_logical_wiring;

/* The compiled wiring of a factor graph (pgmax/infer/inferer.py:65-118). */
typedef struct pgx_wiring {
  int64_t num_edge_states;             /* E_s: length of ftov_msgs */
  int64_t num_var_states;              /* V_s: length of evidence */
  int64_t num_log_potentials;          /* C: length of log_potentials */
  /* per edge state (pgmax/infer/inferer.py:84-91) */
  const int32_t* var_states_for_edge_states;
  const int32_t* edge_indices_for_edge_states;
  const int32_t* factor_indices_for_edge_states;
  /* variable index of each variable state (pgmax/fgraph/fgraph.py:317-326) */
  const int32_t* evidence_to_vars;
  /* [type][0..1]: message range and log-potential range of each factor
   * type (pgmax/fgraph/fgraph.py:185-232) */
  int64_t msgs_range[4][2];
  int64_t potentials_range[4][2];
  /* EnumFactor wiring (pgmax/factor/enum.py:54-72), rows relative to the
   * Enum ranges, or enum_num_rows = PGX_ENUM_FULL_TABLES */
  int64_t enum_num_rows;
  const int32_t* enum_factor_configs_indices;
  const int32_t* enum_factor_configs_edge_states;
  /* OR, AND, Pool wirings, indexed by PGX_OR / PGX_AND / PGX_POOL */
  pgx_logical_wiring logical[4];
} pgx_wiring;

typedef struct pgx_graph pgx_graph;

/* Compile the wiring into device-resident tables.
 * Replaces the per-run index work of pgmax/infer/inferer.py:65-118 and the
 * static arguments of FAC_TO_VAR_UPDATES (pgmax/infer/bp.py:107-123). */
int pgx_graph_create(const pgx_wiring* wiring, pgx_graph** out);

int pgx_graph_destroy(pgx_graph* graph);

/* Validate and compile a wiring on the host only (no GPU needed); fills
 * the same 8 sizes as pgx_graph_info. */
int pgx_wiring_check(const pgx_wiring* wiring, int64_t* out8);

/* Sizes of the compiled graph: [E_s, V_s, C, num_edges, num_variables,
 * num_factors, num_tasks, num_blocks]. */
int pgx_graph_info(const pgx_graph* graph, int64_t* out8);

/* One BP run: pgmax/infer/bp.py:63-155 (run_with_diffs).
 * Batch b reads log_potentials + b*lp_stride, evidence + b*ev_stride,
 * ftov_in + b*ftov_in_stride (a stride of 0 shares one array across the
 * batch, as an un-mapped argument of jax.vmap would) and writes
 * ftov_out + b*E_s and deltas + b*max(num_iters,1). num_iters <= 1 runs one
 * iteration (pgmax/infer/bp.py:142-146). ftov_out must not alias ftov_in. */
int pgx_run(pgx_graph* graph,
            const float* log_potentials, int64_t lp_stride,
            const float* evidence, int64_t ev_stride,
            const float* ftov_in, int64_t ftov_in_stride,
            float* ftov_out, float* deltas,
            int32_t batch, int32_t num_iters, float damping,
            float temperature, void* stream);

/* pgx_run without any allocation, for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc). No counterpart in the reference. The same arguments and results
 * as pgx_run, with the run's scratch (message buffers, variable sums,
 * partial sums, delta slots) in a device array the library holds on each GPU
 * it runs on until the process ends: 64 MiB, so a run needing more scratch
 * (messages of batch x E_s floats, twice for planar layouts, plus batch x
 * V_s variable sums, ...) returns PGX_EINVAL with the size it would need.
 * The array is shared by every call on a GPU and keeps the last call's
 * contents: calls that overlap on one GPU (different streams, threads)
 * race. Batches above 65535 are refused; profiling and graph capture
 * options are ignored. */
int pgx_run_noalloc(pgx_graph* graph,
                    const float* log_potentials, int64_t lp_stride,
                    const float* evidence, int64_t ev_stride,
                    const float* ftov_in, int64_t ftov_in_stride,
                    float* ftov_out, float* deltas,
                    int32_t batch, int32_t num_iters, float damping,
                    float temperature, void* stream);

/* BP run to convergence, per independent task (connected component, or
 * merged run of small components; see the "merge" option): every task of
 * every batch element stops after the first iteration whose max-abs
 * message change (pgmax/infer/bp.py:136, over the task's messages) is
 * <= tol, or after max_iters. ftov_out then holds exactly the messages a
 * pgx_run with that task's iteration count would give. task_deltas
 * (batch x num_tasks x max(max_iters,1), zero past the stop) and
 * iters_done (batch x num_tasks) are written in the task order of
 * pgx_graph_tasks. No reference counterpart: the reference always runs a
 * fixed count with one global delta. */
int pgx_run_converge(pgx_graph* graph,
                     const float* log_potentials, int64_t lp_stride,
                     const float* evidence, int64_t ev_stride,
                     const float* ftov_in, int64_t ftov_in_stride,
                     float* ftov_out, float* task_deltas, int32_t* iters_done,
                     int32_t batch, int32_t max_iters, float damping,
                     float temperature, float tol, void* stream);

/* Autotuner: times every execution shape this graph allows (one workgroup
 * per task with 256 / 512 / 1024 threads, per-phase launches, with and
 * without factor-side variable sums and vector accesses) on the given
 * inputs for probe_iters iterations, and makes the fastest the graph's
 * shape for later pgx_run (converge = 0) or pgx_run_converge (converge = 1)
 * calls with the same batch size and temperature class (0 or > 0), as long
 * as the "mode", "block" and "tile" options are 0. Per-phase shapes are
 * tried with 64, 128 and 256 factors per workgroup, the complete bipartite
 * iteration with tiles of 32 rows x 32, 64 or 128 columns and 256 or 512
 * threads, and per-phase runs with square pairwise tables both on S lanes
 * per factor and one factor per thread, with and without runs of
 * consecutive workgroups per XCD. chosen (7 int64, may be NULL) receives
 * [mode, block, fuse, vec, tile, lanes, xcd]. Synchronises the host. No
 * reference counterpart. */
int pgx_autotune(pgx_graph* graph,
                 const float* log_potentials, int64_t lp_stride,
                 const float* evidence, int64_t ev_stride,
                 const float* ftov_in, int64_t ftov_in_stride,
                 int32_t batch, float damping, float temperature,
                 int32_t converge, int32_t probe_iters, void* stream,
                 int64_t* chosen);

/* The tasks of a compiled graph in execution order: 4 ints each,
 * [first variable, end variable, first block, end block]. out holds
 * 4 * num_tasks ints (num_tasks from pgx_graph_info). */
int pgx_graph_tasks(const pgx_graph* graph, int32_t* out);

/* Beliefs: evidence + scatter-add of ftov per variable state
 * (pgmax/infer/inferer.py:211-225). beliefs is batch x V_s. */
int pgx_get_beliefs(pgx_graph* graph,
                    const float* evidence, int64_t ev_stride,
                    const float* ftov, int64_t ftov_stride,
                    float* beliefs, int32_t batch, void* stream);

/* MAP decoding of flat beliefs: index of the first maximal state of every
 * variable (pgmax/infer/inferer.py:251-264). map_states is batch x NV. */
int pgx_decode_map(pgx_graph* graph, const float* beliefs,
                   int32_t* map_states, int32_t batch, void* stream);

/* Beliefs, MAP states and marginals in one call: beliefs as pgx_get_beliefs
 * (pgmax/infer/inferer.py:218-222); then, in one pass over each variable's
 * beliefs, map_states (batch x NV, the first maximal state; a NaN counts as
 * maximal; pgmax/infer/inferer.py:251-264) and marginals (batch x V_s,
 * exp(b - max) / sum over the variable's states, summed in state order, the
 * softmax of pgmax/infer/bp.py:263-288). beliefs is required (it holds the
 * intermediate); map_states and marginals may be NULL. */
int pgx_beliefs_decode(pgx_graph* graph,
                       const float* evidence, int64_t ev_stride,
                       const float* ftov, int64_t ftov_stride,
                       float* beliefs, int32_t* map_states, float* marginals,
                       int32_t batch, void* stream);

/* The factor-to-variable update of one factor type over the whole graph:
 * FAC_TO_VAR_UPDATES[type](vtof_msgs[range], log_potentials[range],
 * temperature, normalize) of pgmax/factor/enum.py:397-475,
 * pgmax/factor/logical.py:491-779, pgmax/factor/pool.py:275-474.
 * vtof and ftov are full-length (E_s) arrays; only the type's message range
 * of ftov is written. normalize=0 gives OR, AND and Pool their unnormalised
 * output (pgmax/factor/logical.py:769-778, pgmax/factor/pool.py:461-473); Enum
 * factors ignore it, as the reference does (pgmax/factor/enum.py:398-446). */
int pgx_fac_to_var(pgx_graph* graph, int32_t factor_type,
                   const float* vtof, const float* log_potentials,
                   float* ftov, float temperature, int32_t normalize,
                   void* stream);

/* ---- reverse mode --------------------------------------------------------
 * The vector-Jacobian product of pgx_run, for graphs whose factors are all
 * Enum factors: what jax.grad through BP.run computes
 * (pgmax/infer/bp.py:63-155, differentiated in examples/grid_mrf.ipynb
 * cells 15-16). pgx_vjp_create takes the graph and the reference-layout
 * arrays the reverse pass walks: var_states_for_edge_states and
 * edge_indices_for_edge_states (E_s each, pgmax/infer/inferer.py:84-88)
 * and the Enum rows (factor_configs_indices / _edge_states,
 * pgmax/factor/enum.py:54-72; full tables listed row by row), ordered by
 * config. A graph with OR or AND factors also needs pgx_vjp_set_logical for
 * each of the two types it holds, one with Pool factors pgx_vjp_set_pool. */
typedef struct pgx_vjp pgx_vjp;
int pgx_vjp_create(pgx_graph* graph, const int32_t* var_states_for_edge_states,
                   const int32_t* edge_indices_for_edge_states, const int32_t* row_cfg,
                   const int32_t* row_es, int64_t num_rows, pgx_vjp** out);
int pgx_vjp_destroy(pgx_vjp* vjp);

/* The OR (factor_type PGX_OR, edge_states_offset 1) or AND (PGX_AND, -1)
 * factors of the graph, as LogicalWiring holds them
 * (pgmax/factor/logical.py:457-488): parents_factor_indices and
 * parents_msg_indices [num_parents], children_edge_states [num_factors],
 * message indices relative to the type's message range. */
int pgx_vjp_set_logical(pgx_vjp* vjp, int32_t factor_type, const int32_t* parents_factor_indices,
                        const int32_t* parents_msg_indices, int64_t num_parents,
                        const int32_t* children_edge_states, int64_t num_factors,
                        int32_t edge_states_offset);

/* The Pool factors of the graph, as PoolWiring holds them
 * (pgmax/factor/pool.py:275-326): pool_choices_factor_indices and
 * pool_choices_msg_indices [num_choices], pool_indicators_edge_states
 * [num_factors], relative to the Pool message range. */
int pgx_vjp_set_pool(pgx_vjp* vjp, const int32_t* choices_factor_indices,
                     const int32_t* choices_msg_indices, int64_t num_choices,
                     const int32_t* indicators_edge_states, int64_t num_factors);

/* Replays pgx_run(lp, ev, ftov_in, num_iters, damping, temperature) one
 * iteration at a time (kept in the handle's scratch: (num_iters + 1) x batch
 * x E_s floats, at most 64 GiB), then pulls g_out (batch x E_s, the
 * cotangent of the final messages) back through every iteration. Writes
 * g_lp, g_ev, g_ftov_in with the batch layout of their inputs: batch x size
 * for a batched input, size (summed over the batch) for a shared one
 * (stride 0). A maximum attained by several entries shares its derivative
 * equally; max and clip pass half at equality. No reference counterpart as
 * a C call: the reference gets this from JAX's autodiff. */
int pgx_run_vjp(pgx_vjp* vjp, const float* log_potentials, int64_t lp_stride,
                const float* evidence, int64_t ev_stride, const float* ftov_in,
                int64_t ftov_in_stride, const float* g_out, float* g_lp, float* g_ev,
                float* g_ftov_in, int32_t batch, int32_t num_iters, float damping,
                float temperature, void* stream);

/* EnumFactor update on raw rows, for any (config, edge-state) row list
 * (pgmax/factor/enum.py:451-475 as written, no structure assumed):
 *   summary[c] = (sum of vtof[row_es[r]] over rows r of config c) + lp[c]
 *   ftov[es]   = max / T-logsumexp of summary over the rows of es - vtof[es]
 * cfg_row_ptr[num_configs+1] delimits each config's rows in row_es[];
 * es_row_ptr[num_es+1] delimits each edge state's entries in es_row_cfg[],
 * the configs of its rows in row order. summary is device scratch of
 * num_configs floats. All pointers are device pointers. */
int pgx_enum_rows_fac_to_var(const int32_t* cfg_row_ptr, const int32_t* row_es,
                             const int32_t* es_row_ptr, const int32_t* es_row_cfg,
                             int64_t num_configs, int64_t num_es,
                             const float* vtof, const float* log_potentials,
                             float* summary, float* ftov, float temperature,
                             void* stream);

/* Engine options: "mode" (0 auto, 1 one workgroup per task, 2 per-phase
 * launches, 3 one cooperative launch for all iterations of a graph whose
 * factor side recomputes variable sums, see "fuse"), "block" (threads per
 * workgroup of the task loop and of the complete bipartite iteration:
 * 128, 256 or 512 there; 0 auto), "profile" (1: record
 * HIP events around the BP iterations of every pgx_run; 2: also around every
 * kernel launch), "fuse" (1, the
 * default: graphs of pairwise factors whose variables have at most 8 edges
 * recompute variable sums on the factor side -- without tuning only for
 * runs of at most 2^24 edge states; 2: at every size --, and complete
 * bipartite binary graphs (a binary RBM) form them from chunk sums the factor
 * kernel leaves behind -- one launch per iteration either way; 0: separate
 * variable phase),
 * "graph" (1: pgx_run records its launches as a HIP graph once per set of
 * arguments -- pointers, strides, sizes, damping, temperature, options --
 * and replays it on later calls with the same arguments; at most 8 graphs
 * per pgx_graph, freed with it; default 0),
 * "tile" (factors and threads per workgroup of the per-phase factor
 * kernels: 64, 128 or 256; columns per 32-row tile of the complete
 * bipartite iteration: 32, 64 or 128; 0 auto), "planar" (read by pgx_graph_create: 1,
 * the default, stores the messages of each block of 2x2 Enum factors as two
 * planes, first edges then second edges, inside the engine; results are
 * returned in the reference layout either way),
 * "vec" (bit 0: 8-byte loads of binary edge messages, bit 1: 16-byte loads
 * of 2x2 log-potential tables, each used only where the buffers are aligned,
 * bit 2: those loads non-temporal; default 3), "pack" (1, the default:
 * complete bipartite binary graphs, and in the per-phase shape graphs of
 * 2x2 factors whose variables have at most 4 edges, iterate on one float per
 * binary message -- normalisation leaves one state exactly 0 -- decoded to
 * the same values; 0: two floats, except in the complete bipartite
 * iteration), "merge" (read by pgx_graph_create: neighbouring components with fewer
 * edge + variable states than this share a task; 0 keeps every component
 * its own task; default 4096), "lanes" (1, the default: in the per-phase
 * shape, square pairwise Enum tables of 3..8 states run on S lanes per
 * factor, one lane per table row and column; 0: one factor per thread),
 * "xcd" (per-phase kernels: each of the 8 XCDs runs runs of this many
 * consecutive workgroups, so neighbouring tiles share an L2; 0 or 1, the
 * default: round-robin order), "share" (1, the default: in runs of 8 or
 * more iterations, one pass per run finds the lane tiles whose factors all
 * hold the same log-potential table, bit for bit, and those tiles read one
 * table instead of one per factor -- same values, same results; 0: every
 * factor reads its own), "mpnn_rows" (1, the default: pgx_mpnn_predict
 * works in variable-major row order when no variable has more than 16 edges,
 * each variable's sum taken inside its tile; 0: edge-major rows with a
 * separate variable-sum pass, as pgx_mpnn_grad's forward pass; both give the
 * same bits). Returns the previous value. */
int64_t pgx_set_option(const char* key, int64_t value);

/* With "profile" 1 or 2: waits for the recorded runs and returns the summed
 * GPU time of their BP iterations (milliseconds, HIP events on the run's
 * stream) and how many runs were recorded, then clears the record. */
int pgx_profile_read(double* total_ms, int64_t* num_runs);

/* Kernel ids of pgx_profile_kernels. */
#define PGX_KERNEL_TASK_LOOP 0
#define PGX_KERNEL_VAR_PHASE 1
#define PGX_KERNEL_VAR_COMBINE 2
#define PGX_KERNEL_FAC_FUSED 3
#define PGX_KERNEL_FAC_PHASE 4   /* + block kind: 2x2 Enum, pairwise Enum,
                                    general Enum, OR, AND, Pool */
#define PGX_KERNEL_BIPARTITE 10 /* complete bipartite binary graphs: whole iteration */
/* pgx_run_vjp on pairwise graphs: variable sums and factor step of the
 * forward replay, factor and variable side of the reverse step */
#define PGX_KERNEL_VJP_VAR_SUMS 11
#define PGX_KERNEL_VJP_PAIR_FWD 12
#define PGX_KERNEL_VJP_PAIR_BWD 13
#define PGX_KERNEL_VJP_VAR_BWD 14
#define PGX_KERNEL_COUNT 15

/* With "profile" 2 every launch of pgx_run / pgx_run_converge (and of the
 * pairwise reverse pass of pgx_run_vjp) is bracketed
 * by HIP events on the run's stream (the back-to-back iteration launches of
 * the complete bipartite shape by one pair around all of them, so their
 * average includes the gaps between them but no event records). Fills total_ms[PGX_KERNEL_COUNT] and
 * launches[PGX_KERNEL_COUNT] (either may be NULL) with the summed GPU time
 * and launch count per kernel id since the last call, then clears them. */
int pgx_profile_kernels(double* total_ms, int64_t* launches);

/* ---- learned warm start ------------------------------------------------
 * No counterpart in the reference: a 3-layer message-passing network
 * (hidden width 64, BF16 activations on MFMA, FP32 accumulation) that
 * predicts initial ftov messages (the BPArrays.ftov_msgs input of
 * pgmax/infer/bp_state.py:172-255) for graphs whose variables are binary
 * and whose factors are pairwise Enum factors with full 2x2 tables.
 * Parameters are one flat FP32 device array of PGX_MPNN_NUM_PARAMS floats:
 * W0[8][64], b0[64], then for l = 1..3 W_l[192][64], b_l[64], then
 * w_h[64], b_h. */
#define PGX_MPNN_NUM_PARAMS 37697

typedef struct pgx_mpnn pgx_mpnn;

/* Network state for a compiled graph (PGX_EUNSUP if the graph is not
 * pairwise binary). The graph must outlive nothing: the network copies
 * the topology it needs. */
int pgx_mpnn_create(const pgx_graph* graph, pgx_mpnn** out);
int pgx_mpnn_destroy(pgx_mpnn* net);

/* Predicted normalised messages for `batch` inputs (strides as pgx_run);
 * ftov_out is batch x E_s. */
int pgx_mpnn_predict(pgx_mpnn* net, const float* params,
                     const float* log_potentials, int64_t lp_stride,
                     const float* evidence, int64_t ev_stride,
                     float* ftov_out, int32_t batch, void* stream);

/* Mean squared error of the predicted log-ratios against those of
 * target_ftov (batch x E_s, stride target_stride) and its gradient with
 * respect to every parameter: grads (PGX_MPNN_NUM_PARAMS floats) and loss
 * (1 float), both device memory. Deterministic for a given batch size. */
int pgx_mpnn_grad(pgx_mpnn* net, const float* params,
                  const float* log_potentials, int64_t lp_stride,
                  const float* evidence, int64_t ev_stride,
                  const float* target_ftov, int64_t target_stride,
                  float* grads, float* loss, int32_t batch, void* stream);

/* One Adam step (Kingma & Ba 2015) on n parameters; step counts from 1. */
int pgx_mpnn_adam(float* params, const float* grads, float* m, float* v,
                  int64_t n, int32_t step, float lr, float beta1, float beta2,
                  float eps, void* stream);

/* No counterpart in the reference: checks, on the device it runs on, two
 * numerical building blocks the kernels replace their textbook forms with,
 * over the float bit patterns first .. first + count - 1, adding the number
 * of mismatches to *mismatches (device memory, one uint64):
 *   which = 0: the hardware FP32 -> BF16 conversion of the warm-start
 *              kernels against integer round-to-nearest-even (add 0x7fff
 *              plus the kept LSB), for each pattern and its negation; NaNs
 *              are not compared;
 *   which = 1: the logarithm's quotient f / (2 + f), f = m - 1 for the
 *              pattern m, against the IEEE division.
 * For tests; returns PGX_EINVAL for another `which` or a null pointer. */
int pgx_check_numerics(int32_t which, uint32_t first, uint32_t count,
                       unsigned long long* mismatches, void* stream);

/* Text of the last error raised on the calling thread ("" if none). */
const char* pgx_last_error(void);

#ifdef __cplusplus
}
#endif

#endif  /* PGX_H_ */
