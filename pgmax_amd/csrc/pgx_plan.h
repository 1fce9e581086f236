// Host-side compilation of the reference's flat wiring into the engine's
// device layout ("plan"). Shared between the plan compiler (pgx_plan.cpp)
// and the kernels (pgx_kernels.hip).
#ifndef PGX_PLAN_H_
#define PGX_PLAN_H_

#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/pgx.h"

namespace pgx {

// Factor block kinds. A block is a run of factors of one task that share one
// code path and whose edge / edge-state / log-potential offsets advance
// arithmetically (Enum) or through the factor CSR (logical, pool).
enum BlockKind : int32_t {
  kEnumBin2 = 0,   // pairwise, 2 x 2 states, all 4 configs in row-major order
  kEnumPair = 1,   // pairwise, S0 x S1 <= 8 x 8, all configs in row-major order
  kEnumGeneral = 2,
  kOr = 3,
  kAnd = 4,
  kPool = 5,
};

struct Block {
  int32_t kind;
  int32_t task;
  int32_t f0;      // first factor (global id)
  int32_t count;   // number of factors
  int32_t e0;      // first edge of factor f0
  int32_t es0;     // first edge state of factor f0
  int32_t nv;      // Enum: edges per factor
  int32_t esf;     // Enum: edge states per factor
  int32_t c0;      // Enum: absolute log-potential index of factor f0, config 0
  int32_t nc;      // Enum: configs per factor
  int32_t tbl;     // Enum general: offset of the config table in `tables`
  int32_t s0, s1;  // Enum pairwise: state counts
  int32_t off;     // logical: +1 / -1 (Pool: +1)
  int32_t pad[2];  // pad[0] bit 0: Enum 2x2 block with 16-byte aligned log-potentials,
                   // bit 1: 2x2 block stored planar (see Plan::perm),
                   // bit 2: OR/AND block run by lane groups (wide factors),
                   // bit 3: general Enum block run by one wavefront per factor,
                   // bit 4: OR/AND block run by one wavefront per factor
};

// Long sums (incident edges of a variable, parents of a logical factor) run
// in runs of kVarChunk consecutive terms, one lane per run, and the run sums
// (in groups of kVarChunk) are added in order -- the order the CPU oracle
// uses too.
constexpr int32_t kVarChunk = 32;
constexpr int32_t kWideRuns = 32;  // runs of parents one lane group of a logical factor covers
constexpr int32_t kWaveLogicalParents = 256;  // parents of a logical factor on one wavefront
// General Enum factors run one wavefront per factor when their config
// summaries and edge-state vtof fit the wavefront's LDS.
constexpr int32_t kWaveMaxConfigs = 1024;
constexpr int32_t kWaveMaxStates = 256;
constexpr int64_t kPos4MaxBytes = int64_t(1) << 30;        // per-variable descriptors
constexpr int64_t kEdgePos4MaxBytes = int64_t(64) << 20;   // per-edge descriptors

struct Task {
  int32_t v0, v1;   // variables [v0, v1)
  int32_t b0, b1;   // blocks [b0, b1)
  int32_t h0, h1;   // entries [h0, h1) of heavy_vars
  int64_t work;     // edge states, for ordering
};

struct Plan {
  int64_t num_es = 0, num_vs = 0, num_lp = 0;
  int32_t num_edges = 0, num_vars = 0, num_factors = 0;
  int64_t msgs_range[4][2] = {{0, 0}};
  int64_t pot_range[4][2] = {{0, 0}};
  std::vector<int32_t> var_vs0;    // [NV+1] first variable state of each variable
  std::vector<int32_t> var_eptr;   // [NV+1] CSR over incident edges
  std::vector<int32_t> var_adj_es; // [E] first edge state of each incident edge, ascending
  std::vector<int32_t> edge_es0;   // [E+1] first edge state of each edge
  std::vector<int32_t> edge_vs0;   // [E] first variable state of each edge's variable
  std::vector<int32_t> fac_e0;     // [F+1] first edge of each factor
  std::vector<int32_t> tables;     // Enum general config tables (layout at Block::tbl:
                                   // [nv edge offsets][esf owning edge][nc * nv rows]
                                   // [esf + 1 list offsets][configs of each edge state])
  std::vector<int32_t> heavy_vars; // variables of degree > kVarChunk, grouped by task
  std::vector<int32_t> heavy_cptr; // [NH+1] first chunk of each heavy variable
  std::vector<int32_t> heavy_pofs; // [NH+1] first partial-sum slot (chunks x states)
  // Engine message layout. With a planar layout every 2x2 Enum block of n
  // factors keeps the two states of each factor's first edge in one plane
  // [es0, es0 + 2n) and those of its second edge in the next [es0 + 2n,
  // es0 + 4n), so consecutive factors touch consecutive 8-byte pairs on both
  // the factor and the variable side. perm[es] is the engine position of the
  // reference's edge state es (empty: identity); var_adj_es holds engine
  // positions, var_adj_es_ref the reference's.
  std::vector<int32_t> perm;
  std::vector<int32_t> var_adj_es_ref;
  std::vector<int32_t> heavy_chunk_var;  // [NHC] heavy-variable index of each chunk
  std::vector<int32_t> heavy_chunk_k;    // [NHC] chunk number within its variable
  // Variables of more than kVarChunk chunks also sum their chunk sums in
  // groups of kVarChunk: heavy_gpofs[h] is the first group-sum slot (after
  // all chunk slots) or -1; (vh_group_h, vh_group_j) enumerate the groups.
  // Chunks split for the per-phase variable kernel: "row" chunks hold a
  // binary variable's 32 edges at consecutive positions (one 8-lane group
  // reads them together), "col" chunks the rest (one thread each).
  std::vector<int32_t> heavy_chunk_stride;  // [NHC] position step of a full chunk whose
                                            // edges are evenly spaced, else 0
  // Chunk descriptors, 4 ints each, so a kernel reads one record per chunk:
  // [first adjacency index (row chunks: first engine position), len | ns << 8,
  //  partial-sum slot, position step or 0]. chunk_desc in chunk order,
  // chunk_desc_ref the same with step 0 (reference layout), col_desc /
  // row_desc in the order of heavy_col_chunks / heavy_row_chunks.
  std::vector<int32_t> chunk_desc, chunk_desc_ref, col_desc, row_desc;
  // Heavy-variable descriptors, 4 ints each: [first variable state,
  // chunks | ns << 16, first chunk slot, first group slot or -1].
  std::vector<int32_t> heavy_desc;
  std::vector<int32_t> heavy_col_chunks;
  std::vector<int32_t> heavy_row_chunks;
  std::vector<int32_t> heavy_gpofs;
  std::vector<int32_t> vh_group_h;
  std::vector<int32_t> vh_group_j;
  int64_t partial_size = 0;              // chunk + group slots per batch element
  // (a task's chunks occupy [heavy_cptr[h0], heavy_cptr[h1]) in chunk-major order)
  std::vector<Block> blocks;
  std::vector<Task> tasks;         // ordered by descending work
  int32_t max_esf = 0;             // largest Enum edge-state count per factor
  std::vector<int32_t> edge_var;   // [E] variable of each edge
  int32_t max_degree = 0;          // most edges on one variable
  bool pairwise_only = false;      // every factor is a full-table pairwise Enum factor
  // Variables of at most 4 edges: their edges' engine positions in edge
  // order, 4 ints per variable (-1 past the degree; -2 in slot 0 for more
  // than 4 edges), so a variable sum needs one index load. edge_pos4 holds
  // the entry of each edge's variable (graphs whose factor side recomputes
  // variable sums). Empty past kPos4MaxBytes / kEdgePos4MaxBytes.
  std::vector<int32_t> var_pos4, edge_pos4;
  // Every factor a planar 2x2 Enum factor and every variable binary with at
  // most 4 edges: the per-phase shape may keep one float per message.
  bool packable = false;
  bool even2 = true;               // every 2-state edge and variable starts at an even index
  // Complete bipartite binary graphs (a binary RBM): the only block is a
  // planar 2x2 block whose factor i = a * bip_nb + b joins row variable a
  // (first edge) and column variable b (second edge), and every variable
  // has more than kVarChunk edges. Then a row's edges in edge order are its
  // factors b = 0, 1, ... and a column's its factors a = 0, 1, ..., so a
  // tile of kVarChunk rows by a multiple of kVarChunk columns holds whole
  // chunks of both sides. bip_row[2a], bip_row[2a + 1]: first variable state
  // and first chunk slot of row a; bip_col the same for columns.
  // bip_unary = 1: a block of one unary factor per unit (2 configs, row
  // units' first, then column units') precedes the pairwise block, and is
  // every unit's first edge (the RBM with its biases as unary Enum factors,
  // benchmark/rbm_lib.py:140-168). A unit's edges are then its unary factor
  // and its pairwise factors, and the tiles run over a grid of (na + 1) x
  // (nb + 1) with the unary factors as row 0 and column 0, which keeps every
  // 32-edge run inside one tile. bip_pb: the pairwise block; bip_uc0: the
  // first log-potential of the unary block.
  int32_t bip_na = 0, bip_nb = 0;
  int32_t bip_unary = 0, bip_pb = 0, bip_uc0 = 0;
  std::vector<int32_t> bip_row, bip_col;
  // When both tables are affine (entry i = x0 + i * dx), bip_aff holds
  // [row state x0, dx, row slot x0, dx, column state x0, dx, column slot x0,
  // dx] and kernels compute the entries instead of loading them.
  bool bip_affine = false;
  int32_t bip_aff[8] = {0, 0, 0, 0, 0, 0, 0, 0};
};

// Builds `plan` from the wiring. Neighbouring components whose combined
// edge + variable states stay below merge_below share one task (0: never).
// Returns PGX_OK or an error code with `err`.
int build_plan(const pgx_wiring& w, Plan* plan, std::string* err, int64_t merge_below,
               bool planar = false);

}  // namespace pgx

#endif  // PGX_PLAN_H_
