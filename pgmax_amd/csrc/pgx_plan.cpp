// Compiles the reference's flat wiring (pgmax/infer/inferer.py:65-118 and
// the per-type get_inference_arguments) into the engine's plan: CSR tables
// for the variable and factor sides, deduplicated Enum config tables,
// independent components ("tasks") and arithmetic factor blocks.
//
// The wiring must have the structure pgmax's compile_wiring functions give
// it (pgmax/factor/factor.py:197-225, pgmax/factor/enum.py:363-394,
// pgmax/factor/logical.py:457-488): edges of a factor consecutive, edge
// states of an edge consecutive, Enum rows config-major with the j-th row of
// a config in the factor's j-th edge, logical parents then child. Anything
// else is rejected with PGX_EUNSUP and a message naming the first violation.

#include "pgx_plan.h"

#include <algorithm>
#include <array>
#include <map>
#include <numeric>
#include <sstream>

namespace pgx {
namespace {

struct UnionFind {
  std::vector<int32_t> p;
  explicit UnionFind(int32_t n) : p(n) { std::iota(p.begin(), p.end(), 0); }
  int32_t find(int32_t x) {
    while (p[x] != x) {
      p[x] = p[p[x]];
      x = p[x];
    }
    return x;
  }
  void unite(int32_t a, int32_t b) {
    a = find(a);
    b = find(b);
    if (a != b) p[std::max(a, b)] = std::min(a, b);
  }
};

#define PGX_REQUIRE(cond, code, msg)     \
  do {                                   \
    if (!(cond)) {                       \
      std::ostringstream _os;            \
      _os << msg;                        \
      *err = _os.str();                  \
      return code;                       \
    }                                    \
  } while (0)

struct EnumFactorInfo {
  int32_t c0 = 0, nc = 0;
  bool full = false;           // implicit full table (no rows given)
  std::vector<int32_t> table;  // nc * nv edge-state offsets within the factor
};

// Row-major full table of a factor (last variable fastest): the config
// order of np.mgrid[...].transpose(...).reshape(...) in
// pgmax/fgroup/enum.py:303-309 and of itertools.product.
std::vector<int32_t> full_table(const std::vector<int32_t>& edge_off,
                                const std::vector<int32_t>& states, int32_t nc) {
  const int32_t nv = int32_t(states.size());
  std::vector<int32_t> t(size_t(nc) * nv);
  for (int32_t c = 0; c < nc; ++c) {
    int32_t rem = c;
    for (int32_t j = nv - 1; j >= 0; --j) {
      t[size_t(c) * nv + j] = edge_off[j] + rem % states[j];
      rem /= states[j];
    }
  }
  return t;
}

}  // namespace

int build_plan(const pgx_wiring& w, Plan* P, std::string* err, int64_t merge_below,
               bool planar) {
  const int64_t Es = w.num_edge_states, Vs = w.num_var_states, C = w.num_log_potentials;
  PGX_REQUIRE(Es >= 0 && Vs >= 0 && C >= 0, PGX_EINVAL, "negative sizes");
  PGX_REQUIRE(Es < (int64_t(1) << 31) && Vs < (int64_t(1) << 31) && C < (int64_t(1) << 31),
              PGX_EUNSUP, "arrays longer than 2^31 elements are not supported");
  P->num_es = Es;
  P->num_vs = Vs;
  P->num_lp = C;
  for (int t = 0; t < 4; ++t)
    for (int k = 0; k < 2; ++k) {
      P->msgs_range[t][k] = w.msgs_range[t][k];
      P->pot_range[t][k] = w.potentials_range[t][k];
    }

  // ---- variables -------------------------------------------------------
  const int32_t* e2v = w.evidence_to_vars;
  int32_t NV = 0;
  if (Vs > 0) {
    PGX_REQUIRE(e2v != nullptr, PGX_EINVAL, "evidence_to_vars is null");
    PGX_REQUIRE(e2v[0] == 0, PGX_EUNSUP, "evidence_to_vars must start at variable 0");
    for (int64_t i = 1; i < Vs; ++i)
      PGX_REQUIRE(e2v[i] == e2v[i - 1] || e2v[i] == e2v[i - 1] + 1, PGX_EUNSUP,
                  "evidence_to_vars must number variables consecutively (entry " << i << ")");
    NV = e2v[Vs - 1] + 1;
  }
  P->num_vars = NV;
  P->var_vs0.assign(NV + 1, 0);
  for (int64_t i = Vs - 1; i >= 0; --i) P->var_vs0[e2v[i]] = int32_t(i);
  P->var_vs0[NV] = int32_t(Vs);

  // ---- edges -------------------------------------------------------------
  const int32_t* vsf = w.var_states_for_edge_states;
  const int32_t* ei = w.edge_indices_for_edge_states;
  const int32_t* fi = w.factor_indices_for_edge_states;
  int32_t E = 0;
  if (Es > 0) {
    PGX_REQUIRE(vsf && ei && fi, PGX_EINVAL, "edge-state arrays are null");
    PGX_REQUIRE(ei[0] == 0, PGX_EUNSUP, "edge indices must start at 0");
    for (int64_t i = 1; i < Es; ++i)
      PGX_REQUIRE(ei[i] == ei[i - 1] || ei[i] == ei[i - 1] + 1, PGX_EUNSUP,
                  "edge states of an edge must be consecutive (edge state " << i << ")");
    E = ei[Es - 1] + 1;
  }
  P->num_edges = E;
  P->edge_es0.assign(E + 1, 0);
  for (int64_t i = Es - 1; i >= 0; --i) P->edge_es0[ei[i]] = int32_t(i);
  P->edge_es0[E] = int32_t(Es);
  P->edge_vs0.assign(E, 0);
  std::vector<int32_t> edge_var(E), edge_fac(E);
  for (int32_t e = 0; e < E; ++e) {
    const int32_t a = P->edge_es0[e], b = P->edge_es0[e + 1];
    const int32_t vs = vsf[a];
    PGX_REQUIRE(vs >= 0 && vs < Vs, PGX_EINVAL, "variable state out of range at edge " << e);
    const int32_t v = e2v[vs];
    PGX_REQUIRE(vs == P->var_vs0[v] && (b - a) == P->var_vs0[v + 1] - P->var_vs0[v],
                PGX_EUNSUP, "edge " << e << " does not cover all states of its variable");
    for (int32_t k = a; k < b; ++k) {
      PGX_REQUIRE(vsf[k] == vs + (k - a), PGX_EUNSUP, "edge " << e << " states out of order");
      PGX_REQUIRE(fi[k] == fi[a], PGX_EUNSUP, "edge " << e << " spans two factors");
    }
    edge_var[e] = v;
    edge_fac[e] = fi[a];
    P->edge_vs0[e] = vs;
  }

  // ---- factors -----------------------------------------------------------
  int32_t F = 0;
  if (E > 0) {
    PGX_REQUIRE(edge_fac[0] == 0, PGX_EUNSUP, "factor indices must start at 0");
    for (int32_t e = 1; e < E; ++e)
      PGX_REQUIRE(edge_fac[e] == edge_fac[e - 1] || edge_fac[e] == edge_fac[e - 1] + 1,
                  PGX_EUNSUP, "edges of a factor must be consecutive (edge " << e << ")");
    F = edge_fac[E - 1] + 1;
  }
  P->num_factors = F;
  P->fac_e0.assign(F + 1, 0);
  for (int32_t e = E - 1; e >= 0; --e) P->fac_e0[edge_fac[e]] = e;
  P->fac_e0[F] = E;

  // factor ranges of each type (types are laid out in registry order)
  int32_t tf0[4], tf1[4];
  {
    int64_t expect = 0;
    int32_t f = 0;
    for (int t = 0; t < 4; ++t) {
      const int64_t m0 = w.msgs_range[t][0], m1 = w.msgs_range[t][1];
      PGX_REQUIRE(m0 == expect && m1 >= m0, PGX_EINVAL, "message ranges must tile [0, E_s)");
      tf0[t] = f;
      while (f < F && P->edge_es0[P->fac_e0[f]] < m1) ++f;
      tf1[t] = f;
      PGX_REQUIRE((f < F && P->edge_es0[P->fac_e0[f]] == m1) || (f == F && m1 == Es),
                  PGX_EINVAL, "message range of type " << t << " does not end on a factor boundary");
      expect = m1;
    }
    PGX_REQUIRE(expect == Es && f == F, PGX_EINVAL, "message ranges do not cover all factors");
  }

  // ---- Enum factors ------------------------------------------------------
  const int32_t nfe = tf1[PGX_ENUM] - tf0[PGX_ENUM];
  std::vector<EnumFactorInfo> einfo(nfe);
  if (w.enum_num_rows < 0) {
    PGX_REQUIRE(w.enum_num_rows == PGX_ENUM_FULL_TABLES, PGX_EINVAL,
                "enum_num_rows must be >= 0 or PGX_ENUM_FULL_TABLES");
    int64_t next = 0;
    for (int32_t fl = 0; fl < nfe; ++fl) {
      const int32_t f = tf0[PGX_ENUM] + fl;
      int64_t nc = 1;
      for (int32_t e = P->fac_e0[f]; e < P->fac_e0[f + 1]; ++e) {
        nc *= P->edge_es0[e + 1] - P->edge_es0[e];
        PGX_REQUIRE(nc < (int64_t(1) << 31), PGX_EUNSUP, "Enum factor " << fl
                    << " has more than 2^31 configurations");
      }
      einfo[fl].c0 = int32_t(next);
      einfo[fl].nc = int32_t(nc);
      einfo[fl].full = true;
      next += nc;
      PGX_REQUIRE(next < (int64_t(1) << 31), PGX_EUNSUP, "too many Enum configurations");
    }
    const int64_t p0 = w.potentials_range[PGX_ENUM][0], p1 = w.potentials_range[PGX_ENUM][1];
    PGX_REQUIRE(p1 - p0 == next, PGX_EINVAL, "Enum log-potential range has "
                << (p1 - p0) << " entries for " << next << " full-table configs");
    for (auto& info : einfo) info.c0 += int32_t(p0);
  } else {
    const int64_t R = w.enum_num_rows;
    const int32_t* ci = w.enum_factor_configs_indices;
    const int32_t* ce = w.enum_factor_configs_edge_states;
    PGX_REQUIRE(R == 0 || (ci && ce), PGX_EINVAL, "Enum row arrays are null");
    const int64_t m0 = w.msgs_range[PGX_ENUM][0];
    int64_t r = 0;
    int32_t next_cfg = 0;
    int32_t cur = 0;  // enum-local factor index
    while (r < R) {
      PGX_REQUIRE(ci[r] == next_cfg, PGX_EUNSUP, "Enum configs must be numbered in row order");
      const int64_t es_first = ce[r] + m0;
      PGX_REQUIRE(es_first >= m0 && es_first < w.msgs_range[PGX_ENUM][1], PGX_EINVAL,
                  "Enum edge state out of range at row " << r);
      const int32_t e_first = int32_t(std::upper_bound(P->edge_es0.begin(), P->edge_es0.end(),
                                                       int32_t(es_first)) - P->edge_es0.begin()) - 1;
      const int32_t f = edge_fac[e_first];
      const int32_t fl = f - tf0[PGX_ENUM];
      PGX_REQUIRE(fl >= cur && fl < nfe, PGX_EUNSUP, "Enum configs must be grouped by factor");
      if (fl > cur) cur = fl;
      EnumFactorInfo& info = einfo[fl];
      if (info.nc == 0) info.c0 = next_cfg;
      const int32_t e0 = P->fac_e0[f], nv = P->fac_e0[f + 1] - e0;
      PGX_REQUIRE(r + nv <= R, PGX_EUNSUP, "truncated Enum config at row " << r);
      for (int32_t j = 0; j < nv; ++j) {
        PGX_REQUIRE(ci[r + j] == next_cfg, PGX_EUNSUP, "Enum config " << next_cfg
                    << " must have one row per variable of its factor");
        const int32_t es = int32_t(ce[r + j] + m0);
        const int32_t a = P->edge_es0[e0 + j], b = P->edge_es0[e0 + j + 1];
        PGX_REQUIRE(es >= a && es < b, PGX_EUNSUP, "Enum row " << r + j
                    << " is not in the factor's edge " << j);
        info.table.push_back(es - P->edge_es0[e0]);
      }
      info.nc += 1;
      r += nv;
      ++next_cfg;
    }
    const int64_t p0 = w.potentials_range[PGX_ENUM][0], p1 = w.potentials_range[PGX_ENUM][1];
    PGX_REQUIRE(p1 - p0 == next_cfg, PGX_EINVAL, "Enum log-potential range has "
                << (p1 - p0) << " entries for " << next_cfg << " configs");
    for (auto& info : einfo) info.c0 += int32_t(p0);
  }

  // ---- logical / pool factors -------------------------------------------
  for (int t = PGX_OR; t <= PGX_POOL; ++t) {
    const pgx_logical_wiring& lw = w.logical[t];
    const int32_t nf = tf1[t] - tf0[t];
    PGX_REQUIRE(lw.num_factors == nf, PGX_EINVAL, "type " << t << " wiring has "
                << lw.num_factors << " factors, the message range holds " << nf);
    if (nf == 0) continue;
    const int32_t off = lw.edge_states_offset;
    PGX_REQUIRE(off == 1 || (off == -1 && t != PGX_POOL), PGX_EINVAL,
                "edge_states_offset must be 1 (OR, Pool) or -1 (AND)");
    const int32_t rel = (t == PGX_POOL) ? 0 : (1 - off) / 2;
    const int64_t m0 = w.msgs_range[t][0];
    int64_t p = 0;
    for (int32_t fl = 0; fl < nf; ++fl) {
      const int32_t f = tf0[t] + fl;
      const int32_t e0 = P->fac_e0[f], ne = P->fac_e0[f + 1] - e0;
      PGX_REQUIRE(ne >= 2, PGX_EUNSUP, "logical factor " << fl << " needs a parent and a child");
      for (int32_t k = 0; k < ne; ++k)
        PGX_REQUIRE(P->edge_es0[e0 + k + 1] - P->edge_es0[e0 + k] == 2, PGX_EUNSUP,
                    "logical factor variables must be binary");
      for (int32_t k = 0; k < ne - 1; ++k, ++p) {
        PGX_REQUIRE(p < lw.num_parents && lw.parents_factor_indices[p] == fl, PGX_EUNSUP,
                    "parents of factor " << fl << " must be consecutive, in edge order");
        PGX_REQUIRE(lw.parents_msg_indices[p] + m0 == P->edge_es0[e0 + k] + rel, PGX_EUNSUP,
                    "parent message index of factor " << fl << " does not match its edge " << k);
      }
      PGX_REQUIRE(lw.children_edge_states[fl] + m0 == P->edge_es0[e0 + ne - 1] + rel, PGX_EUNSUP,
                  "child message index of factor " << fl << " does not match its last edge");
    }
    PGX_REQUIRE(p == lw.num_parents, PGX_EINVAL, "type " << t << " has extra parents");
  }

  // ---- components -> tasks ----------------------------------------------
  UnionFind uf(NV);
  for (int32_t f = 0; f < F; ++f)
    for (int32_t e = P->fac_e0[f] + 1; e < P->fac_e0[f + 1]; ++e)
      uf.unite(edge_var[P->fac_e0[f]], edge_var[e]);
  std::vector<int32_t> comp(NV);
  for (int32_t v = 0; v < NV; ++v) comp[v] = uf.find(v);
  // runs of components in variable order
  std::vector<int32_t> run_of_comp(NV, -1);
  std::vector<std::pair<int32_t, int32_t>> var_runs;
  bool contiguous = true;
  for (int32_t v = 0; v < NV; ++v) {
    if (v == 0 || comp[v] != comp[v - 1]) {
      if (run_of_comp[comp[v]] != -1) { contiguous = false; break; }
      run_of_comp[comp[v]] = int32_t(var_runs.size());
      var_runs.push_back({v, v + 1});
    } else {
      var_runs.back().second = v + 1;
    }
  }
  const int32_t nruns = contiguous ? int32_t(var_runs.size()) : 1;
  // factor range of each run for each type
  std::vector<std::array<std::pair<int32_t, int32_t>, 4>> frange(nruns);
  for (auto& fr : frange)
    for (auto& x : fr) x = {0, 0};
  if (contiguous) {
    for (int t = 0; t < 4 && contiguous; ++t) {
      int32_t prev_run = -1;
      std::vector<char> seen(nruns, 0);
      for (int32_t f = tf0[t]; f < tf1[t]; ++f) {
        const int32_t run = run_of_comp[comp[edge_var[P->fac_e0[f]]]];
        if (run != prev_run) {
          if (seen[run]) { contiguous = false; break; }
          seen[run] = 1;
          frange[run][t] = {f, f + 1};
          prev_run = run;
        } else {
          frange[run][t].second = f + 1;
        }
      }
    }
  }
  std::vector<Task> tasks;
  std::vector<std::array<std::pair<int32_t, int32_t>, 4>> task_frange;
  if (!contiguous || NV == 0) {
    Task t{0, NV, 0, 0, 0, 0, 0};
    tasks.push_back(t);
    std::array<std::pair<int32_t, int32_t>, 4> fr;
    for (int k = 0; k < 4; ++k) fr[k] = {tf0[k], tf1[k]};
    task_frange.push_back(fr);
  } else {
    // merge neighbouring runs while small, when all their ranges are adjacent
    const int64_t kMergeBelow = merge_below;
    for (int32_t r = 0; r < nruns; ++r) {
      int64_t work = var_runs[r].second - var_runs[r].first;
      for (int t = 0; t < 4; ++t)
        work += P->edge_es0[P->fac_e0[frange[r][t].second]] - P->edge_es0[P->fac_e0[frange[r][t].first]];
      bool merged = false;
      if (!tasks.empty() && kMergeBelow > 0 && tasks.back().work + work <= kMergeBelow &&
          tasks.back().v1 == var_runs[r].first) {
        bool adj = true;
        for (int t = 0; t < 4; ++t) {
          const auto& a = task_frange.back()[t];
          const auto& b = frange[r][t];
          if (a.first == a.second || b.first == b.second) continue;
          if (a.second != b.first) adj = false;
        }
        if (adj) {
          tasks.back().v1 = var_runs[r].second;
          tasks.back().work += work;
          for (int t = 0; t < 4; ++t) {
            auto& a = task_frange.back()[t];
            const auto& b = frange[r][t];
            if (b.first == b.second) continue;
            if (a.first == a.second) a = b; else a.second = b.second;
          }
          merged = true;
        }
      }
      if (!merged) {
        tasks.push_back(Task{var_runs[r].first, var_runs[r].second, 0, 0, 0, 0, work});
        task_frange.push_back(frange[r]);
      }
    }
  }

  // ---- blocks --------------------------------------------------------------
  std::map<std::vector<int32_t>, int32_t> table_ids;
  for (size_t ti = 0; ti < tasks.size(); ++ti) {
    Task& task = tasks[ti];
    task.b0 = int32_t(P->blocks.size());
    int64_t work = task.v1 - task.v0;
    for (int t = 0; t < 4; ++t) {
      const int32_t fa = task_frange[ti][t].first, fb = task_frange[ti][t].second;
      if (fa == fb) continue;
      work += P->edge_es0[P->fac_e0[fb]] - P->edge_es0[P->fac_e0[fa]];
      if (t != PGX_ENUM) {
        Block b{};
        b.kind = t == PGX_OR ? kOr : (t == PGX_AND ? kAnd : kPool);
        b.task = int32_t(ti);
        b.f0 = fa;
        b.count = fb - fa;
        b.e0 = P->fac_e0[fa];
        b.es0 = P->edge_es0[b.e0];
        b.off = w.logical[t].edge_states_offset;
        P->blocks.push_back(b);
        continue;
      }
      for (int32_t f = fa; f < fb; ++f) {
        const EnumFactorInfo& info = einfo[f - tf0[PGX_ENUM]];
        const int32_t e0 = P->fac_e0[f], nv = P->fac_e0[f + 1] - e0;
        const int32_t es0 = P->edge_es0[e0], esf = P->edge_es0[e0 + nv] - es0;
        Block b{};
        b.task = int32_t(ti);
        b.f0 = f;
        b.count = 1;
        b.e0 = e0;
        b.es0 = es0;
        b.nv = nv;
        b.esf = esf;
        b.c0 = info.c0;
        b.nc = info.nc;
        b.kind = kEnumGeneral;
        if (nv == 2) {
          const int32_t s0 = P->edge_es0[e0 + 1] - es0, s1 = esf - s0;
          bool full = info.nc == s0 * s1 && s0 <= 8 && s1 <= 8;
          for (int32_t c = 0; full && !info.full && c < info.nc; ++c)
            full = info.table[2 * c] == c / s1 && info.table[2 * c + 1] == s0 + c % s1;
          if (full) {
            b.kind = (s0 == 2 && s1 == 2) ? kEnumBin2 : kEnumPair;
            b.s0 = s0;
            b.s1 = s1;
          }
        }
        if (b.kind == kEnumGeneral) {
          // table layout: [nv edge offsets][esf owning edge][nc*nv row offsets]
          std::vector<int32_t> key;
          key.reserve(2 + nv + esf + info.table.size());
          key.push_back(nv);
          key.push_back(info.nc);
          for (int32_t j = 0; j < nv; ++j) key.push_back(P->edge_es0[e0 + j] - es0);
          for (int32_t j = 0; j < nv; ++j)
            for (int32_t k = P->edge_es0[e0 + j]; k < P->edge_es0[e0 + j + 1]; ++k) key.push_back(j);
          if (info.full) {
            std::vector<int32_t> off(nv), ns(nv);
            for (int32_t j = 0; j < nv; ++j) {
              off[j] = P->edge_es0[e0 + j] - es0;
              ns[j] = P->edge_es0[e0 + j + 1] - P->edge_es0[e0 + j];
            }
            const std::vector<int32_t> t = full_table(off, ns, info.nc);
            key.insert(key.end(), t.begin(), t.end());
          } else {
            key.insert(key.end(), info.table.begin(), info.table.end());
          }
          auto it = table_ids.find(key);
          if (it == table_ids.end()) {
            const int32_t id = int32_t(P->tables.size());
            P->tables.insert(P->tables.end(), key.begin() + 2, key.end());
            // then, per edge state k, the configs whose row for k's edge is k,
            // in increasing order: [esf + 1 offsets][config ids]
            const int32_t* owner = key.data() + 2 + nv;
            const int32_t* rows = owner + esf;
            std::vector<std::vector<int32_t>> lists(esf);
            for (int32_t c = 0; c < info.nc; ++c)
              for (int32_t j = 0; j < nv; ++j) {
                const int32_t k = rows[c * nv + j];
                if (k >= 0 && k < esf && owner[k] == j) lists[k].push_back(c);
              }
            int32_t acc = 0;
            for (int32_t k = 0; k <= esf; ++k) {
              P->tables.push_back(acc);
              if (k < esf) acc += int32_t(lists[k].size());
            }
            for (const auto& l : lists) P->tables.insert(P->tables.end(), l.begin(), l.end());
            PGX_REQUIRE(P->tables.size() < (size_t(1) << 31), PGX_EUNSUP, "Enum tables too large");
            it = table_ids.emplace(std::move(key), id).first;
          }
          b.tbl = it->second;
          P->max_esf = std::max(P->max_esf, esf);
        }
        if (!P->blocks.empty()) {
          Block& prev = P->blocks.back();
          if (prev.task == b.task && prev.kind == b.kind && prev.nv == b.nv &&
              prev.esf == b.esf && prev.nc == b.nc && prev.tbl == b.tbl &&
              prev.s0 == b.s0 && prev.s1 == b.s1 && int32_t(P->blocks.size()) > task.b0 &&
              prev.e0 + prev.count * prev.nv == b.e0 &&
              prev.es0 + prev.count * prev.esf == b.es0 &&
              prev.c0 + prev.count * prev.nc == b.c0) {
            prev.count += 1;
            continue;
          }
        }
        P->blocks.push_back(b);
      }
    }
    task.b1 = int32_t(P->blocks.size());
    task.work = work;
  }
  // ---- variable-side CSR ---------------------------------------------------
  P->var_eptr.assign(NV + 1, 0);
  for (int32_t e = 0; e < E; ++e) P->var_eptr[edge_var[e] + 1] += 1;
  for (int32_t v = 0; v < NV; ++v) P->var_eptr[v + 1] += P->var_eptr[v];
  P->var_adj_es.assign(E, 0);
  std::vector<int32_t> fill(P->var_eptr.begin(), P->var_eptr.end() - 1);
  for (int32_t e = 0; e < E; ++e) P->var_adj_es[fill[edge_var[e]]++] = P->edge_es0[e];

  P->var_adj_es_ref = P->var_adj_es;
  if (planar) {
    bool any = false;
    P->perm.resize(size_t(Es));
    std::iota(P->perm.begin(), P->perm.end(), 0);
    for (Block& b : P->blocks) {
      if (b.kind != kEnumBin2 || b.count < 2) continue;
      b.pad[0] |= 2;
      any = true;
      for (int32_t i = 0; i < b.count; ++i)
        for (int32_t s = 0; s < 2; ++s) {
          P->perm[b.es0 + 4 * i + s] = b.es0 + 2 * i + s;
          P->perm[b.es0 + 4 * i + 2 + s] = b.es0 + 2 * b.count + 2 * i + s;
        }
    }
    if (any) {
      for (auto& es : P->var_adj_es) es = P->perm[es];
    } else {
      P->perm.clear();
    }
  }

  P->edge_var = edge_var;
  for (int32_t v = 0; v < NV; ++v)
    P->max_degree = std::max(P->max_degree, P->var_eptr[v + 1] - P->var_eptr[v]);
  P->pairwise_only = !P->blocks.empty();
  for (Block& b : P->blocks) {
    if (b.kind != kEnumBin2 && b.kind != kEnumPair) P->pairwise_only = false;
    if (b.kind == kEnumBin2 && b.c0 % 4 == 0) b.pad[0] |= 1;
    // general Enum factors with a wavefront's worth of rows and tables that
    // fit the per-wavefront LDS run one wavefront per factor
    if (b.kind == kEnumGeneral && int64_t(b.nc) * b.nv >= 64 && b.nc <= kWaveMaxConfigs &&
        b.esf <= kWaveMaxStates)
      b.pad[0] |= 8;
    if (b.kind == kOr || b.kind == kAnd) {
      // factors of more than one run of parents go to lane groups (all of
      // the block's factors have at most kWideRuns runs)
      int32_t maxp = 0;
      for (int32_t i = 0; i < b.count; ++i)
        maxp = std::max(maxp, P->fac_e0[b.f0 + i + 1] - P->fac_e0[b.f0 + i] - 1);
      // (one wavefront per factor when all have at most 256 parents)
      if (maxp > kVarChunk && maxp <= kWaveLogicalParents) b.pad[0] |= 16;
      else if (maxp > kVarChunk && maxp <= kWideRuns * kVarChunk) b.pad[0] |= 4;
    }
  }
  for (int32_t e = 0; e < E && P->even2; ++e)
    if (P->edge_es0[e + 1] - P->edge_es0[e] == 2 && (P->edge_es0[e] & 1)) P->even2 = false;
  for (int32_t v = 0; v < NV && P->even2; ++v)
    if (P->var_vs0[v + 1] - P->var_vs0[v] == 2 && (P->var_vs0[v] & 1)) P->even2 = false;

  // position descriptors of low-degree variables
  if (int64_t(NV) * 16 <= kPos4MaxBytes) {
    P->var_pos4.assign(size_t(NV) * 4, -1);
    for (int32_t v = 0; v < NV; ++v) {
      const int32_t p0 = P->var_eptr[v], deg = P->var_eptr[v + 1] - p0;
      if (deg > 4) {
        P->var_pos4[size_t(v) * 4] = -2;
        continue;
      }
      for (int32_t k = 0; k < deg; ++k) P->var_pos4[size_t(v) * 4 + k] = P->var_adj_es[p0 + k];
    }
    if (P->pairwise_only && int64_t(E) * 16 <= kEdgePos4MaxBytes) {
      P->edge_pos4.resize(size_t(E) * 4);
      for (int32_t e = 0; e < E; ++e)
        for (int k = 0; k < 4; ++k)
          P->edge_pos4[size_t(e) * 4 + k] = P->var_pos4[size_t(edge_var[e]) * 4 + k];
    }
  }

  P->packable = !P->var_pos4.empty() && !P->perm.empty() && P->even2 && P->max_degree <= 4 &&
                !P->blocks.empty();
  for (const Block& b : P->blocks)
    if (b.kind != kEnumBin2 || !(b.pad[0] & 2)) P->packable = false;
  for (int32_t v = 0; v < NV && P->packable; ++v)
    if (P->var_vs0[v + 1] - P->var_vs0[v] != 2) P->packable = false;

  // heavy variables, grouped by task (tasks own contiguous variable ranges)
  for (Task& t : tasks) {
    t.h0 = int32_t(P->heavy_vars.size());
    for (int32_t v = t.v0; v < t.v1; ++v)
      if (P->var_eptr[v + 1] - P->var_eptr[v] > kVarChunk) P->heavy_vars.push_back(v);
    t.h1 = int32_t(P->heavy_vars.size());
  }
  P->heavy_cptr.assign(1, 0);
  P->heavy_pofs.assign(1, 0);
  for (int32_t h = 0; h < int32_t(P->heavy_vars.size()); ++h) {
    const int32_t v = P->heavy_vars[h];
    const int32_t nch = (P->var_eptr[v + 1] - P->var_eptr[v] + kVarChunk - 1) / kVarChunk;
    const int32_t ns = P->var_vs0[v + 1] - P->var_vs0[v];
    P->heavy_cptr.push_back(P->heavy_cptr.back() + nch);
    P->heavy_pofs.push_back(P->heavy_pofs.back() + nch * ns);
  }
  // group sums of variables with more than kVarChunk chunks
  int64_t gslot = P->heavy_pofs.back();
  P->heavy_gpofs.assign(P->heavy_vars.size(), -1);
  for (int32_t h = 0; h < int32_t(P->heavy_vars.size()); ++h) {
    const int32_t v = P->heavy_vars[h];
    const int32_t nch = P->heavy_cptr[h + 1] - P->heavy_cptr[h];
    const int32_t ng = (nch + kVarChunk - 1) / kVarChunk;
    if (ng <= 1) continue;
    const int32_t ns = P->var_vs0[v + 1] - P->var_vs0[v];
    P->heavy_gpofs[h] = int32_t(gslot);
    for (int32_t j = 0; j < ng; ++j) {
      P->vh_group_h.push_back(h);
      P->vh_group_j.push_back(j);
    }
    gslot += int64_t(ng) * ns;
  }
  PGX_REQUIRE(gslot < (int64_t(1) << 31), PGX_EUNSUP, "too many partial sums");
  P->partial_size = gslot;

  // Chunk order inside each task: chunk number first, variable second, so
  // neighbouring threads sum the same chunk of neighbouring variables (whose
  // edges often sit next to each other, e.g. the factors (i, j), (i, j+1)).
  for (const Task& t : tasks) {
    int32_t kmax = 0;
    for (int32_t h = t.h0; h < t.h1; ++h)
      kmax = std::max(kmax, P->heavy_cptr[h + 1] - P->heavy_cptr[h]);
    for (int32_t k = 0; k < kmax; ++k)
      for (int32_t h = t.h0; h < t.h1; ++h)
        if (k < P->heavy_cptr[h + 1] - P->heavy_cptr[h]) {
          P->heavy_chunk_var.push_back(h);
          P->heavy_chunk_k.push_back(k);
        }
  }

  for (int32_t hc = 0; hc < int32_t(P->heavy_chunk_var.size()); ++hc) {
    const int32_t h = P->heavy_chunk_var[hc], k = P->heavy_chunk_k[hc];
    const int32_t v = P->heavy_vars[h];
    const int32_t a = P->var_eptr[v] + k * kVarChunk;
    const int32_t z = std::min(a + kVarChunk, P->var_eptr[v + 1]);
    bool row = P->var_vs0[v + 1] - P->var_vs0[v] == 2;
    for (int32_t p = a; row && p + 1 < z; ++p) row = P->var_adj_es[p + 1] == P->var_adj_es[p] + 2;
    (row ? P->heavy_row_chunks : P->heavy_col_chunks).push_back(hc);
    int32_t stride = 0;
    if (z - a == kVarChunk) {
      stride = P->var_adj_es[a + 1] - P->var_adj_es[a];
      for (int32_t p = a; stride > 0 && p + 1 < z; ++p)
        if (P->var_adj_es[p + 1] - P->var_adj_es[p] != stride) stride = 0;
    }
    P->heavy_chunk_stride.push_back(std::max(stride, 0));
  }

  {
    const int32_t nhc = int32_t(P->heavy_chunk_var.size());
    auto desc_of = [&](int32_t hc, bool row, bool ref) {
      const int32_t h = P->heavy_chunk_var[hc], k = P->heavy_chunk_k[hc];
      const int32_t v = P->heavy_vars[h];
      const int32_t ns = P->var_vs0[v + 1] - P->var_vs0[v];
      const int32_t a = P->var_eptr[v] + k * kVarChunk;
      const int32_t len = std::min(a + kVarChunk, P->var_eptr[v + 1]) - a;
      return std::array<int32_t, 4>{row ? P->var_adj_es[a] : a, len | (ns << 8),
                                    P->heavy_pofs[h] + k * ns,
                                    (row || ref) ? 0 : P->heavy_chunk_stride[hc]};
    };
    for (int32_t h = 0; h < int32_t(P->heavy_vars.size()); ++h) {
      const int32_t v = P->heavy_vars[h];
      PGX_REQUIRE(P->var_vs0[v + 1] - P->var_vs0[v] < (int32_t(1) << 15), PGX_EUNSUP,
                  "variable " << v << " has too many states for chunk descriptors");
    }
    for (int32_t hc = 0; hc < nhc; ++hc) {
      for (int32_t x : desc_of(hc, false, false)) P->chunk_desc.push_back(x);
      for (int32_t x : desc_of(hc, false, true)) P->chunk_desc_ref.push_back(x);
    }
    for (int32_t hc : P->heavy_col_chunks)
      for (int32_t x : desc_of(hc, false, false)) P->col_desc.push_back(x);
    for (int32_t hc : P->heavy_row_chunks)
      for (int32_t x : desc_of(hc, true, false)) P->row_desc.push_back(x);
    for (int32_t h = 0; h < int32_t(P->heavy_vars.size()); ++h) {
      const int32_t v = P->heavy_vars[h];
      const int32_t ns = P->var_vs0[v + 1] - P->var_vs0[v];
      const int32_t nch = P->heavy_cptr[h + 1] - P->heavy_cptr[h];
      PGX_REQUIRE(nch < (1 << 16), PGX_EUNSUP,
                  "variable " << v << " has more than " << (kVarChunk << 16) << " edges");
      P->heavy_desc.insert(P->heavy_desc.end(),
                           {P->var_vs0[v], nch | (ns << 16), P->heavy_pofs[h], P->heavy_gpofs[h]});
    }
  }

  // complete bipartite binary graph? (optionally with one unary factor per
  // unit first: Plan::bip_unary)
  auto unary_block = [&](const Block& u) {
    // one variable, 2 states, configs [0] and [1]: the general table holds
    // [edge offset 0][owning edge 0, 0][rows 0, 1]
    return u.kind == kEnumGeneral && u.nv == 1 && u.esf == 2 && u.nc == 2 && u.es0 == 0 &&
           u.e0 == 0 && P->tables[u.tbl] == 0 && P->tables[u.tbl + 1] == 0 &&
           P->tables[u.tbl + 2] == 0 && P->tables[u.tbl + 3] == 0 && P->tables[u.tbl + 4] == 1;
  };
  const int32_t pb = !P->even2 ? -1
                     : P->blocks.size() == 1 ? 0
                     : (P->blocks.size() == 2 && unary_block(P->blocks[0])) ? 1
                                                                             : -1;
  if (pb >= 0) {
    const Block& b = P->blocks[pb];
    const int32_t n = b.count;
    const int32_t nu = pb == 1 ? P->blocks[0].count : 0;  // unary factors
    int32_t nb = 0;
    // (the pairwise block holds every other edge state: first-edge pairs at
    // [2 nu, 2 nu + 2n), second-edge pairs at [2 nu + 2n, 2 nu + 4n), which
    // the one-float form relies on)
    if (b.kind == kEnumBin2 && (b.pad[0] & 2) && n + nu == P->num_factors && b.es0 == 2 * nu &&
        int64_t(4) * n + 2 * nu == P->num_es) {
      const int32_t r0 = edge_var[b.e0];
      while (nb < n && edge_var[b.e0 + 2 * nb] == r0) ++nb;
    }
    const int32_t na = nb > 0 ? n / nb : 0;
    const int32_t ud = nu > 0 ? 1 : 0;  // a unit's unary edge
    bool ok = nb > kVarChunk && na > kVarChunk && int64_t(na) * nb == n && na + nb == NV &&
              (nu == 0 || nu == na + nb);
    for (int32_t i = 0; ok && i < n; ++i)
      ok = edge_var[b.e0 + 2 * i] == edge_var[b.e0 + 2 * (i - i % nb)] &&
           edge_var[b.e0 + 2 * i + 1] == edge_var[b.e0 + 2 * (i % nb) + 1];
    std::vector<int32_t> heavy_of(NV, -1);
    for (int32_t h = 0; h < int32_t(P->heavy_vars.size()); ++h) heavy_of[P->heavy_vars[h]] = h;
    auto info = [&](int32_t v, std::vector<int32_t>* dst) {
      dst->push_back(P->var_vs0[v]);
      dst->push_back(P->heavy_pofs[heavy_of[v]]);
    };
    // with every factor in the block, degrees nb (rows) and na (columns) and
    // na + nb variables, all row and column variables are distinct
    // (with unary factors: unit a's is factor a, column unit c's factor
    // na + c, and, coming first in edge order, each unit's first edge)
    for (int32_t a = 0; ok && a < na; ++a) {
      const int32_t v = edge_var[b.e0 + 2 * a * nb];
      ok = P->var_eptr[v + 1] - P->var_eptr[v] == nb + ud &&
           P->var_vs0[v + 1] - P->var_vs0[v] == 2 && (nu == 0 || edge_var[a] == v);
      if (ok) info(v, &P->bip_row);
    }
    for (int32_t c = 0; ok && c < nb; ++c) {
      const int32_t v = edge_var[b.e0 + 2 * c + 1];
      ok = P->var_eptr[v + 1] - P->var_eptr[v] == na + ud &&
           P->var_vs0[v + 1] - P->var_vs0[v] == 2 && (nu == 0 || edge_var[na + c] == v);
      if (ok) info(v, &P->bip_col);
    }
    if (ok) {
      P->bip_na = na;
      P->bip_nb = nb;
      P->bip_unary = ud;
      P->bip_pb = pb;
      P->bip_uc0 = nu > 0 ? P->blocks[0].c0 : 0;
      // affine tables? (true when the row and column variables are numbered
      // consecutively, as variable groups are)
      auto affine = [](const std::vector<int32_t>& t, int32_t n, int32_t* out) {
        for (int c = 0; c < 2; ++c) {
          out[2 * c] = t[c];
          out[2 * c + 1] = n > 1 ? t[2 + c] - t[c] : 0;
          for (int32_t i = 0; i < n; ++i)
            if (int64_t(t[2 * i + c]) != int64_t(out[2 * c]) + int64_t(i) * out[2 * c + 1])
              return false;
        }
        return true;
      };
      int32_t aff[8];
      if (affine(P->bip_row, na, aff) && affine(P->bip_col, nb, aff + 4)) {
        P->bip_affine = true;
        std::copy(aff, aff + 8, P->bip_aff);
      }
    } else {
      P->bip_row.clear();
      P->bip_col.clear();
    }
  }

  // drop tasks without factors, order by descending work
  std::vector<Task> kept;
  for (const Task& t : tasks)
    if (t.b1 > t.b0) kept.push_back(t);
  std::stable_sort(kept.begin(), kept.end(),
                   [](const Task& a, const Task& b) { return a.work > b.work; });
  P->tasks = kept;
  return PGX_OK;
}

}  // namespace pgx
