"""Batches of uneven factor graphs in one launch.

A batch of independent graphs of different sizes, state counts and factor
types is compiled into ONE disjoint-union wiring with a graph-offset table:
the union has the layout a FactorGraph gets when the graphs' variable groups
and factor groups are added one graph after the other (variables graph by
graph; messages and log-potentials factor type by factor type, graph by
graph inside a type, as pgmax/fgraph/fgraph.py:185-232 orders them). The
engine finds the graphs again as connected components and gives each its
own workgroup, so a batch runs in one launch with per-graph deltas and
per-graph early exit (``BatchedBP.run_converge``).

``grid_wiring`` builds the wiring of a 4-neighbour pairwise grid directly
with NumPy (no per-variable Python objects); tests check it is identical to
what FactorGraph + PairwiseFactorGroup compile.
"""

import ctypes
import dataclasses
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from pgmax_amd import _arrays
from pgmax_amd import _lib
from pgmax_amd import engine
from pgmax_amd.factor import FAC_TO_VAR_UPDATES

ENUM, OR, AND, POOL = 0, 1, 2, 3


@dataclasses.dataclass
class FlatWiring:
  """One factor graph's compiled wiring in the reference's flat layout."""

  num_var_states: int
  num_log_potentials: int
  var_states_for_edge_states: np.ndarray
  edge_indices_for_edge_states: np.ndarray
  factor_indices_for_edge_states: np.ndarray
  evidence_to_vars: np.ndarray
  msgs_range: List[Tuple[int, int]]
  potentials_range: List[Tuple[int, int]]
  enum_ci: Optional[np.ndarray]  # None: every Enum factor is a full table
  enum_ce: Optional[np.ndarray]
  logical: Dict[int, Tuple[np.ndarray, np.ndarray, np.ndarray, int]]

  @property
  def full_tables(self) -> bool:
    return self.enum_ci is None

  def enum_rows(self) -> Tuple[np.ndarray, np.ndarray]:
    """Explicit Enum rows; full tables are listed in row-major config order."""
    if not self.full_tables:
      return self.enum_ci, self.enum_ce
    m0, m1 = self.msgs_range[ENUM]
    ei = np.asarray(self.edge_indices_for_edge_states[m0:m1])
    fi = np.asarray(self.factor_indices_for_edge_states[m0:m1])
    if ei.size == 0:
      return np.zeros(0, np.int64), np.zeros(0, np.int64)
    starts = np.flatnonzero(np.r_[True, ei[1:] != ei[:-1]])  # first state of each edge
    sizes = np.diff(np.r_[starts, ei.size])  # states per edge
    efac = fi[starts]
    fstart = np.flatnonzero(np.r_[True, efac[1:] != efac[:-1]])  # first edge of each factor
    arity = np.diff(np.r_[fstart, efac.size])
    # configs and rows per factor, numbered in factor order
    log_nc = np.add.reduceat(np.log2(sizes.astype(np.float64)), fstart)
    nc = np.rint(np.exp2(log_nc)).astype(np.int64)
    cfg0 = np.r_[0, np.cumsum(nc)[:-1]]
    row0 = np.r_[0, np.cumsum(nc * arity)[:-1]]
    ci = np.empty(int((nc * arity).sum()), np.int64)
    ce = np.empty_like(ci)
    # factors with the same state counts share one row-major config grid
    keys = {}
    for f_idx, (a, e0) in enumerate(zip(arity.tolist(), fstart.tolist())):
      keys.setdefault(tuple(sizes[e0:e0 + a].tolist()), []).append(f_idx)
    for shape, members in keys.items():
      members = np.asarray(members, np.int64)
      a = len(shape)
      grid = np.indices(shape).reshape(a, -1).T  # (configs, a), row-major
      n = grid.shape[0]
      edge_first = starts[fstart[members][:, None] + np.arange(a)]  # (factors, a)
      rows = row0[members][:, None] + np.arange(n * a)  # (factors, n * a)
      ce[rows] = (edge_first[:, None, :] + grid[None]).reshape(len(members), -1)
      ci[rows] = (cfg0[members][:, None, None] + np.arange(n)[None, :, None]
                  + np.zeros((1, 1, a), np.int64)).reshape(len(members), -1)
    return ci, ce

  @property
  def num_edge_states(self) -> int:
    return int(self.var_states_for_edge_states.shape[0])

  @property
  def num_vars(self) -> int:
    return int(self.evidence_to_vars[-1]) + 1 if self.evidence_to_vars.size else 0

  @staticmethod
  def from_context(ctx) -> "FlatWiring":
    """From a pgmax_amd InfererContext (pgmax/infer/inferer.py:65-118)."""
    types = list(FAC_TO_VAR_UPDATES.keys())
    args = ctx.inference_arguments
    logical = {}
    for code, t in ((OR, types[1]), (AND, types[2])):
      a = args[t]
      logical[code] = (np.asarray(a["parents_factor_indices"]),
                       np.asarray(a["parents_msg_indices"]),
                       np.asarray(a["children_edge_states"]), int(a["edge_states_offset"]))
    a = args[types[3]]
    logical[POOL] = (np.asarray(a["pool_choices_factor_indices"]),
                     np.asarray(a["pool_choices_msg_indices"]),
                     np.asarray(a["pool_indicators_edge_states"]), 1)
    fg_state = ctx.bp_state.fg_state
    return FlatWiring(
        num_var_states=int(fg_state.num_var_states),
        num_log_potentials=int(fg_state.log_potentials.shape[0]),
        var_states_for_edge_states=np.asarray(ctx.var_states_for_edge_states),
        edge_indices_for_edge_states=np.asarray(ctx.edge_indices_for_edge_states),
        factor_indices_for_edge_states=np.asarray(ctx.factor_indices_for_edge_states),
        evidence_to_vars=np.asarray(ctx.evidence_to_vars),
        msgs_range=[tuple(int(v) for v in ctx.factor_type_to_msgs_range[t]) for t in types],
        potentials_range=[tuple(int(v) for v in ctx.factor_type_to_potentials_range[t])
                          for t in types],
        enum_ci=np.asarray(args[types[0]]["factor_configs_indices"]),
        enum_ce=np.asarray(args[types[0]]["factor_configs_edge_states"]),
        logical=logical,
    )

  def to_engine_wiring(self) -> engine.Wiring:
    return engine.Wiring(
        num_var_states=self.num_var_states,
        num_log_potentials=self.num_log_potentials,
        var_states_for_edge_states=self.var_states_for_edge_states,
        edge_indices_for_edge_states=self.edge_indices_for_edge_states,
        factor_indices_for_edge_states=self.factor_indices_for_edge_states,
        evidence_to_vars=self.evidence_to_vars,
        msgs_range=self.msgs_range,
        potentials_range=self.potentials_range,
        enum_args="full" if self.full_tables else (self.enum_ci, self.enum_ce),
        logical_args=self.logical,
    )


def grid_pairs(height: int, width: int, torus: bool) -> np.ndarray:
  """Variable pairs of a 4-neighbour grid in raster order.

  For each (i, j): the pair with (i+1, j), then with (i, j+1); on a torus the
  neighbours wrap, otherwise pairs leaving the grid are skipped. This is the
  factor order of examples/ising_model.ipynb cell 12 of the reference.
  """
  i, j = np.meshgrid(np.arange(height), np.arange(width), indexing="ij")
  i, j = i.ravel(), j.ravel()
  v = i * width + j
  if torus:
    down = ((i + 1) % height) * width + j
    right = i * width + (j + 1) % width
    return np.stack([np.stack([v, down], 1), np.stack([v, right], 1)], 1).reshape(-1, 2)
  pairs = np.full((v.shape[0], 2, 2), -1, dtype=np.int64)
  ok_d = i + 1 < height
  ok_r = j + 1 < width
  pairs[:, 0, 0] = v
  pairs[:, 0, 1] = np.where(ok_d, v + width, -1)
  pairs[:, 1, 0] = v
  pairs[:, 1, 1] = np.where(ok_r, v + 1, -1)
  keep = np.stack([ok_d, ok_r], 1).ravel()
  return pairs.reshape(-1, 2)[keep]


def pairwise_wiring(num_vars: int, num_states: int, pairs: np.ndarray,
                    full_tables: bool = False) -> FlatWiring:
  """Wiring of one PairwiseFactorGroup over ``num_vars`` variables of S states.

  Same arrays as FactorGraph([NDVarArray]) + PairwiseFactorGroup(pairs)
  compiles (pgmax/factor/enum.py:191-270, config a * S + b = (a, b)).
  With ``full_tables`` the Enum rows are left implicit (the engine's
  PGX_ENUM_FULL_TABLES form), which saves 2 * S^2 ints per factor.
  """
  s = int(num_states)
  nf = pairs.shape[0]
  vs_of_edge = (pairs.reshape(-1) * s)
  es = np.arange(nf * 2 * s, dtype=np.int64)
  edge = es // s
  vsfes = vs_of_edge[edge] + es % s
  if full_tables:
    ci = ce = None
  else:
    c = np.arange(s * s)
    fc = np.arange(nf * s * s, dtype=np.int64)
    ci = np.repeat(fc, 2)
    f_of = np.repeat(np.arange(nf, dtype=np.int64) * 2 * s, s * s)
    ce = np.stack([f_of + np.tile(c // s, nf), f_of + s + np.tile(c % s, nf)], 1).reshape(-1)
  n_es = nf * 2 * s
  n_lp = nf * s * s
  return FlatWiring(
      num_var_states=num_vars * s,
      num_log_potentials=n_lp,
      var_states_for_edge_states=vsfes,
      edge_indices_for_edge_states=edge,
      factor_indices_for_edge_states=es // (2 * s),
      evidence_to_vars=np.repeat(np.arange(num_vars, dtype=np.int64), s),
      msgs_range=[(0, n_es), (n_es, n_es), (n_es, n_es), (n_es, n_es)],
      potentials_range=[(0, n_lp), (n_lp, n_lp), (n_lp, n_lp), (n_lp, n_lp)],
      enum_ci=ci,
      enum_ce=ce,
      logical={t: (np.zeros(0, np.int64), np.zeros(0, np.int64), np.zeros(0, np.int64),
                   1 if t != AND else -1) for t in (OR, AND, POOL)},
  )


def grid_wiring(height: int, width: int, num_states: int, torus: bool = False,
                full_tables: bool = False) -> FlatWiring:
  """4-neighbour pairwise grid of ``num_states``-state variables."""
  return pairwise_wiring(height * width, num_states, grid_pairs(height, width, torus),
                         full_tables=full_tables)


@dataclasses.dataclass
class UnionLayout:
  """Where each graph's flat arrays land in the union."""

  var_offsets: np.ndarray       # [G+1] variables
  vs_offsets: np.ndarray        # [G+1] variable states (evidence)
  es_index: List[np.ndarray]    # per graph: union position of each local edge state
  lp_index: List[np.ndarray]    # per graph: union position of each local log-potential
  num_es: int
  num_lp: int
  num_vs: int


def concatenate(wirings: Sequence[FlatWiring]) -> Tuple[FlatWiring, UnionLayout]:
  """Disjoint union of graphs in the layout of one FactorGraph holding them all.

  The union keeps implicit full Enum tables when every graph has them, and
  lists rows explicitly otherwise.
  """
  G = len(wirings)
  full = all(w.full_tables for w in wirings)
  nvars = np.array([w.num_vars for w in wirings], dtype=np.int64)
  nvs = np.array([w.num_var_states for w in wirings], dtype=np.int64)
  var_off = np.concatenate([[0], np.cumsum(nvars)])
  vs_off = np.concatenate([[0], np.cumsum(nvs)])
  es_index = [np.empty(w.num_edge_states, np.int64) for w in wirings]
  lp_index = [np.empty(w.num_log_potentials, np.int64) for w in wirings]
  vsf, ei, fi = [], [], []
  msgs_range, pot_range = [], []
  enum_ci, enum_ce = [], []
  logical = {}
  es_pos = lp_pos = edge_base = fac_base = 0
  for t in range(4):
    t_es0, t_lp0 = es_pos, lp_pos
    cfg_base = 0
    t_fac_base = 0
    pf, pm, ce_, off = [], [], [], 1 if t != AND else -1
    for g, w in enumerate(wirings):
      m0, m1 = w.msgs_range[t]
      p0, p1 = w.potentials_range[t]
      n = m1 - m0
      es_index[g][m0:m1] = es_pos + np.arange(n)
      lp_index[g][p0:p1] = lp_pos + np.arange(p1 - p0)
      if n:
        le = w.edge_indices_for_edge_states[m0:m1]
        lf = w.factor_indices_for_edge_states[m0:m1]
        vsf.append(w.var_states_for_edge_states[m0:m1] + vs_off[g])
        ei.append(le - le[0] + edge_base)
        fi.append(lf - lf[0] + fac_base)
        nf_local = int(lf[-1] - lf[0] + 1)
        edge_base += int(le[-1] - le[0] + 1)
        fac_base += nf_local
        if t == ENUM:
          if not full:
            w_ci, w_ce = w.enum_rows()
            enum_ci.append(w_ci + cfg_base)
            enum_ce.append(w_ce + (es_pos - t_es0))
            cfg_base += int(w_ci[-1]) + 1 if w_ci.size else 0
        else:
          a_pfi, a_pmi, a_ces, a_off = w.logical[t]
          pf.append(np.asarray(a_pfi) + t_fac_base)
          pm.append(np.asarray(a_pmi) + (es_pos - t_es0))
          ce_.append(np.asarray(a_ces) + (es_pos - t_es0))
          off = a_off
          t_fac_base += nf_local
      es_pos += n
      lp_pos += p1 - p0
    msgs_range.append((t_es0, es_pos))
    pot_range.append((t_lp0, lp_pos))
    if t != ENUM:
      cat = lambda xs: np.concatenate(xs) if xs else np.zeros(0, np.int64)
      logical[t] = (cat(pf), cat(pm), cat(ce_), off)
  cat = lambda xs: np.concatenate(xs) if xs else np.zeros(0, np.int64)
  union = FlatWiring(
      num_var_states=int(vs_off[-1]),
      num_log_potentials=int(lp_pos),
      var_states_for_edge_states=cat(vsf),
      edge_indices_for_edge_states=cat(ei),
      factor_indices_for_edge_states=cat(fi),
      evidence_to_vars=cat([w.evidence_to_vars + var_off[g] for g, w in enumerate(wirings)]),
      msgs_range=msgs_range,
      potentials_range=pot_range,
      enum_ci=None if full else cat(enum_ci),
      enum_ce=None if full else cat(enum_ce),
      logical=logical,
  )
  layout = UnionLayout(var_offsets=var_off, vs_offsets=vs_off, es_index=es_index,
                       lp_index=lp_index, num_es=int(es_pos), num_lp=int(lp_pos),
                       num_vs=int(vs_off[-1]))
  return union, layout


class BatchedBP:
  """Loopy BP over a batch of independent, uneven graphs in one launch.

  Every graph must be connected (it is then one engine task); results per
  graph equal running ``BP`` on that graph alone, bit for bit.
  """

  def __init__(self, wirings: Sequence[FlatWiring], temperature: float = 0.0):
    self.wirings = list(wirings)
    self.temperature = float(temperature)
    self.union, self.layout = concatenate(self.wirings)
    prev = _lib.set_option("merge", 0)
    try:
      self.graph = engine.Graph(self.union.to_engine_wiring())
    finally:
      _lib.set_option("merge", prev)
    tasks = np.zeros(4 * max(self.graph.num_tasks, 1), dtype=np.int32)
    _lib.check(_lib.load().pgx_graph_tasks(self.graph.handle, tasks.ctypes.data), "pgx_graph_tasks")
    tasks = tasks.reshape(-1, 4)[:self.graph.num_tasks]
    self.task_graph = np.searchsorted(self.layout.var_offsets, tasks[:, 0], side="right") - 1
    if np.bincount(self.task_graph, minlength=len(self.wirings)).max(initial=0) > 1:
      raise ValueError("every graph of a BatchedBP must be connected")
    self._es_perm = None

  @property
  def num_graphs(self) -> int:
    return len(self.wirings)

  def close(self) -> None:
    """Releases the compiled graph's device memory now."""
    self.graph.close()

  # ---- packing per-graph arrays into the union layout -------------------
  def _pack(self, parts, index, size):
    torch = _lib.require_gpu()
    out = torch.empty((size,), dtype=torch.float32, device="cuda")
    for g, p in enumerate(parts):
      if index[g].size:
        out[torch.as_tensor(index[g], device="cuda")] = torch.as_tensor(
            np.asarray(_arrays.to_numpy(p), dtype=np.float32).reshape(-1), device="cuda")
    return out

  def pack_log_potentials(self, parts):
    return self._pack(parts, self.layout.lp_index, self.layout.num_lp)

  def pack_ftov(self, parts):
    return self._pack(parts, self.layout.es_index, self.layout.num_es)

  def pack_evidence(self, parts):
    torch = _lib.require_gpu()
    return torch.cat([torch.as_tensor(np.asarray(_arrays.to_numpy(p), np.float32).reshape(-1))
                      for p in parts]).to("cuda")

  def unpack_ftov(self, union):
    idx = [np.asarray(i) for i in self.layout.es_index]
    return [union[..., i] for i in idx]

  def unpack_beliefs(self, union):
    o = self.layout.vs_offsets
    return [union[..., o[g]:o[g + 1]] for g in range(self.num_graphs)]

  # ---- runs ---------------------------------------------------------------
  def _inputs(self, log_potentials, ftov_msgs, evidence):
    torch = _lib.require_gpu()
    g = self.graph
    lp = engine._as_batched(log_potentials, g.num_lp, "log_potentials")[0]
    ev = engine._as_batched(evidence, g.num_vs, "evidence")[0]
    if ftov_msgs is None:
      ftov_msgs = torch.zeros((g.num_es,), dtype=torch.float32, device="cuda")
    ft = engine._as_batched(ftov_msgs, g.num_es, "ftov_msgs")[0]
    for t, n in ((lp, g.num_lp), (ev, g.num_vs), (ft, g.num_es)):
      if t.ndim != 1:
        raise ValueError("BatchedBP takes one union array per input (no extra batch axis)")
    return lp, ft, ev

  def autotune(self, log_potentials, ftov_msgs, evidence, damping: float = 0.5,
               temperature: Optional[float] = None, converge: bool = False,
               probe_iters: int = 10):
    """The project's autotuner (pgx_autotune) on this union: times the
    one-workgroup-per-graph and per-phase shapes and keeps the fastest."""
    t = self.temperature if temperature is None else float(temperature)
    lp, ft, ev = self._inputs(log_potentials, ftov_msgs, evidence)
    return engine.tune_graph(self.graph, lp, 0, ev, 0, ft, 0, 1, damping, t, converge,
                             probe_iters)

  def run(self, log_potentials, ftov_msgs, evidence, num_iters: int, damping: float = 0.5,
          temperature: Optional[float] = None):
    """Fixed iteration count; returns (union ftov, global deltas per iteration)."""
    torch = _lib.require_gpu()
    t = self.temperature if temperature is None else float(temperature)
    lp, ft, ev = self._inputs(log_potentials, ftov_msgs, evidence)
    g = self.graph
    n_it = max(int(num_iters), 1)
    out = torch.empty((g.num_es,), device="cuda", dtype=torch.float32)
    deltas = torch.empty((n_it,), device="cuda", dtype=torch.float32)
    _lib.check(_lib.load().pgx_run(
        g.handle, _lib.ptr(lp), 0, _lib.ptr(ev), 0, _lib.ptr(ft), 0, _lib.ptr(out),
        _lib.ptr(deltas), 1, n_it, float(damping), t, _lib.current_stream()), "pgx_run")
    return out, deltas

  def run_converge(self, log_potentials, ftov_msgs, evidence, max_iters: int,
                   tol: float = 1e-4, damping: float = 0.5,
                   temperature: Optional[float] = None):
    """Each graph runs until its max-abs message change is <= tol (or max_iters).

    Returns (union ftov, per-graph deltas [G, max_iters] zero past the stop,
    per-graph iteration counts [G]).
    """
    torch = _lib.require_gpu()
    t = self.temperature if temperature is None else float(temperature)
    lp, ft, ev = self._inputs(log_potentials, ftov_msgs, evidence)
    g = self.graph
    n_it = max(int(max_iters), 1)
    out = torch.empty((g.num_es,), device="cuda", dtype=torch.float32)
    task_deltas = torch.empty((g.num_tasks, n_it), device="cuda", dtype=torch.float32)
    iters = torch.empty((g.num_tasks,), device="cuda", dtype=torch.int32)
    _lib.check(_lib.load().pgx_run_converge(
        g.handle, _lib.ptr(lp), 0, _lib.ptr(ev), 0, _lib.ptr(ft), 0, _lib.ptr(out),
        _lib.ptr(task_deltas), _lib.ptr(iters), 1, n_it, float(damping), t, float(tol),
        _lib.current_stream()), "pgx_run_converge")
    order = torch.as_tensor(np.argsort(self.task_graph), device="cuda")
    return out, task_deltas[order], iters[order]

  def beliefs(self, evidence, ftov_msgs):
    torch = _lib.require_gpu()
    g = self.graph
    ev = engine._as_batched(evidence, g.num_vs, "evidence")[0]
    ft = engine._as_batched(ftov_msgs, g.num_es, "ftov_msgs")[0]
    out = torch.empty((g.num_vs,), device="cuda", dtype=torch.float32)
    _lib.check(_lib.load().pgx_get_beliefs(g.handle, _lib.ptr(ev), 0, _lib.ptr(ft), 0,
                                           _lib.ptr(out), 1, _lib.current_stream()),
               "pgx_get_beliefs")
    return out
