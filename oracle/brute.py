"""Brute-force enumeration over tiny factor graphs -- TEST INFRASTRUCTURE ONLY.

Exact log-partition, marginals and MAP of the distribution
``p(x) ∝ exp(sum_v ev[v, x_v] + sum_f lp_f(x_f))`` defined by flat
EnumFactor wiring. On a tree, converged sum-product BP (temperature 1)
reproduces these marginals and max-product BP the max-marginals, which is a
check on the oracle that does not depend on the reference's code.
"""

import itertools
from typing import Dict, List, Tuple

import numpy as np


def enum_factor_tables(graph: Dict, log_potentials: np.ndarray):
  """Per-factor {tuple(var_state_indices): log_potential} for Enum factors."""
  m0, m1 = graph["msgs_range"]["enum"]
  p0, _ = graph["potentials_range"]["enum"]
  a = graph["args"]["enum"]
  vs = graph["var_states_for_edge_states"][m0:m1]
  cfg = a["factor_configs_indices"]
  es = a["factor_configs_edge_states"]
  rows_by_cfg: Dict[int, List[int]] = {}
  for c, e in zip(cfg, es):
    rows_by_cfg.setdefault(int(c), []).append(int(vs[e]))
  tables: Dict[Tuple[int, ...], List] = {}
  # group configs by the variables they touch (one factor each)
  by_factor: Dict[int, Dict[Tuple[int, ...], float]] = {}
  fi = graph.get("factor_indices_for_edge_states")
  for c, states in rows_by_cfg.items():
    first_es = int(es[np.searchsorted(cfg, c)])
    f = int(fi[m0 + first_es]) if fi is not None else c
    by_factor.setdefault(f, {})[tuple(states)] = float(log_potentials[p0 + c])
  return by_factor


def joint_log_weights(num_states: List[int], var_starts: List[int],
                      evidence: np.ndarray, factor_tables):
  """log weight of every joint assignment (rows in itertools.product order)."""
  assignments = np.array(list(itertools.product(*[range(s) for s in num_states])))
  logw = np.zeros(assignments.shape[0])
  vstate = assignments + np.asarray(var_starts)[None, :]
  logw += evidence[vstate].sum(axis=1)
  state_to_var = {}
  for v, (st, ns) in enumerate(zip(var_starts, num_states)):
    for s in range(ns):
      state_to_var[st + s] = v
  for table in factor_tables.values():
    keys = list(table.keys())
    vars_of_factor = [state_to_var[s] for s in keys[0]]
    firsts = [var_starts[v] for v in vars_of_factor]
    lookup = {tuple(k[j] - firsts[j] for j in range(len(k))): w for k, w in table.items()}
    sub = assignments[:, vars_of_factor]
    logw += np.array([lookup.get(tuple(r), -np.inf) for r in sub])
  return assignments, logw


def exact_marginals(num_states, var_starts, evidence, factor_tables, temperature=1.0):
  """Per-variable (soft-)max-marginals; temperature 0 gives max-marginals."""
  assignments, logw = joint_log_weights(num_states, var_starts, evidence, factor_tables)
  out = []
  for v, ns in enumerate(num_states):
    vals = []
    for s in range(ns):
      sel = logw[assignments[:, v] == s]
      if temperature == 0.0:
        vals.append(sel.max())
      else:
        m = sel.max()
        vals.append(m + temperature * np.log(np.sum(np.exp((sel - m) / temperature))))
    vals = np.array(vals)
    if temperature == 0.0:
      out.append(vals - vals.max())
    else:
      p = np.exp((vals - vals.max()) / temperature)
      out.append(p / p.sum())
  return out, assignments[np.argmax(logw)]
