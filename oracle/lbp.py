"""CPU oracle for loopy BP -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg may import this module, and only as the checker. The product path
(``pgmax_amd``) never calls it.

A NumPy restatement of the reference's BP iteration, operation by operation,
in a chosen float dtype: ``np.float32`` reproduces the reference's FP32 op
order (sequential scatter-adds via ``np.add.at``, NaN-propagating maxima via
``np.maximum.at``), ``np.float64`` is the high-precision oracle the FP32
results are compared with.

Pinning (see tests/test_oracle.py): the PGMax end-to-end golden of 84
max-product messages and 12 MAP states (tests/test_pgmax.py:35-421 of the
reference); exact marginals / MAP on trees by brute-force enumeration;
the 2-state chain transfer-matrix closed form; OR/AND/Pool closed forms
equal to their EnumFactor expansions (reference tests/factor/test_or.py,
test_and.py, test_pool.py) at every temperature; the clipping invariant
(reference tests/test_clipping.py).
"""

from typing import Dict, Tuple

import numpy as np

from oracle import f32math

MSG_NEG_INF = -1e32  # pgmax/utils/__init__.py:26
LOG_POTENTIAL_MAX_ABS = 1e6  # pgmax/utils/__init__.py:32
NEG_INF = -np.inf  # pgmax/utils/__init__.py:37
TEMPERATURE_STABILITY_THRE = 0.5  # pgmax/factor/logical.py:33
VAR_SUM_CHUNK = 32  # edges per partial sum of a variable state


def _f(dtype, x):
  return np.asarray(x, dtype=dtype)


def _m(dt):
  """Transcendentals for a dtype: the kernels' FP32 algorithms for float32
  (oracle/f32math.py, bit-identical to pgmax_amd/csrc/pgx_math.h), NumPy's
  for float64."""
  if np.dtype(dt) == np.float32:
    from oracle import f32math  # pylint: disable=g-import-not-at-top
    return f32math
  return np


# ---------------------------------------------------------------------------
# Segment helpers: pgmax/factor/update_utils.py
# ---------------------------------------------------------------------------


def get_maxes_and_argmaxes(data, labels, num_labels):
  """Per-label max and LAST argmax among ties (update_utils.py:25-64)."""
  n = data.shape[0]
  maxes = np.full((num_labels,), NEG_INF, dtype=data.dtype)
  np.maximum.at(maxes, labels, data)
  only_maxes_pos = np.arange(n, dtype=np.int64) - n * np.where(
      data != maxes[labels], 1, 0
  )
  argmaxes = np.full((num_labels,), np.iinfo(np.int32).min, dtype=np.int64)
  np.maximum.at(argmaxes, labels, only_maxes_pos)
  return maxes, argmaxes


def logsumexps_with_temp(data, labels, num_labels, temperature, maxes=None):
  """Per-label T*log(sum exp(x/T)) (update_utils.py:67-98)."""
  dt = data.dtype
  if maxes is None:
    maxes = np.full((num_labels,), NEG_INF, dtype=dt)
    np.maximum.at(maxes, labels, data)
  sums = np.zeros((num_labels,), dtype=dt)
  with np.errstate(invalid="ignore", over="ignore"):
    np.add.at(sums, labels, _m(dt).exp((data - maxes[labels]) / dt.type(temperature)))
  with np.errstate(divide="ignore", invalid="ignore"):
    return dt.type(temperature) * _m(dt).log(sums) + maxes


def log1mexp(x):
  """log(1 - exp(-x)) for x >= 0 (update_utils.py:134-147)."""
  dt = x.dtype
  m = _m(dt)
  with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
    return np.where(
        x <= dt.type(np.log(2)), m.log(-m.expm1(-x)), m.log1p(-m.exp(-x))
    )


def logaddexp_with_temp(data1, data2, temperature):
  """T*log(exp(a/T) + exp(b/T)) (update_utils.py:150-166)."""
  dt = data1.dtype
  t = dt.type(temperature)
  maxes = np.maximum(data1, data2)
  mins = np.minimum(data1, data2)
  m = _m(dt)
  with np.errstate(invalid="ignore", over="ignore"):
    return t * m.log1p(m.exp((mins - maxes) / t)) + maxes


def logminusexp_with_temp(data1, data2, temperature, eps=1e-30):
  """T*log(exp(a/T) - exp(b/T)) for a >= b, NEG_INF when a < b + eps (update_utils.py:169-190)."""
  dt = data1.dtype
  t = dt.type(temperature)
  with np.errstate(invalid="ignore", over="ignore"):
    return np.where(
        data1 >= data2 + dt.type(eps),
        t * log1mexp((data1 - data2) / t) + data1,
        dt.type(NEG_INF),
    )


# ---------------------------------------------------------------------------
# Variable side and normalisation: pgmax/infer/bp.py
# ---------------------------------------------------------------------------


def chunked_segment_sum(values, segments, num_segments):
  """Per-segment sums in the engine's fixed order.

  The values of a segment, in array order, are summed in runs of
  VAR_SUM_CHUNK (sequentially, from 0); runs of VAR_SUM_CHUNK run sums form
  groups (summed sequentially, from 0); the group sums are added in order,
  from 0. Up to VAR_SUM_CHUNK values this is the plain sequential sum, and
  up to VAR_SUM_CHUNK**2 values it is "run sums added in order".
  """
  values = np.asarray(values)
  seg = np.asarray(segments)
  dt = values.dtype
  counts = np.bincount(seg, minlength=num_segments)
  if counts.max(initial=0) <= VAR_SUM_CHUNK:
    sums = np.zeros(num_segments, dtype=dt)
    np.add.at(sums, seg, values)
    return sums
  order = np.argsort(seg, kind="stable")
  first = np.concatenate([[0], np.cumsum(counts)])[:-1]
  rank = np.empty_like(order)
  rank[order] = np.arange(seg.shape[0]) - first[seg[order]]
  n1 = (counts + VAR_SUM_CHUNK - 1) // VAR_SUM_CHUNK
  c0 = np.concatenate([[0], np.cumsum(n1)])[:-1]
  run_sums = np.zeros(int(n1.sum()), dtype=dt)
  np.add.at(run_sums, c0[seg] + rank // VAR_SUM_CHUNK, values)
  run_seg = np.repeat(np.arange(num_segments), n1)
  run_rank = np.arange(run_sums.shape[0]) - c0[run_seg]
  n2 = (n1 + VAR_SUM_CHUNK - 1) // VAR_SUM_CHUNK
  g0 = np.concatenate([[0], np.cumsum(n2)])[:-1]
  group_sums = np.zeros(int(n2.sum()), dtype=dt)
  np.add.at(group_sums, g0[run_seg] + run_rank // VAR_SUM_CHUNK, run_sums)
  sums = np.zeros(num_segments, dtype=dt)
  np.add.at(sums, np.repeat(np.arange(num_segments), n2), group_sums)
  return sums


def var_state_sums(ftov_msgs, evidence, var_states_for_edge_states):
  """evidence + scatter_add(ftov) per variable state (bp.py:217, inferer.py:218-222).

  Accumulation order: the messages of a variable state are summed in
  edge-state order starting from 0, then the evidence is added. XLA does not
  fix the order of its scatter-add; this order reproduces the reference's
  84-message golden (tests/test_pgmax.py:63-152) within 2.4e-7, whereas
  adding onto the evidence first is 1.13e-6 away from it. Variables with
  more than VAR_SUM_CHUNK edges sum in runs and groups of runs
  (chunked_segment_sum), so many lanes can share one variable; up to
  VAR_SUM_CHUNK edges this is the plain sequential sum. The HIP kernels use
  the same order.
  """
  n_vs = np.shape(evidence)[-1]
  return chunked_segment_sum(ftov_msgs, var_states_for_edge_states, n_vs) + evidence


def pass_var_to_fac_messages(ftov_msgs, evidence, var_states_for_edge_states):
  """vtof = (evidence + scatter_add(ftov))[vs] - ftov (bp.py:191-219)."""
  var_sums = var_state_sums(ftov_msgs, evidence, var_states_for_edge_states)
  return var_sums[var_states_for_edge_states] - ftov_msgs


def normalize_and_clip_msgs(msgs, edge_indices_for_edge_states, num_edges):
  """Subtract each edge's max, clip at MSG_NEG_INF (bp.py:222-260)."""
  max_by_edges = np.full((num_edges,), NEG_INF, dtype=msgs.dtype)
  np.maximum.at(max_by_edges, edge_indices_for_edge_states, msgs)
  with np.errstate(invalid="ignore"):
    norm = msgs - max_by_edges[edge_indices_for_edge_states]
  return np.maximum(norm, msgs.dtype.type(MSG_NEG_INF))


# ---------------------------------------------------------------------------
# Factor updates
# ---------------------------------------------------------------------------


def pass_enum_fac_to_var_messages(
    vtof_msgs, factor_configs_indices, factor_configs_edge_states,
    log_potentials, num_val_configs, temperature,
):
  """EnumFactor -> variable messages (pgmax/factor/enum.py:451-475)."""
  dt = vtof_msgs.dtype
  summary = np.zeros((num_val_configs,), dtype=dt)
  np.add.at(summary, factor_configs_indices, vtof_msgs[factor_configs_edge_states])
  summary = summary + log_potentials
  rows = summary[factor_configs_indices]
  maxes = np.full((vtof_msgs.shape[0],), -np.inf, dtype=dt)
  np.maximum.at(maxes, factor_configs_edge_states, rows)
  if temperature == 0.0:
    out = maxes
  else:
    out = logsumexps_with_temp(
        rows, factor_configs_edge_states, vtof_msgs.shape[0], temperature, maxes
    )
  with np.errstate(invalid="ignore"):
    return out - vtof_msgs


def _bincount_first(factor_indices, num_factors):
  counts = np.bincount(factor_indices, minlength=num_factors)[:num_factors]
  first = np.concatenate([[0], np.cumsum(counts)])[:-1]
  return counts, first


def pass_logical_fac_to_var_messages(
    vtof_msgs, parents_factor_indices, parents_msg_indices,
    children_edge_states, edge_states_offset, temperature, normalize=True,
):
  """OR/AND closed forms (pgmax/factor/logical.py:561-779)."""
  dt = vtof_msgs.dtype
  nf = children_edge_states.shape[0]
  pfi = parents_factor_indices
  off = edge_states_offset
  p_rel = vtof_msgs[parents_msg_indices + off]
  p_oth = vtof_msgs[parents_msg_indices]
  p_diff = p_rel - p_oth
  c_rel = vtof_msgs[children_edge_states + off]
  c_oth = vtof_msgs[children_edge_states]

  # per-factor sums over parents in the engine's chunked order (plain
  # sequential order up to VAR_SUM_CHUNK parents)
  s_oth = chunked_segment_sum(p_oth, pfi, nf)
  children_msgs_other = s_oth

  m1, a1 = get_maxes_and_argmaxes(p_diff, pfi, nf)
  wo = np.array(p_diff, copy=True)
  wo[a1] = NEG_INF
  m2 = np.full((nf,), NEG_INF, dtype=dt)
  np.maximum.at(m2, pfi, wo)

  zero = dt.type(0.0)
  if temperature == 0.0:
    mx = np.maximum(p_oth, p_rel)
    s_mx = chunked_segment_sum(mx, pfi, nf)
    par_rel = s_mx[pfi] + c_rel[pfi] - mx
    mz = np.minimum(zero, m1)
    children_msgs_relevant = s_mx + mz
    opt1 = c_oth[pfi] + s_oth[pfi] - p_oth
    opt2 = par_rel + mz[pfi]
    opt2_am = par_rel[a1] + np.minimum(zero, m2)
    opt2[a1] = opt2_am
    par_oth = np.maximum(opt1, opt2)
  else:
    lse = logaddexp_with_temp(p_rel, p_oth, temperature)
    s_lse = chunked_segment_sum(lse, pfi, nf)
    with np.errstate(invalid="ignore"):
      s_lse_wo = s_lse[pfi] - lse
      s_oth_wo = s_oth[pfi] - p_oth
    par_rel = c_rel[pfi] + s_lse_wo
    children_msgs_relevant = logminusexp_with_temp(s_lse, s_oth, temperature, eps=1e-4)
    if temperature < TEMPERATURE_STABILITY_THRE:
      children_msgs_relevant = np.maximum(
          children_msgs_relevant,
          logaddexp_with_temp(s_oth + m1, s_oth + m2, temperature),
      )
    opt1 = c_oth[pfi] + s_oth_wo
    opt2 = c_rel[pfi] + s_lse_wo
    opt3 = c_rel[pfi] + s_oth_wo
    la12 = logaddexp_with_temp(opt1, opt2, temperature)
    par_oth = logminusexp_with_temp(la12, opt3, temperature, eps=1e-4)
    if temperature < TEMPERATURE_STABILITY_THRE:
      x = s_oth_wo + m1[pfi]
      x[a1] = s_oth_wo[a1] + m2
      lb = logaddexp_with_temp(opt1, c_rel[pfi] + x, temperature)
      par_oth = np.maximum(par_oth, lb)

  counts, first = _bincount_first(pfi, nf)
  single = counts == 1
  par_rel = np.array(par_rel, copy=True)
  par_oth = np.array(par_oth, copy=True)
  par_rel[first] = np.where(single, c_rel, par_rel[first])
  par_oth[first] = np.where(single, c_oth, par_oth[first])

  out = np.zeros_like(vtof_msgs)
  with np.errstate(invalid="ignore"):
    if normalize:
      out[parents_msg_indices + off] = par_rel - par_oth
      out[children_edge_states + off] = children_msgs_relevant - children_msgs_other
    else:
      out[parents_msg_indices + off] = par_rel
      out[parents_msg_indices] = par_oth
      out[children_edge_states + off] = children_msgs_relevant
      out[children_edge_states] = children_msgs_other
  return out


def pass_pool_fac_to_var_messages(
    vtof_msgs, pool_choices_factor_indices, pool_choices_msg_indices,
    pool_indicators_edge_states, temperature, normalize=True,
):
  """Pool closed forms (pgmax/factor/pool.py:328-474)."""
  dt = vtof_msgs.dtype
  nf = pool_indicators_edge_states.shape[0]
  cfi = pool_choices_factor_indices
  cmi = pool_choices_msg_indices
  ies = pool_indicators_edge_states
  pc_diff = vtof_msgs[cmi + 1] - vtof_msgs[cmi]
  pc_zero = vtof_msgs[cmi]
  pi_diff = vtof_msgs[ies + 1] - vtof_msgs[ies]
  pi_one = vtof_msgs[ies + 1]

  s_zero = np.zeros((nf,), dtype=dt)
  np.add.at(s_zero, cfi, pc_zero)
  pc_msgs_ones = s_zero[cfi] + pi_one[cfi] - pc_zero
  pi_msgs_zeros = s_zero

  m1, a1 = get_maxes_and_argmaxes(pc_diff, cfi, nf)
  if temperature == 0.0:
    wo = np.array(pc_diff, copy=True)
    wo[a1] = NEG_INF
    m2 = np.full((nf,), NEG_INF, dtype=dt)
    np.maximum.at(m2, cfi, wo)
    pc_msgs_diffs = np.minimum(pi_diff, -m1)[cfi]
    pc_msgs_diffs[a1] = np.minimum(pi_diff, -m2)
    pi_msgs_diffs = m1
  else:
    pi_msgs_diffs = logsumexps_with_temp(pc_diff, cfi, nf, temperature, maxes=m1)
    f_lse = logaddexp_with_temp(pi_msgs_diffs, -pi_diff, temperature)
    pc_msgs_diffs = -logminusexp_with_temp(f_lse[cfi], pc_diff, temperature)
    ind = np.array(pc_diff, copy=True)
    ind[a1] = -pi_diff
    pc_msgs_diffs_am = -logsumexps_with_temp(ind, cfi, nf, temperature)
    pc_msgs_diffs[a1] = pc_msgs_diffs_am

  counts, first = _bincount_first(cfi, nf)
  single = counts == 1
  pc_msgs_diffs = np.array(pc_msgs_diffs, copy=True)
  pc_msgs_diffs[first] = np.where(single, pi_diff, pc_msgs_diffs[first])
  pc_msgs_ones = np.array(pc_msgs_ones, copy=True)
  pc_msgs_ones[first] = np.where(single, pi_one, pc_msgs_ones[first])

  out = np.zeros_like(vtof_msgs)
  with np.errstate(invalid="ignore"):
    if normalize:
      out[cmi + 1] = pc_msgs_diffs
      out[ies + 1] = pi_msgs_diffs
    else:
      out[cmi + 1] = pc_msgs_ones
      out[cmi] = pc_msgs_ones - pc_msgs_diffs
      out[ies + 1] = pi_msgs_zeros + pi_msgs_diffs
      out[ies] = pi_msgs_zeros
  return out


# ---------------------------------------------------------------------------
# Driver: pgmax/infer/bp.py:63-155
# ---------------------------------------------------------------------------

ENUM, OR, AND, POOL = "enum", "or", "and", "pool"
TYPE_ORDER = (ENUM, OR, AND, POOL)  # pgmax/factor/__init__.py:37-44


def run_with_diffs(
    graph: Dict,
    log_potentials: np.ndarray,
    ftov_msgs: np.ndarray,
    evidence: np.ndarray,
    num_iters: int,
    damping: float = 0.5,
    temperature: float = 0.0,
    dtype=np.float32,
) -> Tuple[np.ndarray, np.ndarray]:
  """Reference BP loop (bp.py:63-155) over the flat wiring in ``graph``.

  ``graph`` holds ``var_states_for_edge_states``, ``edge_indices_for_edge_states``,
  ``num_edges``, ``msgs_range[type]``, ``potentials_range[type]`` and
  ``args[type]`` (the per-type inference arguments, indices relative to the
  type's message range).
  Returns (final ftov messages, per-iteration deltas).
  """
  dt = np.dtype(dtype)
  lp = np.clip(_f(dt, log_potentials), -LOG_POTENTIAL_MAX_ABS, LOG_POTENTIAL_MAX_ABS)
  ev = _f(dt, evidence)
  vs = graph["var_states_for_edge_states"]
  ei = graph["edge_indices_for_edge_states"]
  ne = graph["num_edges"]
  msgs = normalize_and_clip_msgs(_f(dt, ftov_msgs), ei, ne)
  d = dt.type(damping)
  omd = dt.type(1 - damping)
  deltas = []
  for _ in range(max(int(num_iters), 1)):
    vtof = pass_var_to_fac_messages(msgs, ev, vs)
    new = np.zeros_like(vtof)
    for t in TYPE_ORDER:
      m0, m1 = graph["msgs_range"][t]
      if m0 == m1:
        continue
      p0, p1 = graph["potentials_range"][t]
      a = graph["args"][t]
      v = vtof[m0:m1]
      if t == ENUM:
        new[m0:m1] = pass_enum_fac_to_var_messages(
            v, a["factor_configs_indices"], a["factor_configs_edge_states"],
            lp[p0:p1], a["num_val_configs"], temperature,
        )
      elif t in (OR, AND):
        new[m0:m1] = pass_logical_fac_to_var_messages(
            v, a["parents_factor_indices"], a["parents_msg_indices"],
            a["children_edge_states"], a["edge_states_offset"], temperature,
        )
      else:
        new[m0:m1] = pass_pool_fac_to_var_messages(
            v, a["pool_choices_factor_indices"], a["pool_choices_msg_indices"],
            a["pool_indicators_edge_states"], temperature,
        )
    with np.errstate(invalid="ignore"):
      new = d * msgs + omd * new
    new = normalize_and_clip_msgs(new, ei, ne)
    with np.errstate(invalid="ignore"):
      deltas.append(np.max(np.abs(new - msgs)))
    msgs = new
  return msgs, np.array(deltas, dtype=dt)


def get_flat_beliefs(graph: Dict, ftov_msgs, evidence, dtype=np.float32):
  """evidence + scatter_add(ftov) (pgmax/infer/inferer.py:218-222)."""
  dt = np.dtype(dtype)
  return var_state_sums(
      _f(dt, ftov_msgs), _f(dt, evidence), graph["var_states_for_edge_states"]
  )


def map_and_marginals(flat_beliefs, var_vs0):
  """MAP state (first maximal state; a NaN counts as maximal) and marginals
  (exp(b - max) / sum of exp(b - max) per variable, summed in state order) of
  FP32 log-beliefs, as decode_map_states (pgmax/infer/inferer.py:251-264) and
  get_marginals (pgmax/infer/bp.py:263-288, jax.nn.softmax) define them.
  var_vs0[v] is variable v's first state, var_vs0[-1] the state count."""
  x = np.asarray(flat_beliefs, np.float32)
  vs0 = np.asarray(var_vs0, np.int64)
  ns = np.diff(vs0)
  nv = ns.shape[0]
  m = np.full(nv, -np.inf, np.float32)
  best = np.zeros(nv, np.int64)
  done = np.zeros(nv, bool)  # a NaN was met: it is the max and the argmax
  for k in range(int(ns.max(initial=0))):
    live = (k < ns) & ~done
    xk = np.where(live, x[np.minimum(vs0[:-1] + k, x.shape[0] - 1)], np.float32(0))
    take = live & ((k == 0) | (xk > m) | np.isnan(xk))
    m = np.where(take, xk, m)
    best = np.where(take, k, best)
    done |= live & np.isnan(xk)
  var_of = np.repeat(np.arange(nv), ns)
  with np.errstate(invalid="ignore"):
    e = f32math.exp(x - m[var_of])
    sums = np.zeros(nv, np.float32)
    for k in range(int(ns.max(initial=0))):
      live = k < ns
      sums[live] = sums[live] + e[vs0[:-1][live] + k]
    marg = (e / sums[var_of]).astype(np.float32)
  return best.astype(np.int32), marg


def graph_from_context(ctx) -> Dict:
  """Flat wiring dict from a pgmax_amd InfererContext (host contract object)."""
  from pgmax_amd.factor import FAC_TO_VAR_UPDATES  # pylint: disable=g-import-not-at-top
  names = dict(zip(FAC_TO_VAR_UPDATES.keys(), TYPE_ORDER))
  g = {
      "var_states_for_edge_states": np.asarray(ctx.var_states_for_edge_states),
      "edge_indices_for_edge_states": np.asarray(ctx.edge_indices_for_edge_states),
      "factor_indices_for_edge_states": np.asarray(ctx.factor_indices_for_edge_states),
      "num_edges": int(ctx.num_edges),
      "msgs_range": {},
      "potentials_range": {},
      "args": {},
  }
  for t, name in names.items():
    g["msgs_range"][name] = tuple(ctx.factor_type_to_msgs_range[t])
    g["potentials_range"][name] = tuple(ctx.factor_type_to_potentials_range[t])
    g["args"][name] = {
        k: (np.asarray(v) if isinstance(v, np.ndarray) else v)
        for k, v in ctx.inference_arguments[t].items()
    }
  return g


def graph_from_flat(flat) -> Dict:
  """Flat wiring dict from a pgmax_amd.batch.FlatWiring (Enum rows listed
  explicitly, full tables expanded in row-major config order)."""
  ci, ce = flat.enum_rows()
  g = {
      "var_states_for_edge_states": np.asarray(flat.var_states_for_edge_states),
      "edge_indices_for_edge_states": np.asarray(flat.edge_indices_for_edge_states),
      "factor_indices_for_edge_states": np.asarray(flat.factor_indices_for_edge_states),
      "num_edges": int(flat.edge_indices_for_edge_states[-1]) + 1
                   if flat.edge_indices_for_edge_states.size else 0,
      "msgs_range": dict(zip(TYPE_ORDER, (tuple(r) for r in flat.msgs_range))),
      "potentials_range": dict(zip(TYPE_ORDER, (tuple(r) for r in flat.potentials_range))),
      "args": {"enum": {"factor_configs_indices": np.asarray(ci),
                        "factor_configs_edge_states": np.asarray(ce),
                        "num_val_configs": int(np.asarray(ci).max()) + 1 if np.size(ci) else 0}},
  }
  keys = ("parents_factor_indices", "parents_msg_indices", "children_edge_states",
          "edge_states_offset")
  pool_keys = ("pool_choices_factor_indices", "pool_choices_msg_indices",
               "pool_indicators_edge_states")
  for code, name in ((1, "or"), (2, "and"), (3, "pool")):
    a = flat.logical.get(code)
    if a is None:
      continue
    if name == "pool":
      g["args"][name] = dict(zip(pool_keys, a[:3]))
    else:
      g["args"][name] = dict(zip(keys, a))
  return g
