"""Reverse mode of BP for graphs of Enum factors -- TEST INFRASTRUCTURE ONLY.

The vector-Jacobian product of the final messages of ``run``
(pgmax/infer/bp.py:63-155) with respect to the log-potentials, the evidence
and the initial messages: what ``jax.grad`` through ``bp.run`` gives in the
reference's learning example (examples/grid_mrf.ipynb cells 15-16), derived
by hand step by step:

  forward, per iteration (Enum factors only):
    vs  = evidence + segment_sum(m, var_state)          (bp.py:217)
    vt  = vs[var_state] - m                             (bp.py:218)
    s_c = clip(lp_c) + sum of vt over the rows of c     (enum.py:451-460)
    o_k = T log sum_{c ∋ k} exp(s_c / T)  (max at T=0)  (update_utils.py:67-98)
    f_k = o_k - vt_k                                    (enum.py:474)
    x   = d m + (1 - d) f                               (bp.py:127)
    m'  = max(x - max_edge(x), -1e32)                   (bp.py:249-259)
  and m_0 = the same normalisation of the initial messages (bp.py:92-96).

Conventions where the reference's derivative is not unique (a maximum
attained by several entries, a clip at its boundary): the derivative of a
maximum is shared equally by the entries that attain it (as lax.reduce_max's
rule does) and max(a, b) / clip pass half at equality (as lax.max does).
Away from ties and boundaries these choices do not matter; tests pin the
product against central finite differences there.
"""

import numpy as np

MSG_NEG_INF = -1e32
LP_MAX = 1e6


def _normalize(x, edge, n_edges):
  mx = np.full(n_edges, -np.inf)
  np.maximum.at(mx, edge, x)
  return np.maximum(x - mx[edge], MSG_NEG_INF), mx


def _normalize_bwd(x, gy, edge, n_edges):
  """VJP of max(x - max_edge(x), -1e32)."""
  y_raw = x - _normalize(x, edge, n_edges)[1][edge]
  keep = np.where(y_raw > MSG_NEG_INF, 1.0, np.where(y_raw == MSG_NEG_INF, 0.5, 0.0))
  gk = gy * keep
  mx = np.full(n_edges, -np.inf)
  np.maximum.at(mx, edge, x)
  at_max = (x == mx[edge]).astype(np.float64)
  cnt = np.zeros(n_edges)
  np.add.at(cnt, edge, at_max)
  tot = np.zeros(n_edges)
  np.add.at(tot, edge, gk)
  return gk - at_max / np.maximum(cnt[edge], 1.0) * tot[edge]


def var_sum_pullback(g, gvt):
  """Cotangent of the variable sums: each sum feeds the vt of its edge
  states, so it collects their cotangents."""
  gvs = np.zeros(g.num_vs)
  np.add.at(gvs, g.vsi, gvt)
  return gvs


class EnumGraph:
  """The arrays of a graph: var_state[E_s], edge[E_s], the Enum rows
  (config, edge state; configs number the log-potentials) and, optionally,
  OR / AND factor types (``logical``: LogicalType records) and Pool factors
  (``pool``: PoolType records), after the Enum edge states, which are the
  first ``num_enum_es``."""

  def __init__(self, var_state, edge, row_cfg, row_es, num_vs, num_cfg, logical=(),
               num_enum_es=None, pool=()):
    self.vsi = np.asarray(var_state, np.int64)
    self.edge = np.asarray(edge, np.int64)
    self.n_edges = int(self.edge.max()) + 1 if self.edge.size else 0
    self.row_cfg = np.asarray(row_cfg, np.int64)
    self.row_es = np.asarray(row_es, np.int64)
    self.num_vs = int(num_vs)
    self.num_cfg = int(num_cfg)
    self.num_es = self.vsi.shape[0]
    self.logical = list(logical)
    self.pool = list(pool)
    self.num_enum_es = self.num_es if num_enum_es is None else int(num_enum_es)


class PoolType:
  """Pool factors whose messages start at m0: choices (factor, message
  index relative to m0) and indicators (edge state relative to m0), as
  pgmax/factor/pool.py:275-326 lays them out."""

  def __init__(self, m0, choices_factor_indices, choices_msg_indices, indicators_edge_states):
    self.m0 = int(m0)
    self.cfi = np.asarray(choices_factor_indices, np.int64)
    self.cmi = np.asarray(choices_msg_indices, np.int64)
    self.ies = np.asarray(indicators_edge_states, np.int64)


class LogicalType:
  """OR (off = +1) or AND (off = -1) factors whose messages start at m0:
  parents (factor, message index relative to m0), children (edge state
  relative to m0), as pgmax/factor/logical.py:457-488 lays them out."""

  def __init__(self, m0, parents_factor_indices, parents_msg_indices, children_edge_states,
               edge_states_offset):
    self.m0 = int(m0)
    self.pfi = np.asarray(parents_factor_indices, np.int64)
    self.pmi = np.asarray(parents_msg_indices, np.int64)
    self.ces = np.asarray(children_edge_states, np.int64)
    self.off = int(edge_states_offset)


# ---- OR / AND updates with their partial derivatives ------------------------
# pgmax/factor/logical.py:561-779 (oracle/lbp.py pass_logical_fac_to_var_messages
# states the same formulas); eps of the log-minus-exps 1e-4, stability
# branch below T = 0.5 (logical.py:33).

LME_EPS = 1e-4
STABILITY_T = 0.5


def _lae(a, b, t):
  """T log(e^(a/T) + e^(b/T)) and its partials."""
  with np.errstate(over="ignore", invalid="ignore", divide="ignore"):
    mx, mn = max(a, b), min(a, b)
    val = t * np.log1p(np.exp((mn - mx) / t)) + mx
    da = 1.0 / (1.0 + np.exp((b - a) / t))
  return val, da, 1.0 - da


def _lme(a, b, t):
  """T log(e^(a/T) - e^(b/T)) if a >= b + eps, else -inf; and its partials."""
  if not a >= b + LME_EPS:
    return -np.inf, 0.0, 0.0
  q = np.exp((b - a) / t)
  return a + t * np.log1p(-q), 1.0 / (1.0 - q), -q / (1.0 - q)


def _max2(a, b):
  """max(a, b) and its partials (half each at equality)."""
  if a > b:
    return a, 1.0, 0.0
  if b > a:
    return b, 0.0, 1.0
  return a, 0.5, 0.5


def _dmin0(m):
  """d min(0, m) / dm, half at 0."""
  return 1.0 if m < 0 else (0.5 if m == 0 else 0.0)


def _top2(d):
  """m1, the last index attaining it, m2 = max over the other indices
  (update_utils.py:25-64), and the cotangent maps of m1 and m2 onto d
  (shared equally by the indices attaining each)."""
  n = d.shape[0]
  m1 = d.max()
  a1 = int(np.flatnonzero(d == m1)[-1])
  rest = np.delete(np.arange(n), a1)
  m2 = d[rest].max() if n > 1 else -np.inf
  w1 = (d == m1).astype(np.float64)
  w1 /= w1.sum()
  w2 = np.zeros(n)
  if n > 1:
    hit = rest[d[rest] == m2]
    w2[hit] = 1.0 / hit.size
  return m1, a1, m2, w1, w2


def logical_factor(rel, oth, c_rel, c_oth, t):
  """One OR / AND factor: (parent outputs u[n] (relevant minus other), child
  output w) and a function mapping their cotangents (gu, gw) to those of
  (rel, oth, c_rel, c_oth)."""
  n = rel.shape[0]
  d = rel - oth
  m1, a1, m2, w1, w2 = _top2(d)
  s_oth = oth.sum()
  u = np.empty(n)
  if t == 0:
    mx = np.maximum(oth, rel)
    p_oth = np.where(oth > rel, 1.0, np.where(oth == rel, 0.5, 0.0))
    s_mx = mx.sum()
    par_rel = s_mx + c_rel - mx
    opt1 = c_oth + s_oth - oth
    mz = np.where(np.arange(n) == a1, min(0.0, m2), min(0.0, m1))
    opt2 = par_rel + mz
    pm = [_max2(opt1[k], opt2[k]) for k in range(n)]
    par_oth = np.array([x[0] for x in pm])
    u[:] = par_rel - par_oth
    w = (s_mx + min(0.0, m1)) - s_oth
  else:
    lae = [_lae(rel[k], oth[k], t) for k in range(n)]
    lse = np.array([x[0] for x in lae])
    s_lse = lse.sum()
    s_lse_wo, s_oth_wo = s_lse - lse, s_oth - oth
    par_rel = c_rel + s_lse_wo
    cr0 = _lme(s_lse, s_oth, t)
    crs = _lae(s_oth + m1, s_oth + m2, t) if t < STABILITY_T else None
    cmx = _max2(cr0[0], crs[0]) if crs else (cr0[0], 1.0, 0.0)
    w = cmx[0] - s_oth
    opt1, opt2, opt3 = c_oth + s_oth_wo, c_rel + s_lse_wo, c_rel + s_oth_wo
    la12 = [_lae(opt1[k], opt2[k], t) for k in range(n)]
    po0 = [_lme(la12[k][0], opt3[k], t) for k in range(n)]
    xm = s_oth_wo + np.where(np.arange(n) == a1, m2, m1)
    lb = [_lae(opt1[k], c_rel + xm[k], t) for k in range(n)] if t < STABILITY_T else None
    pm = [_max2(po0[k][0], lb[k][0]) if lb else (po0[k][0], 1.0, 0.0) for k in range(n)]
    u[:] = par_rel - np.array([x[0] for x in pm])
  single = n == 1
  if single:
    u[:] = c_rel - c_oth

  def bwd(gu, gw):
    g_rel, g_oth = np.zeros(n), np.zeros(n)
    g_crel, g_coth, g_m1, g_m2 = 0.0, 0.0, 0.0, 0.0
    if t == 0:
      g_smx, g_soth = 0.0, 0.0
      g_mx = np.zeros(n)
      if single:
        g_crel += gu[0]
        g_coth -= gu[0]
      else:
        g_par_rel = gu.copy()
        g_par_oth = -gu
        g_opt1 = g_par_oth * np.array([x[1] for x in pm])
        g_opt2 = g_par_oth * np.array([x[2] for x in pm])
        g_par_rel += g_opt2
        g_mz = g_opt2
        g_coth += g_opt1.sum()
        g_soth += g_opt1.sum()
        g_oth -= g_opt1
        g_smx += g_par_rel.sum()
        g_crel += g_par_rel.sum()
        g_mx -= g_par_rel
        g_m1 += (g_mz.sum() - g_mz[a1]) * _dmin0(m1)
        g_m2 += g_mz[a1] * _dmin0(m2)
      g_soth -= gw
      g_smx += gw
      g_m1 += gw * _dmin0(m1)
      g_oth += g_soth
      g_mx += g_smx
      g_oth += g_mx * p_oth
      g_rel += g_mx * (1.0 - p_oth)
    else:
      g_slse, g_soth = 0.0, 0.0
      g_lse = np.zeros(n)
      if single:
        g_crel += gu[0]
        g_coth -= gu[0]
      else:
        g_par_rel = gu.copy()
        g_par_oth = -gu
        g_po0 = g_par_oth * np.array([x[1] for x in pm])
        g_lb = g_par_oth * np.array([x[2] for x in pm])
        g_la12 = g_po0 * np.array([x[1] for x in po0])
        g_opt3 = g_po0 * np.array([x[2] for x in po0])
        g_opt1 = g_la12 * np.array([x[1] for x in la12])
        g_opt2 = g_la12 * np.array([x[2] for x in la12])
        g_soth_wo = np.zeros(n)
        if lb:
          g_opt1 += g_lb * np.array([x[1] for x in lb])
          g_x = g_lb * np.array([x[2] for x in lb])
          g_crel += g_x.sum()
          g_soth_wo += g_x
          g_m1 += g_x.sum() - g_x[a1]
          g_m2 += g_x[a1]
        g_coth += g_opt1.sum()
        g_soth_wo += g_opt1 + g_opt3
        g_crel += g_opt2.sum() + g_opt3.sum() + g_par_rel.sum()
        g_slse_wo = g_opt2 + g_par_rel
        g_slse += g_slse_wo.sum()
        g_lse -= g_slse_wo
        g_soth += g_soth_wo.sum()
        g_oth -= g_soth_wo
      g_soth -= gw
      g_cr0, g_crs = gw * cmx[1], gw * cmx[2]
      g_slse += g_cr0 * cr0[1]
      g_soth += g_cr0 * cr0[2]
      if crs:
        g_m1 += g_crs * crs[1]
        g_m2 += g_crs * crs[2]
        g_soth += g_crs * (crs[1] + crs[2])
      g_lse += g_slse
      g_oth += g_soth
      g_rel += g_lse * np.array([x[1] for x in lae])
      g_oth += g_lse * np.array([x[2] for x in lae])
    g_d = g_m1 * w1 + g_m2 * w2
    g_rel += g_d
    g_oth -= g_d
    return g_rel, g_oth, g_crel, g_coth

  return u, w, bwd


def _lse(x, t):
  """T log sum exp(x / T) and its gradient (the softmax weights)."""
  m = x.max()
  e = np.exp((x - m) / t)
  return t * np.log(e.sum()) + m, e / e.sum()


def _dmin(a, b):
  """d min(a, b) / da, half at equality."""
  return 1.0 if a < b else (0.5 if a == b else 0.0)


def pool_factor(zero, one, i_zero, i_one, t):
  """One Pool factor (pgmax/factor/pool.py:328-474, normalised outputs):
  (choice outputs u[n], indicator output w) and the map from their
  cotangents to those of (zero, one, i_zero, i_one)."""
  n = zero.shape[0]
  d = one - zero
  pid = i_one - i_zero
  m1, a1, m2, w1, w2 = _top2(d)
  u = np.empty(n)
  if t == 0:
    mm = np.where(np.arange(n) == a1, m2, m1)
    u[:] = np.minimum(pid, -mm)
    w = m1
  else:
    w, sm = _lse(d, t)
    f_lse, fa, fb = _lae(w, -pid, t)
    lme = [_lme0(f_lse, d[k], t) for k in range(n)]
    u[:] = [-x[0] for x in lme]
    ind = d.copy()
    ind[a1] = -pid
    am, smi = _lse(ind, t)
    u[a1] = -am
  single = n == 1
  if single:
    u[:] = pid

  def bwd(gu, gw):
    g_d = np.zeros(n)
    g_pid = 0.0
    if t == 0:
      g_m1, g_m2 = gw, 0.0
      if single:
        g_pid += gu[0]
      else:
        for k in range(n):
          m = m2 if k == a1 else m1
          g_pid += gu[k] * _dmin(pid, -m)
          gm = -gu[k] * _dmin(-m, pid)
          if k == a1:
            g_m2 += gm
          else:
            g_m1 += gm
      g_d += g_m1 * w1 + g_m2 * w2
    else:
      g_w = gw
      if single:
        g_pid += gu[0]
      else:
        g_flse = 0.0
        for k in range(n):
          if k == a1:
            g_ind = -gu[k] * smi
            g_pid -= g_ind[a1]
            g_ind[a1] = 0.0
            g_d += g_ind
          else:
            g_flse -= gu[k] * lme[k][1]
            g_d[k] -= gu[k] * lme[k][2]
        g_w += g_flse * fa
        g_pid -= g_flse * fb
      g_d += g_w * sm
    return g_d, -g_d, g_pid, -g_pid  # (one, zero, indicator one, indicator zero)

  return u, w, bwd


def _lme0(a, b, t):
  """T log(e^(a/T) - e^(b/T)) if a >= b + 1e-30 (pool.py's default eps),
  else -inf; and its partials."""
  if not a >= b + 1e-30:
    return -np.inf, 0.0, 0.0
  q = np.exp((b - a) / t)
  return a + t * np.log1p(-q), 1.0 / (1.0 - q), -q / (1.0 - q)


def _logical_step(g, vt, t):
  """Every logical factor's outputs written into f (relevant slots; the
  other slots 0) and their pull-back functions."""
  f = np.zeros(g.num_es)
  fns = []
  for L in g.logical:
    for fi in range(L.ces.shape[0]):
      par = np.flatnonzero(L.pfi == fi)
      pr = L.m0 + L.pmi[par] + L.off
      po = L.m0 + L.pmi[par]
      cr, co = L.m0 + L.ces[fi] + L.off, L.m0 + L.ces[fi]
      u, w, bwd = logical_factor(vt[pr], vt[po], vt[cr], vt[co], t)
      f[pr] = u
      f[cr] = w
      fns.append((pr, po, cr, co, bwd))
  for P in getattr(g, "pool", ()):
    for fi in range(P.ies.shape[0]):
      ch = np.flatnonzero(P.cfi == fi)
      one, zero = P.m0 + P.cmi[ch] + 1, P.m0 + P.cmi[ch]
      i_one, i_zero = P.m0 + P.ies[fi] + 1, P.m0 + P.ies[fi]
      u, w, bwd = pool_factor(vt[zero], vt[one], vt[i_zero], vt[i_one], t)
      f[one] = u
      f[i_one] = w
      fns.append((one, zero, i_one, i_zero, bwd))
  return f, fns


def _step(g, m, lp_c, ev, d, t):
  """One iteration from messages m; returns the quantities the VJP needs."""
  vs = ev.copy()
  np.add.at(vs, g.vsi, m)
  vt = vs[g.vsi] - m
  s = lp_c.copy()
  np.add.at(s, g.row_cfg, vt[g.row_es])
  rows_s = s[g.row_cfg]
  mx = np.full(g.num_es, -np.inf)
  np.maximum.at(mx, g.row_es, rows_s)
  if t > 0:
    z = np.zeros(g.num_es)
    with np.errstate(invalid="ignore", over="ignore", divide="ignore"):
      np.add.at(z, g.row_es, np.exp((rows_s - mx[g.row_es]) / t))
      out = t * np.log(z) + mx
  else:
    z = None
    out = mx
  with np.errstate(invalid="ignore"):
    f = out - vt
  fns = []
  if g.logical or g.pool:
    f_log, fns = _logical_step(g, vt, t)
    f[g.num_enum_es:] = f_log[g.num_enum_es:]
  x = d * m + (1.0 - d) * f
  return vs, vt, s, rows_s, mx, z, out, x, fns


def forward(g, lp, ev, ftov_in, num_iters, damping, temperature):
  """(final messages, tape of messages m_0 .. m_n), float64."""
  lp_c = np.clip(np.asarray(lp, np.float64), -LP_MAX, LP_MAX)
  ev = np.asarray(ev, np.float64)
  m = _normalize(np.asarray(ftov_in, np.float64), g.edge, g.n_edges)[0]
  tape = [m]
  for _ in range(max(int(num_iters), 1)):
    x = _step(g, m, lp_c, ev, damping, temperature)[-2]
    m = _normalize(x, g.edge, g.n_edges)[0]
    tape.append(m)
  return m, tape


def vjp(g, lp, ev, ftov_in, num_iters, damping, temperature, g_out):
  """(g_lp, g_ev, g_ftov_in): the cotangent g_out of the final messages
  pulled back through the run."""
  lp = np.asarray(lp, np.float64)
  lp_c = np.clip(lp, -LP_MAX, LP_MAX)
  ev = np.asarray(ev, np.float64)
  d, t = float(damping), float(temperature)
  _, tape = forward(g, lp, ev, ftov_in, num_iters, d, t)
  g_lp = np.zeros(g.num_cfg)
  g_ev = np.zeros(g.num_vs)
  gm_next = np.asarray(g_out, np.float64)
  for it in range(len(tape) - 2, -1, -1):
    m = tape[it]
    vs, vt, s, rows_s, mx, z, out, x, fns = _step(g, m, lp_c, ev, d, t)
    gx = _normalize_bwd(x, gm_next, g.edge, g.n_edges)
    gm = d * gx
    gf = (1.0 - d) * gx
    # Enum: f = out - vt, out_k = LSE over the configs of edge state k
    gvt = np.zeros(g.num_es)
    gvt[:g.num_enum_es] = -gf[:g.num_enum_es]
    # OR / AND / Pool: f = the closed forms of vt (other slots constant 0)
    for pr, po, cr, co, bwd in fns:
      g_rel, g_oth, g_crel, g_coth = bwd(gf[pr], gf[cr])
      np.add.at(gvt, pr, g_rel)
      np.add.at(gvt, po, g_oth)
      gvt[cr] += g_crel
      gvt[co] += g_coth
    if t > 0:
      with np.errstate(invalid="ignore", over="ignore"):
        w = np.exp((rows_s - out[g.row_es]) / t)
      w = np.where(np.isfinite(out[g.row_es]), w, 0.0)
    else:
      hit = (rows_s == mx[g.row_es]).astype(np.float64)
      cnt = np.zeros(g.num_es)
      np.add.at(cnt, g.row_es, hit)
      w = hit / np.maximum(cnt[g.row_es], 1.0)
    gs = np.zeros(g.num_cfg)
    np.add.at(gs, g.row_cfg, gf[g.row_es] * w)
    clip_w = np.where(np.abs(lp) < LP_MAX, 1.0, np.where(np.abs(lp) == LP_MAX, 0.5, 0.0))
    g_lp += gs * clip_w
    np.add.at(gvt, g.row_es, gs[g.row_cfg])
    # vt = vs[var_state] - m; vs = ev + segment_sum(m)
    gvs = var_sum_pullback(g, gvt)
    g_ev += gvs
    gm = gm - gvt + gvs[g.vsi]
    gm_next = gm
  g_in = _normalize_bwd(np.asarray(ftov_in, np.float64), gm_next, g.edge, g.n_edges)
  return g_lp, g_ev, g_in
