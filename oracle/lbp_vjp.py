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
  """The arrays of an Enum-only graph: var_state[E_s], edge[E_s], and the
  rows (config, edge state); configs number the log-potentials."""

  def __init__(self, var_state, edge, row_cfg, row_es, num_vs, num_cfg):
    self.vsi = np.asarray(var_state, np.int64)
    self.edge = np.asarray(edge, np.int64)
    self.n_edges = int(self.edge.max()) + 1 if self.edge.size else 0
    self.row_cfg = np.asarray(row_cfg, np.int64)
    self.row_es = np.asarray(row_es, np.int64)
    self.num_vs = int(num_vs)
    self.num_cfg = int(num_cfg)
    self.num_es = self.vsi.shape[0]


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
    with np.errstate(invalid="ignore", over="ignore"):
      np.add.at(z, g.row_es, np.exp((rows_s - mx[g.row_es]) / t))
      out = t * np.log(z) + mx
  else:
    z = None
    out = mx
  f = out - vt
  x = d * m + (1.0 - d) * f
  return vs, vt, s, rows_s, mx, z, out, x


def forward(g, lp, ev, ftov_in, num_iters, damping, temperature):
  """(final messages, tape of messages m_0 .. m_n), float64."""
  lp_c = np.clip(np.asarray(lp, np.float64), -LP_MAX, LP_MAX)
  ev = np.asarray(ev, np.float64)
  m = _normalize(np.asarray(ftov_in, np.float64), g.edge, g.n_edges)[0]
  tape = [m]
  for _ in range(max(int(num_iters), 1)):
    x = _step(g, m, lp_c, ev, damping, temperature)[-1]
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
    vs, vt, s, rows_s, mx, z, out, x = _step(g, m, lp_c, ev, d, t)
    gx = _normalize_bwd(x, gm_next, g.edge, g.n_edges)
    gm = d * gx
    gf = (1.0 - d) * gx
    # f = out - vt; out_k = LSE over the configs of edge state k
    gvt = -gf
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
