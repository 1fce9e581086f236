"""FP64 oracle of the warm-start message-passing network -- TEST INFRASTRUCTURE ONLY.

Only ``tests/`` and ``bench.py``'s checks may import this module, as the
checker of the GPU kernels in pgmax_amd/csrc/pgx_mpnn.hip. The network has
no counterpart in the reference: it predicts initial factor-to-variable
messages (the ``ftov_msgs`` entry of BPArrays, pgmax/infer/bp_state.py:172-255
of the reference) for graphs of binary variables and pairwise factors.

Model, per edge e = (factor f, variable v), with u the factor's other
variable and o(e) the factor's other edge:
  x_e   = [ev_v, ev_u, J, h_v, h_u, 1, 0, 0]       (8 input features)
          ev_* = evidence[*, 1] - evidence[*, 0]; with the factor's
          log-potential table t oriented as t[state of v, state of u]:
          J = (t00 + t11 - t01 - t10) / 4, h_v = (t10 + t11 - t00 - t01) / 2,
          h_u = (t01 + t11 - t00 - t10) / 2
  h0_e  = relu(x_e W0 + b0)                                   (64)
  layer l = 1..3:
    z_e = [h_e, sum_{e' at v} h_e' - h_e, h_o(e)]             (192)
    h_e = relu(z_e W_l + b_l)                                 (64)
  y_e   = h3_e . w_h + b_h
The predicted message of edge e is [min(0, -y_e), min(0, y_e)] (a normalised
log-message whose log-ratio is y_e). Training minimises the mean squared
error of y_e against the log-ratio of converged BP messages.

The oracle's backward pass is pinned by central finite differences in FP64
(tests/test_mpnn.py).
"""

from typing import Dict, Tuple

import numpy as np

HIDDEN = 64
IN_FEATURES = 8
LAYERS = 3


def init_params(seed: int = 0) -> Dict[str, np.ndarray]:
  """He-uniform weights, zero biases (FP64)."""
  rng = np.random.default_rng(seed)
  p = {}
  def lin(fan_in, fan_out):
    lim = np.sqrt(6.0 / fan_in)
    return rng.uniform(-lim, lim, size=(fan_in, fan_out))
  p["W0"] = lin(IN_FEATURES, HIDDEN)
  p["b0"] = np.zeros(HIDDEN)
  for l in range(1, LAYERS + 1):
    p[f"W{l}"] = lin(3 * HIDDEN, HIDDEN) / np.sqrt(3.0)
    p[f"b{l}"] = np.zeros(HIDDEN)
  p["wh"] = rng.uniform(-1, 1, size=HIDDEN) * np.sqrt(3.0 / HIDDEN)
  p["bh"] = np.zeros(1)
  return p


def pairwise_binary_graph(var_states_for_edge_states, evidence_to_vars, lp_offset=0):
  """Edge arrays of a graph of pairwise factors over binary variables.

  Factor f owns edges 2f and 2f+1 and edge states 4f..4f+3; its four
  log-potentials start at lp_offset + 4f in row-major order (the first
  variable's state major), as pgmax/fgroup/enum.py:303-309 lays them out.
  Returns (edge_var [E], edge_pos [E] in {0, 1}, lp0 [E], num_vars).
  """
  vsf = np.asarray(var_states_for_edge_states)
  e2v = np.asarray(evidence_to_vars)
  E = vsf.shape[0] // 2
  edge_var = e2v[vsf[0::2]]
  edge_pos = np.arange(E) % 2
  lp0 = lp_offset + 4 * (np.arange(E) // 2)
  return edge_var, edge_pos, lp0, int(e2v[-1]) + 1


def features(edge_var, edge_pos, lp0, log_potentials, evidence):
  """Input features x [E, 8] (see the module docstring)."""
  lp = np.asarray(log_potentials, dtype=np.float64)
  ev = np.asarray(evidence, dtype=np.float64).reshape(-1, 2)
  E = edge_var.shape[0]
  other = np.arange(E) ^ 1
  t = np.stack([lp[lp0], lp[lp0 + 1], lp[lp0 + 2], lp[lp0 + 3]], 1)  # t[a*2 + b], a = first var
  t00, t01, t10, t11 = t[:, 0], t[:, 1], t[:, 2], t[:, 3]
  second = edge_pos == 1
  # orient: rows indexed by this edge's variable
  t01o = np.where(second, t10, t01)
  t10o = np.where(second, t01, t10)
  dv = ev[edge_var, 1] - ev[edge_var, 0]
  du = dv[other]
  x = np.zeros((E, IN_FEATURES))
  x[:, 0] = dv
  x[:, 1] = du
  x[:, 2] = (t00 + t11 - t01o - t10o) / 4
  x[:, 3] = (t10o + t11 - t00 - t01o) / 2
  x[:, 4] = (t01o + t11 - t00 - t10o) / 2
  x[:, 5] = 1.0
  return x


def _relu(a):
  return np.maximum(a, 0.0)


def bf16(a):
  """Round to the nearest BF16 value (ties to even), returned as float64."""
  f = np.asarray(a, dtype=np.float32)
  u = f.view(np.uint32).astype(np.uint64)
  u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
  return u.astype(np.uint32).view(np.float32).astype(np.float64)


def forward(p, x, edge_var, num_vars, emulate_bf16=False):
  """Returns (y [E], cache).

  With emulate_bf16 the activations, the layer inputs and the 192x64 layer
  weights are rounded to BF16 where the GPU kernels store them in BF16; all
  arithmetic stays FP64.
  """
  E = x.shape[0]
  other = np.arange(E) ^ 1
  rnd = bf16 if emulate_bf16 else (lambda a: a)
  pre = x @ p["W0"] + p["b0"]
  h = rnd(_relu(pre))
  cache = {"x": x, "h": [h], "z": []}
  for l in range(1, LAYERS + 1):
    sv = np.zeros((num_vars, HIDDEN))
    np.add.at(sv, edge_var, h)
    z = np.concatenate([h, rnd(sv[edge_var] - h), h[other]], 1)
    h = rnd(_relu(z @ rnd(p[f"W{l}"]) + p[f"b{l}"]))
    cache["z"].append(z)
    cache["h"].append(h)
  y = h @ p["wh"] + p["bh"][0]
  return y, cache


def loss(y, target):
  return float(np.mean((y - target) ** 2))


def backward(p, cache, y, target, edge_var, num_vars, emulate_bf16=False):
  """Gradients of the mean squared error with respect to every parameter.

  With emulate_bf16 the layers' pre-activation gradients and weights are
  rounded to BF16 where the GPU kernels feed them to the matrix cores.
  """
  rnd = bf16 if emulate_bf16 else (lambda a: a)
  E = y.shape[0]
  other = np.arange(E) ^ 1
  dy = 2.0 * (y - target) / E
  g = {}
  h3 = cache["h"][LAYERS]
  g["wh"] = h3.T @ dy
  g["bh"] = np.array([dy.sum()])
  dh = np.outer(dy, p["wh"])
  for l in range(LAYERS, 0, -1):
    hl = cache["h"][l]
    dpre = rnd(dh * (hl > 0))
    z = cache["z"][l - 1]
    g[f"W{l}"] = z.T @ dpre
    g[f"b{l}"] = dpre.sum(0)
    dz = dpre @ rnd(p[f"W{l}"]).T
    d_self, d_var, d_fac = dz[:, :HIDDEN], dz[:, HIDDEN:2 * HIDDEN], dz[:, 2 * HIDDEN:]
    tv = np.zeros((num_vars, HIDDEN))
    np.add.at(tv, edge_var, d_var)
    dh = d_self + tv[edge_var] - d_var + d_fac[other]
  h0 = cache["h"][0]
  dpre0 = dh * (h0 > 0)
  g["W0"] = cache["x"].T @ dpre0
  g["b0"] = dpre0.sum(0)
  return g


def adam_step(p, g, m, v, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
  """One Adam update (Kingma & Ba 2015, Algorithm 1) of every parameter in place."""
  for k in p:
    m[k] = b1 * m[k] + (1 - b1) * g[k]
    v[k] = b2 * v[k] + (1 - b2) * g[k] * g[k]
    mh = m[k] / (1 - b1 ** step)
    vh = v[k] / (1 - b2 ** step)
    p[k] = p[k] - lr * mh / (np.sqrt(vh) + eps)


_ORDER = ["W0", "b0", "W1", "b1", "W2", "b2", "W3", "b3", "wh", "bh"]
_SHAPES = {"W0": (IN_FEATURES, HIDDEN), "b0": (HIDDEN,), "wh": (HIDDEN,), "bh": (1,)}
for _l in range(1, LAYERS + 1):
  _SHAPES[f"W{_l}"] = (3 * HIDDEN, HIDDEN)
  _SHAPES[f"b{_l}"] = (HIDDEN,)


def to_flat(p) -> np.ndarray:
  """The flat parameter layout of include/pgx.h (PGX_MPNN_NUM_PARAMS floats)."""
  return np.concatenate([np.asarray(p[k], np.float64).reshape(-1) for k in _ORDER])


def from_flat(flat) -> Dict[str, np.ndarray]:
  flat = np.asarray(flat, dtype=np.float64)
  out, o = {}, 0
  for k in _ORDER:
    n = int(np.prod(_SHAPES[k]))
    out[k] = flat[o:o + n].reshape(_SHAPES[k]).copy()
    o += n
  assert o == flat.size, (o, flat.size)
  return out


def messages_from_ratio(y):
  """Normalised binary log-messages with log-ratio y: [min(0, -y), min(0, y)]."""
  y = np.asarray(y, dtype=np.float64)
  return np.stack([np.minimum(0.0, -y), np.minimum(0.0, y)], 1).reshape(-1)


def ratio_from_messages(ftov):
  m = np.asarray(ftov, dtype=np.float64).reshape(-1, 2)
  return m[:, 1] - m[:, 0]
