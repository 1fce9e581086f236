"""Seeded synthetic workloads: the five benchmark configurations as flat wirings.

Each builder returns the compiled wiring of a factor graph (the flat arrays
``FactorGraph`` + ``InfererContext`` produce, pgmax/infer/inferer.py:65-118)
together with its float inputs, built with vectorised NumPy so that graphs of
hundreds of millions of edge states are ready in seconds:

* ``potts_batch``: 2,048 open 4-neighbour Potts grids of uneven side and
  state count in one disjoint union (one FactorGraph holding every grid as
  its own variable group and PairwiseFactorGroup);
* ``spin_glass_batch``: toroidal binary spin glasses (one shared wiring,
  a batch of coupling and field draws);
* ``deconvolution_wiring`` / ``deconvolution_problem`` /
  ``deconvolution_evidence``: the OR/AND binary deconvolution of
  examples/pmp_binary_deconvolution.ipynb (cells 12-14) of the reference.
"""

import dataclasses
from typing import List

import numpy as np

from pgmax_amd import batch


@dataclasses.dataclass
class PottsBatch:
  """A batch of open Potts grids: per-graph specs, wirings and union inputs."""

  sides: np.ndarray       # [G] grid side
  states: np.ndarray      # [G] states per variable
  betas: np.ndarray       # [G] diagonal coupling
  wirings: List[batch.FlatWiring]
  log_potentials: np.ndarray  # union layout (graph by graph), float32
  evidence: np.ndarray        # union layout (graph by graph), float32

  @property
  def num_graphs(self) -> int:
    return int(self.sides.shape[0])


def potts_batch(num_graphs: int = 2048, seed: int = 0, side_range=(16, 128),
                state_range=(2, 8), beta_range=(0.2, 1.5)) -> PottsBatch:
  """Uneven open Potts grids, 4-neighbour pairwise factors beta * I_S.

  Draws from ``np.random.default_rng(seed)``, in this order: every graph's
  side ~ U{side_range}, then every graph's state count ~ U{state_range},
  then every graph's beta ~ U(beta_range), then each graph's unaries ~ N(0, 1)
  (one value per variable state, graph by graph). Factors follow
  ``batch.grid_pairs`` (for each (i, j): the pair with (i+1, j), then the
  one with (i, j+1)), configs in row-major order
  (pgmax/fgroup/enum.py:303-309). Every graph is Enum-only, so the union's
  log-potentials and evidence are the graphs' arrays in graph order.
  """
  rng = np.random.default_rng(seed)
  sides = rng.integers(side_range[0], side_range[1] + 1, size=num_graphs)
  states = rng.integers(state_range[0], state_range[1] + 1, size=num_graphs)
  betas = rng.uniform(beta_range[0], beta_range[1], size=num_graphs)
  wirings, lps, evs = [], [], []
  for side, s, beta in zip(sides, states, betas):
    side, s = int(side), int(s)
    w = batch.grid_wiring(side, side, s, torus=False, full_tables=True)
    nf = w.num_log_potentials // (s * s)
    table = (np.float32(beta) * np.eye(s, dtype=np.float32)).reshape(-1)
    wirings.append(w)
    lps.append(np.tile(table, nf))
    evs.append(rng.normal(size=side * side * s).astype(np.float32))
  return PottsBatch(sides=sides, states=states, betas=betas, wirings=wirings,
                    log_potentials=np.concatenate(lps), evidence=np.concatenate(evs))


def spin_glass_batch(side: int = 128, batch_size: int = 1024, seed: int = 0,
                     coupling_std: float = 1.0, unary_std: float = 0.5):
  """Toroidal binary spin glasses sharing one wiring.

  Returns (wiring, log_potentials [B, 4F], evidence [B, 2 side^2]) with
  couplings J ~ N(0, coupling_std) per factor and table [[J, -J], [-J, J]],
  fields ~ N(0, unary_std) per variable state, drawn from
  ``np.random.default_rng(seed)`` (all couplings, then all fields).
  """
  rng = np.random.default_rng(seed)
  w = batch.grid_wiring(side, side, 2, torus=True, full_tables=True)
  nf = w.num_log_potentials // 4
  j = rng.normal(0.0, coupling_std, size=(batch_size, nf)).astype(np.float32)
  lp = np.stack([j, -j, -j, j], -1).reshape(batch_size, -1)
  ev = rng.normal(0.0, unary_std, size=(batch_size, w.num_var_states)).astype(np.float32)
  return w, lp, ev


# ---------------------------------------------------------------------------
# Binary deconvolution (examples/pmp_binary_deconvolution.ipynb cells 12-14 of
# the reference): AND(S, W, SW) per (image, channel, shift, feature, offset)
# and OR(SW..., X) per image pixel, built directly as flat wiring arrays.
# ---------------------------------------------------------------------------


def _deconvolution_factors(n_images, n_chan, im_h, im_w, n_feat, feat_h, feat_w):
  """Variables and factors of the deconvolution graph as global variable
  numbers (groups S, W, SW, X in that order): the AND factors' (S, W, SW)
  triples in the notebook's loop order (img, chan, sh, sw, feat, fh, fw), and
  per X pixel, in order of first appearance, the OR factor's SW parents in
  loop order."""
  s_h, s_w = im_h - feat_h + 1, im_w - feat_w + 1
  shapes = {
      "S": (n_images, n_feat, s_h, s_w),
      "W": (n_chan, n_feat, feat_h, feat_w),
      "SW": (n_images, n_chan, im_h, im_w, n_feat, feat_h, feat_w),
      "X": (n_images, n_chan, im_h, im_w),
  }
  off, o = {}, 0
  for k in ("S", "W", "SW", "X"):
    off[k] = o
    o += int(np.prod(shapes[k]))
  nv = o
  i, c, a, b, f, p, q = np.meshgrid(np.arange(n_images), np.arange(n_chan), np.arange(s_h),
                                    np.arange(s_w), np.arange(n_feat), np.arange(feat_h),
                                    np.arange(feat_w), indexing="ij")
  i, c, a, b, f, p, q = (t.reshape(-1).astype(np.int64) for t in (i, c, a, b, f, p, q))
  h, w = a + p, b + q
  s_var = off["S"] + np.ravel_multi_index((i, f, a, b), shapes["S"])
  w_var = off["W"] + np.ravel_multi_index((c, f, p, q), shapes["W"])
  sw_var = off["SW"] + np.ravel_multi_index((i, c, h, w, f, p, q), shapes["SW"])
  x_var = off["X"] + np.ravel_multi_index((i, c, h, w), shapes["X"])
  uniq, first = np.unique(x_var, return_index=True)
  order_x = uniq[np.argsort(first, kind="stable")]
  rank = np.empty(nv, np.int64)
  rank[order_x] = np.arange(order_x.size)
  par_order = np.argsort(rank[x_var], kind="stable")
  return {
      "shapes": shapes, "offsets": off, "num_vars": nv,
      "and": np.stack([s_var, w_var, sw_var], 1),
      "or_children": order_x,
      "or_parents": sw_var[par_order],
      "or_counts": np.bincount(rank[x_var], minlength=order_x.size),
  }


def deconvolution_factor_graph(n_images, n_chan, im_h, im_w, n_feat, feat_h, feat_w):
  """The deconvolution factor graph built through ``FactorGraph``.

  Variable groups [S, W, SW, X] (NDVarArrays of binary variables), one
  ``ANDFactorGroup`` given as an (n_and, 3, 2) array and one
  ``ORFactorGroup`` given as ``RaggedVariables``, in the factor order of the
  notebook's loops (examples/pmp_binary_deconvolution.ipynb cells 12-14 of
  the reference). Returns (FactorGraph, dict of variable groups).
  """
  from pgmax_amd import fgraph, fgroup, vgroup  # pylint: disable=g-import-not-at-top
  d = _deconvolution_factors(n_images, n_chan, im_h, im_w, n_feat, feat_h, feat_w)
  groups = {k: vgroup.NDVarArray(num_states=2, shape=d["shapes"][k])
            for k in ("S", "W", "SW", "X")}
  base = np.zeros(d["num_vars"], np.int64)
  for k, g in groups.items():
    n = int(np.prod(d["shapes"][k]))
    base[d["offsets"][k]:d["offsets"][k] + n] = hash(g) - d["offsets"][k]

  def pairs(var):
    return np.stack([base[var] + var, np.full(var.shape, 2, np.int64)], axis=-1)

  fg = fgraph.FactorGraph(variable_groups=[groups[k] for k in ("S", "W", "SW", "X")])
  fg.add_factors(fgroup.ANDFactorGroup(variables_for_factors=pairs(d["and"])))
  counts = d["or_counts"]
  n_or = counts.shape[0]
  child_pos = np.cumsum(counts + 1) - 1
  or_vars = np.empty(int((counts + 1).sum()), np.int64)
  mask = np.ones(or_vars.shape[0], bool)
  mask[child_pos] = False
  or_vars[mask] = d["or_parents"]
  or_vars[child_pos] = d["or_children"]
  fg.add_factors(fgroup.ORFactorGroup(variables_for_factors=fgroup.RaggedVariables(
      pairs=pairs(or_vars), sizes=counts + 1)))
  assert n_or == d["or_children"].shape[0]
  return fg, groups


def deconvolution_wiring(n_images, n_chan, im_h, im_w, n_feat, feat_h, feat_w):
  """Flat wiring of the deconvolution factor graph, in the layout FactorGraph
  gives variable groups [S, W, SW, X] plus an ANDFactorGroup and an
  ORFactorGroup built in the notebook's loop order. Returns (FlatWiring,
  dict of variable-group offsets and shapes)."""
  d = _deconvolution_factors(n_images, n_chan, im_h, im_w, n_feat, feat_h, feat_w)
  off, shapes, nv = d["offsets"], d["shapes"], d["num_vars"]
  s_var, w_var, sw_var = d["and"][:, 0], d["and"][:, 1], d["and"][:, 2]
  n_and = s_var.size
  order_x, or_parents, or_counts = d["or_children"], d["or_parents"], d["or_counts"]
  n_or = order_x.size
  # edges: OR factors (parents then child), then AND factors (S, W, SW)
  or_starts = np.concatenate([[0], np.cumsum(or_counts + 1)])[:-1]
  or_edges = np.empty(int((or_counts + 1).sum()), np.int64)
  child_pos = or_starts + or_counts
  mask = np.ones(or_edges.size, bool)
  mask[child_pos] = False
  or_edges[mask] = or_parents
  or_edges[child_pos] = order_x
  or_fac = np.repeat(np.arange(n_or), or_counts + 1)
  and_edges = np.stack([s_var, w_var, sw_var], 1).reshape(-1)
  and_fac = n_or + np.repeat(np.arange(n_and), 3)
  edge_var = np.concatenate([or_edges, and_edges])
  edge_fac = np.concatenate([or_fac, and_fac])
  E = edge_var.size
  es = np.arange(2 * E, dtype=np.int64)
  vsf = 2 * edge_var[es // 2] + es % 2
  n_es_or = 2 * or_edges.size
  # logical wiring, indices relative to each type's message range
  or_pfi = np.repeat(np.arange(n_or), or_counts)
  or_pmi = 2 * np.flatnonzero(mask)
  or_ces = 2 * child_pos
  and_pfi = np.repeat(np.arange(n_and), 2)
  and_edge_local = np.arange(3 * n_and).reshape(n_and, 3)
  and_pmi = 2 * and_edge_local[:, :2].reshape(-1) + 1
  and_ces = 2 * and_edge_local[:, 2] + 1
  n_es = 2 * E
  flat = batch.FlatWiring(
      num_var_states=2 * nv,
      num_log_potentials=0,
      var_states_for_edge_states=vsf,
      edge_indices_for_edge_states=es // 2,
      factor_indices_for_edge_states=edge_fac[es // 2],
      evidence_to_vars=np.repeat(np.arange(nv, dtype=np.int64), 2),
      msgs_range=[(0, 0), (0, n_es_or), (n_es_or, n_es), (n_es, n_es)],
      potentials_range=[(0, 0)] * 4,
      enum_ci=np.zeros(0, np.int64),
      enum_ce=np.zeros(0, np.int64),
      logical={batch.OR: (or_pfi, or_pmi, or_ces, 1),
               batch.AND: (and_pfi, and_pmi, and_ces, -1),
               batch.POOL: (np.zeros(0, np.int64),) * 3 + (1,)},
  )
  return flat, {"offsets": off, "shapes": shapes}


def deconvolution_problem(n_images, im_h, im_w, n_feat_true=4, feat_true=5, p_w=0.2,
                          p_s=0.05, seed=0):
  """Synthetic binary images X = OR-convolution of Bernoulli(p_s) activations
  with Bernoulli(p_w) features (one channel)."""
  rng = np.random.default_rng(seed)
  W = rng.random((n_feat_true, feat_true, feat_true)) < p_w
  sh, sw = im_h - feat_true + 1, im_w - feat_true + 1
  S = rng.random((n_images, n_feat_true, sh, sw)) < p_s
  X = np.zeros((n_images, 1, im_h, im_w), bool)
  for f in range(n_feat_true):
    for p in range(feat_true):
      for q in range(feat_true):
        if W[f, p, q]:
          X[:, 0, p:p + sh, q:q + sw] |= S[:, f]
  return X


def deconvolution_evidence(meta, X, prior_logit=-2.0, x_logit=-230.0):
  """Evidence: log-odds prior_logit on S and W, 0 on SW, and the observed
  image on X with strength |x_logit| (the notebook's uS, uW, uX)."""
  off, shapes = meta["offsets"], meta["shapes"]
  nv = off["X"] + int(np.prod(shapes["X"]))
  ev = np.zeros((nv, 2), np.float32)
  nS, nW = int(np.prod(shapes["S"])), int(np.prod(shapes["W"]))
  ev[off["S"]:off["S"] + nS, 1] = prior_logit
  ev[off["W"]:off["W"] + nW, 1] = prior_logit
  ev[off["X"]:, 0] = (2 * X.reshape(-1).astype(np.float32) - 1) * x_logit
  return ev.reshape(-1)
