"""Reverse-mode BP (pgx_run_vjp, infer.vjp) against the FP64 reverse-mode
oracle, which tests/test_oracle.py pins by central finite differences.

The GPU product is FP32 along the engine's FP32 trajectory, the oracle FP64
along its own, so they agree to a relative tolerance, not bit for bit; the
bars are ten times the measured differences.
"""

import numpy as np
import pytest

from oracle import lbp
from oracle import lbp_vjp
from pgmax_amd import fgraph
from pgmax_amd import fgroup
from pgmax_amd import infer
from pgmax_amd import vgroup
from pgmax_amd.infer import vjp as vjp_mod

pytestmark = pytest.mark.gpu


def _np(x):
  return x.detach().cpu().numpy() if hasattr(x, "detach") else np.asarray(x)


def _graph(seed):
  rng = np.random.default_rng(seed)
  v = vgroup.NDVarArray(num_states=3, shape=(6,))
  fg = fgraph.FactorGraph(v)
  fg.add_factors(fgroup.PairwiseFactorGroup(
      [[v[0], v[1]], [v[1], v[2]], [v[2], v[3]], [v[3], v[0]], [v[0], v[2]], [v[4], v[5]]],
      rng.normal(size=(6, 3, 3))))
  cfg = np.array([[a, b, c] for a in range(3) for b in range(3) for c in range(3)
                  if (a + b + c) % 2 == 0])
  fg.add_factors(fgroup.EnumFactorGroup([[v[1], v[3], v[4]]], cfg,
                                        log_potentials=rng.normal(size=(1, cfg.shape[0]))))
  bp = infer.BP(fg.bp_state, temperature=1.0)
  g = lbp.graph_from_context(bp.context)
  args = g["args"]["enum"]
  eg = lbp_vjp.EnumGraph(g["var_states_for_edge_states"], g["edge_indices_for_edge_states"],
                         args["factor_configs_indices"], args["factor_configs_edge_states"],
                         bp.context.bp_state.fg_state.num_var_states,
                         int(np.asarray(args["factor_configs_indices"]).max()) + 1)
  lp = np.asarray(fg.bp_state.log_potentials.value, np.float32)
  return bp, eg, lp, rng


def _pair_graph(seed, num_states):
  """Pairwise factors only (the reverse pass's register path): 7 variables
  of num_states states, cycles and a variable in four factors."""
  rng = np.random.default_rng(seed)
  v = vgroup.NDVarArray(num_states=num_states, shape=(7,))
  fg = fgraph.FactorGraph(v)
  pairs = [[0, 1], [1, 2], [2, 3], [3, 0], [0, 2], [4, 5], [5, 6], [6, 4], [0, 5]]
  fg.add_factors(fgroup.PairwiseFactorGroup(
      [[v[a], v[b]] for a, b in pairs], rng.normal(size=(len(pairs), num_states, num_states))))
  bp = infer.BP(fg.bp_state, temperature=1.0)
  g = lbp.graph_from_context(bp.context)
  args = g["args"]["enum"]
  eg = lbp_vjp.EnumGraph(g["var_states_for_edge_states"], g["edge_indices_for_edge_states"],
                         args["factor_configs_indices"], args["factor_configs_edge_states"],
                         bp.context.bp_state.fg_state.num_var_states,
                         int(np.asarray(args["factor_configs_indices"]).max()) + 1)
  lp = np.asarray(fg.bp_state.log_potentials.value, np.float32)
  return bp, eg, lp, rng


def _rel_err(got, want):
  return float(np.max(np.abs(got - want)) / max(np.max(np.abs(want)), 1e-12))


def _ft_err(got, want, want_ev):
  """Error of the initial messages' cotangent, relative to the larger of its
  own size and the evidence's. Without damping the output forgets the
  initial messages within a few iterations, so that cotangent is a small
  difference (gvs - gvt, the pull-back of vt = vs[var] - m) of terms the
  size of the evidence's cotangent, and FP32 rounds it at their scale."""
  scale = max(np.max(np.abs(want)), np.max(np.abs(want_ev)), 1e-12)
  return float(np.max(np.abs(got - want)) / scale)


@pytest.mark.parametrize("temperature,damping", [(1.0, 0.5), (1.0, 0.0), (0.4, 0.3), (0.0, 0.5)])
def test_run_vjp_matches_oracle(gpu, temperature, damping):
  """Batched inputs (one set per element) and shared ones (stride 0, whose
  gradients sum over the batch)."""
  bp, eg, lp, rng = _graph(5)
  n, nb = 7, 3
  lpb = (lp[None] + rng.normal(0, 0.3, size=(nb, lp.size))).astype(np.float32)
  ev = rng.normal(size=(nb, eg.num_vs)).astype(np.float32)
  ft = rng.normal(size=(nb, eg.num_es)).astype(np.float32)
  w = rng.normal(size=(nb, eg.num_es)).astype(np.float32)
  run_vjp = vjp_mod.RunVJP(bp.context)
  g_lp, g_ev, g_ft = (_np(x) for x in run_vjp(lpb, ev, ft, n, damping, temperature, w))
  errs = []
  for b in range(nb):
    r_lp, r_ev, r_ft = lbp_vjp.vjp(eg, lpb[b], ev[b], ft[b], n, damping, temperature, w[b])
    errs += [_rel_err(g_lp[b], r_lp), _rel_err(g_ev[b], r_ev), _rel_err(g_ft[b], r_ft)]
  # shared log-potentials and initial messages, batched evidence
  s_lp, s_ev, s_ft = (_np(x) for x in run_vjp(lp, ev, ft[0], n, damping, temperature, w))
  assert s_lp.shape == lp.shape and s_ev.shape == ev.shape and s_ft.shape == ft[0].shape
  refs = [lbp_vjp.vjp(eg, lp, ev[b], ft[0], n, damping, temperature, w[b]) for b in range(nb)]
  errs += [_rel_err(s_lp, sum(r[0] for r in refs)), _rel_err(s_ft, sum(r[2] for r in refs))]
  errs += [_rel_err(s_ev[b], refs[b][1]) for b in range(nb)]
  print(f"T={temperature} d={damping}: max relative error {max(errs):.2e}")
  assert max(errs) < 3e-5  # measured <= 2.4e-6
  run_vjp.close()


@pytest.mark.parametrize("num_states,temperature,damping",
                         [(2, 1.0, 0.5), (2, 0.0, 0.5), (3, 0.4, 0.0), (8, 1.0, 0.3),
                          (8, 0.0, 0.5)])
def test_pairwise_register_path_matches_oracle(gpu, num_states, temperature, damping):
  """Graphs of pairwise factors take the one-thread-per-factor path (forward
  replay and reverse step in registers); batched and shared inputs against
  the FP64 oracle. Element 1 clips log-potentials of its first factor:
  entries beyond the bound pass no cotangent, and the bound itself (+1e6,
  the maximum of its row and column) passes half of its cotangent. Only
  the first is compared everywhere: the FP32 trajectory rounds messages
  near 1e6 by ~0.06, which moves softmax weights (T > 0) and, with 8
  states, can reorder close maxima (T = 0); with 2 states at T = 0 the
  comparison is exact up to rounding."""
  bp, eg, lp, rng = _pair_graph(11 + num_states, num_states)
  n, nb, sq = 9, 3, num_states * num_states
  lpb = (lp[None] + rng.normal(0, 0.3, size=(nb, lp.size))).astype(np.float32)
  beyond = [num_states + 1] + ([2 * num_states + 2] if num_states > 2 else [])
  lpb[1, 0] = 1e6
  lpb[1, beyond] = [-3e6, 2e6][:len(beyond)]
  ev = rng.normal(size=(nb, eg.num_vs)).astype(np.float32)
  ft = rng.normal(size=(nb, eg.num_es)).astype(np.float32)
  w = rng.normal(size=(nb, eg.num_es)).astype(np.float32)
  run_vjp = vjp_mod.RunVJP(bp.context)
  g_lp, g_ev, g_ft = (_np(x) for x in run_vjp(lpb, ev, ft, n, damping, temperature, w))
  errs = []
  for b in range(nb):
    r_lp, r_ev, r_ft = lbp_vjp.vjp(eg, lpb[b], ev[b], ft[b], n, damping, temperature, w[b])
    if b == 1:
      assert np.all(g_lp[b][beyond] == 0) and np.all(r_lp[beyond] == 0)
      if temperature == 0 and num_states == 2:
        assert abs(g_lp[b][0] - r_lp[0]) <= 1e-5 * abs(r_lp[0]) and r_lp[0] != 0
      continue
    errs += [_rel_err(g_lp[b], r_lp), _rel_err(g_ev[b], r_ev), _ft_err(g_ft[b], r_ft, r_ev)]
  assert g_lp.shape == (nb, lp.size) and lp.size == 9 * sq
  s_lp, s_ev, s_ft = (_np(x) for x in run_vjp(lp, ev, ft[0], n, damping, temperature, w))
  refs = [lbp_vjp.vjp(eg, lp, ev[b], ft[0], n, damping, temperature, w[b]) for b in range(nb)]
  errs += [_rel_err(s_lp, sum(r[0] for r in refs)),
           _ft_err(s_ft, sum(r[2] for r in refs), refs[0][1])]
  errs += [_rel_err(s_ev[b], refs[b][1]) for b in range(nb)]
  print(f"S={num_states} T={temperature} d={damping}: max relative error {max(errs):.2e}")
  assert max(errs) < 3e-5
  run_vjp.close()


def _paths_used(run):
  """Reverse-pass kernels that ran in run() (HIP-event records, option
  "profile" = 2)."""
  from pgmax_amd import _lib  # pylint: disable=g-import-not-at-top
  _lib.profile_kernels()
  prev = _lib.set_option("profile", 2)
  try:
    out = run()
  finally:
    _lib.set_option("profile", prev)
  return out, set(_lib.profile_kernels())


@pytest.mark.parametrize("temperature", [1.0, 0.0])
def test_mixed_state_pairwise_graph(gpu, temperature):
  """Square pairwise factors of 2, 3 and 5 states in one graph (as in an
  uneven batch of Potts grids): the register path with one factor list per
  state count, against the FP64 oracle; a graph with a three-variable
  factor takes the general path."""
  rng = np.random.default_rng(21)
  groups = {s: vgroup.NDVarArray(num_states=s, shape=(5,)) for s in (2, 3, 5)}
  fg = fgraph.FactorGraph(list(groups.values()))
  for s, v in groups.items():
    pairs = [[0, 1], [1, 2], [2, 3], [3, 4], [4, 0], [0, 2]]
    fg.add_factors(fgroup.PairwiseFactorGroup([[v[a], v[b]] for a, b in pairs],
                                              rng.normal(size=(len(pairs), s, s))))
  bp = infer.BP(fg.bp_state, temperature=1.0)
  g = lbp.graph_from_context(bp.context)
  args = g["args"]["enum"]
  eg = lbp_vjp.EnumGraph(g["var_states_for_edge_states"], g["edge_indices_for_edge_states"],
                         args["factor_configs_indices"], args["factor_configs_edge_states"],
                         bp.context.bp_state.fg_state.num_var_states,
                         int(np.asarray(args["factor_configs_indices"]).max()) + 1)
  lp = np.asarray(fg.bp_state.log_potentials.value, np.float32)
  n, nb = 8, 2
  lpb = (lp[None] + rng.normal(0, 0.3, size=(nb, lp.size))).astype(np.float32)
  ev = rng.normal(size=(nb, eg.num_vs)).astype(np.float32)
  ft = rng.normal(size=(nb, eg.num_es)).astype(np.float32)
  w = rng.normal(size=(nb, eg.num_es)).astype(np.float32)
  run_vjp = vjp_mod.RunVJP(bp.context)
  res, used = _paths_used(lambda: [_np(x) for x in run_vjp(lpb, ev, ft, n, 0.5, temperature, w)])
  assert {"vjp_pair_fwd", "vjp_pair_bwd"} <= used
  errs = []
  for b in range(nb):
    r = lbp_vjp.vjp(eg, lpb[b], ev[b], ft[b], n, 0.5, temperature, w[b])
    errs += [_rel_err(res[0][b], r[0]), _rel_err(res[1][b], r[1]), _ft_err(res[2][b], r[2], r[1])]
  print(f"mixed S, T={temperature}: max relative error {max(errs):.2e}")
  assert max(errs) < 3e-5
  run_vjp.close()
  # a three-variable factor: the general path
  bp2, eg2, lp2, _ = _graph(5)
  run2 = vjp_mod.RunVJP(bp2.context)
  _, used2 = _paths_used(lambda: run2(lp2, np.zeros(eg2.num_vs, np.float32),
                                      np.zeros(eg2.num_es, np.float32), 3, 0.5, temperature,
                                      np.ones(eg2.num_es, np.float32)))
  assert not used2 & {"vjp_pair_fwd", "vjp_pair_bwd"}
  run2.close()


@pytest.mark.parametrize("temperature", [1.0, 0.0])
def test_unary_and_pairwise_graph(gpu, temperature):
  """The reference's RBM form (biases as unary EnumFactors before the
  pairwise group, benchmark/rbm_lib.py:140-168) takes the register path too
  (unary factors: one config per state), against the FP64 oracle."""
  import graphs  # pylint: disable=g-import-not-at-top
  fg, _, _, lp, ev = graphs.rbm(n_visible=9, n_hidden=4, batch=2, seed=3, std=0.8,
                                unary_factors=True)
  bp = infer.BP(fg.bp_state, temperature=1.0)
  g = lbp.graph_from_context(bp.context)
  args = g["args"]["enum"]
  eg = lbp_vjp.EnumGraph(g["var_states_for_edge_states"], g["edge_indices_for_edge_states"],
                         args["factor_configs_indices"], args["factor_configs_edge_states"],
                         bp.context.bp_state.fg_state.num_var_states,
                         int(np.asarray(args["factor_configs_indices"]).max()) + 1)
  rng = np.random.default_rng(5)
  lp32 = lp.astype(np.float32)
  ev32 = (ev + rng.normal(0, 0.3, size=ev.shape)).astype(np.float32)
  ft = rng.normal(size=(2, eg.num_es)).astype(np.float32)
  w = rng.normal(size=(2, eg.num_es)).astype(np.float32)
  run_vjp = vjp_mod.RunVJP(bp.context)
  res, used = _paths_used(lambda: [_np(x) for x in run_vjp(lp32, ev32, ft, 10, 0.5, temperature,
                                                           w)])
  assert {"vjp_pair_fwd", "vjp_pair_bwd"} <= used
  errs = []
  for b in range(2):
    r = lbp_vjp.vjp(eg, lp32[b], ev32[b], ft[b], 10, 0.5, temperature, w[b])
    errs += [_rel_err(res[0][b], r[0]), _rel_err(res[1][b], r[1]), _ft_err(res[2][b], r[2], r[1])]
  print(f"RBM with unary factors, T={temperature}: max relative error {max(errs):.2e}")
  assert max(errs) < 3e-5
  run_vjp.close()


def test_gradient_of_a_marginal_loss(gpu):
  """The reference's learning loop (examples/grid_mrf.ipynb cell 15): a
  cross-entropy of marginals after sum-product BP, differentiated with
  respect to the log-potentials and the evidence through torch autograd and
  run_differentiable; equal to the oracle's product with the loss's
  cotangent of the messages."""
  import torch  # pylint: disable=g-import-not-at-top
  bp, eg, lp, rng = _graph(6)
  n, damping = 15, 0.0
  ev = rng.normal(size=eg.num_vs).astype(np.float32)
  target = rng.integers(0, 3, size=6)
  lp_t = torch.tensor(lp, device="cuda", requires_grad=True)
  ev_t = torch.tensor(ev, device="cuda", requires_grad=True)
  ft_t = torch.zeros(eg.num_es, device="cuda")
  msgs = vjp_mod.run_differentiable(bp, lp_t, ev_t, ft_t, n, damping, 1.0)
  msgs.retain_grad()
  beliefs = vjp_mod.flat_beliefs(bp.context, ev_t, msgs).reshape(6, 3)
  marg = torch.softmax(beliefs, dim=-1)
  loss = -torch.log(marg[torch.arange(6), torch.as_tensor(target, device="cuda")]).mean()
  loss.backward()
  g_out = _np(msgs.grad).astype(np.float64)
  r_lp, r_ev, _ = lbp_vjp.vjp(eg, lp, ev, np.zeros(eg.num_es), n, damping, 1.0, g_out)
  # the beliefs' own dependence on the evidence (outside the run)
  b64 = ev.astype(np.float64).copy()
  np.add.at(b64, eg.vsi, lbp_vjp.forward(eg, lp, ev, np.zeros(eg.num_es), n, damping, 1.0)[0])
  p = np.exp(b64.reshape(6, 3) - b64.reshape(6, 3).max(-1, keepdims=True))
  p /= p.sum(-1, keepdims=True)
  onehot = np.eye(3)[target]
  direct = ((p - onehot) / 6).reshape(-1)
  e_lp, e_ev = _rel_err(_np(lp_t.grad), r_lp), _rel_err(_np(ev_t.grad), r_ev + direct)
  print(f"marginal loss: relative error lp {e_lp:.2e}, evidence {e_ev:.2e}")
  assert e_lp < 2e-6 and e_ev < 2e-6  # measured 1.2e-7


@pytest.mark.parametrize("temperature", [0.0, 1.0])
def test_enum_update_writes_every_edge_state(gpu, temperature):
  """pgx_fac_to_var on Enum factors writes every edge state (the reverse
  pass relies on it and skips clearing its output): an output filled with
  NaN comes back equal to the oracle's update everywhere."""
  import ctypes  # pylint: disable=g-import-not-at-top
  import torch  # pylint: disable=g-import-not-at-top
  from pgmax_amd import _lib  # pylint: disable=g-import-not-at-top
  bp, eg, lp, rng = _graph(8)
  vt = rng.normal(size=eg.num_es).astype(np.float32)
  vt_d, lp_d = torch.tensor(vt, device="cuda"), torch.tensor(lp, device="cuda")
  out = torch.full((eg.num_es,), float("nan"), device="cuda")
  lib = _lib.load()
  h = bp.context.engine.graph.handle
  _lib.check(lib.pgx_fac_to_var(h, 0, _lib.ptr(vt_d), _lib.ptr(lp_d), _lib.ptr(out),
                                float(temperature), 1, _lib.current_stream()), "pgx_fac_to_var")
  args = bp.context.inference_arguments
  enum_args = [a for a in args.values() if "factor_configs_indices" in a][0]
  want = lbp.pass_enum_fac_to_var_messages(
      vt, enum_args["factor_configs_indices"], enum_args["factor_configs_edge_states"], lp,
      enum_args["num_val_configs"], temperature)
  np.testing.assert_allclose(_np(out), want, atol=1e-5)
  rc = lib.pgx_fac_to_var(h, 0, _lib.ptr(vt_d), ctypes.c_void_p(), _lib.ptr(out), 1.0, 1,
                          _lib.current_stream())
  assert rc != 0


def _logical_graph(seed, big=False):
  """Pairwise Enum, OR and AND factors over binary variables (children that
  are parents elsewhere; with ``big``, OR factors of 40 and 300 parents:
  chunked sums), with the FP64 oracle's view of it."""
  rng = np.random.default_rng(seed)
  n = 400 if big else 12
  v = vgroup.NDVarArray(num_states=2, shape=(n,))
  fg = fgraph.FactorGraph(v)
  fg.add_factors(fgroup.PairwiseFactorGroup(
      [[v[0], v[1]], [v[1], v[2]], [v[3], v[4]], [v[5], v[0]]], rng.normal(size=(4, 2, 2))))
  ors = [[v[0], v[2], v[3], v[6]], [v[4], v[7]], [v[1], v[5], v[8], v[9], v[2], v[10]]]
  if big:
    ors += [[v[int(i)] for i in rng.choice(np.arange(12, n), k, replace=False)] + [v[11]]
            for k in (40, 300)]
  fg.add_factors(fgroup.ORFactorGroup(variables_for_factors=ors))
  fg.add_factors(fgroup.ANDFactorGroup(variables_for_factors=[
      [v[6], v[7], v[11]], [v[9], v[3], v[10], v[8]]]))
  bp = infer.BP(fg.bp_state, temperature=1.0)
  g = lbp.graph_from_context(bp.context)
  args = g["args"]["enum"]
  logical = [lbp_vjp.LogicalType(g["msgs_range"][t][0], g["args"][t]["parents_factor_indices"],
                                 g["args"][t]["parents_msg_indices"],
                                 g["args"][t]["children_edge_states"],
                                 g["args"][t]["edge_states_offset"]) for t in ("or", "and")]
  eg = lbp_vjp.EnumGraph(g["var_states_for_edge_states"], g["edge_indices_for_edge_states"],
                         args["factor_configs_indices"], args["factor_configs_edge_states"],
                         bp.context.bp_state.fg_state.num_var_states,
                         int(np.asarray(args["factor_configs_indices"]).max()) + 1,
                         logical=logical, num_enum_es=g["msgs_range"]["enum"][1])
  lp = np.asarray(fg.bp_state.log_potentials.value, np.float32)
  return bp, eg, lp, rng


@pytest.mark.parametrize("big", [False, True])
@pytest.mark.parametrize("temperature,damping", [(1.0, 0.5), (0.3, 0.2), (0.7, 0.0), (0.0, 0.5)])
def test_logical_graph_matches_oracle(gpu, temperature, damping, big):
  """OR and AND factors in the reverse pass (their closed forms, top-2 and
  stability maxima pulled back factor by factor) with Enum factors beside
  them, batched and shared inputs, against the FP64 oracle (which
  tests/test_oracle.py pins by finite differences on the same graph)."""
  bp, eg, lp, rng = _logical_graph(9, big)
  n, nb = 6, 2
  lpb = (lp[None] + rng.normal(0, 0.3, size=(nb, lp.size))).astype(np.float32)
  ev = rng.normal(size=(nb, eg.num_vs)).astype(np.float32)
  ft = rng.normal(size=(nb, eg.num_es)).astype(np.float32)
  w = rng.normal(size=(nb, eg.num_es)).astype(np.float32)
  run_vjp = vjp_mod.RunVJP(bp.context)
  res = [_np(x) for x in run_vjp(lpb, ev, ft, n, damping, temperature, w)]
  errs = []
  for b in range(nb):
    r = lbp_vjp.vjp(eg, lpb[b], ev[b], ft[b], n, damping, temperature, w[b])
    errs += [_rel_err(res[0][b], r[0]), _rel_err(res[1][b], r[1]), _ft_err(res[2][b], r[2], r[1])]
  s_lp, s_ev, _ = (_np(x) for x in run_vjp(lp, ev, ft[0], n, damping, temperature, w))
  refs = [lbp_vjp.vjp(eg, lp, ev[b], ft[0], n, damping, temperature, w[b]) for b in range(nb)]
  errs += [_rel_err(s_lp, sum(r[0] for r in refs))]
  errs += [_rel_err(s_ev[b], refs[b][1]) for b in range(nb)]
  print(f"OR/AND (big {big}) T={temperature} d={damping}: max relative error {max(errs):.2e}")
  assert max(errs) < 7e-5  # measured <= 6.5e-6
  run_vjp.close()


@pytest.mark.parametrize("temperature,damping", [(1.0, 0.5), (0.3, 0.2), (0.0, 0.5)])
def test_pool_graph_matches_oracle(gpu, temperature, damping):
  """Pool factors in the reverse pass (the argmax choice's log-sum-exp, the
  log-minus-exps of the others, min / max at T = 0) beside Enum factors,
  against the FP64 oracle (pinned by finite differences on the same
  graph)."""
  rng = np.random.default_rng(13)
  v = vgroup.NDVarArray(num_states=2, shape=(12,))
  fg = fgraph.FactorGraph(v)
  fg.add_factors(fgroup.PairwiseFactorGroup(
      [[v[0], v[1]], [v[2], v[3]], [v[4], v[0]]], rng.normal(size=(3, 2, 2))))
  fg.add_factors(fgroup.PoolFactorGroup(variables_for_factors=[
      [v[0], v[2], v[5], v[6]], [v[7], v[8]], [v[1], v[3], v[9], v[4], v[10], v[11]],
      [v[6], v[9], v[2]]]))
  bp = infer.BP(fg.bp_state, temperature=1.0)
  g = lbp.graph_from_context(bp.context)
  args, a = g["args"]["enum"], g["args"]["pool"]
  eg = lbp_vjp.EnumGraph(g["var_states_for_edge_states"], g["edge_indices_for_edge_states"],
                         args["factor_configs_indices"], args["factor_configs_edge_states"],
                         bp.context.bp_state.fg_state.num_var_states,
                         int(np.asarray(args["factor_configs_indices"]).max()) + 1,
                         num_enum_es=g["msgs_range"]["enum"][1],
                         pool=[lbp_vjp.PoolType(g["msgs_range"]["pool"][0],
                                                a["pool_choices_factor_indices"],
                                                a["pool_choices_msg_indices"],
                                                a["pool_indicators_edge_states"])])
  lp = np.asarray(fg.bp_state.log_potentials.value, np.float32)
  n, nb = 6, 2
  lpb = (lp[None] + rng.normal(0, 0.3, size=(nb, lp.size))).astype(np.float32)
  ev = rng.normal(size=(nb, eg.num_vs)).astype(np.float32)
  ft = rng.normal(size=(nb, eg.num_es)).astype(np.float32)
  w = rng.normal(size=(nb, eg.num_es)).astype(np.float32)
  run_vjp = vjp_mod.RunVJP(bp.context)
  res = [_np(x) for x in run_vjp(lpb, ev, ft, n, damping, temperature, w)]
  errs = []
  for b in range(nb):
    r = lbp_vjp.vjp(eg, lpb[b], ev[b], ft[b], n, damping, temperature, w[b])
    errs += [_rel_err(res[0][b], r[0]), _rel_err(res[1][b], r[1]), _ft_err(res[2][b], r[2], r[1])]
  print(f"Pool T={temperature} d={damping}: max relative error {max(errs):.2e}")
  assert max(errs) < 4e-6  # measured <= 3.3e-7
  run_vjp.close()
