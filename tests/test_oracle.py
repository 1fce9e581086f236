"""Pins the CPU oracle (oracle/lbp.py) before it is used to check the HIP engine.

Each check can fail for a plausible mistake in the restatement: the
reference's own golden vector (84 messages after 100 max-product iterations
and 12 MAP states), brute-force enumeration on trees, and the 2-state chain
in closed form.
"""

import os

import numpy as np
import pytest

import graphs
from oracle import brute
from oracle import f32math
from oracle import lbp
from pgmax_amd import factor
from pgmax_amd import fgraph
from pgmax_amd import fgroup
from pgmax_amd import vgroup
from pgmax_amd.infer.inferer import InfererContext

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "e2e_sanity.npz")


def _oracle_inputs(bp_state):
  ctx = InfererContext(bp_state)
  g = lbp.graph_from_context(ctx)
  lp = np.asarray(bp_state.log_potentials.value, dtype=np.float32)
  ftov = np.asarray(bp_state.ftov_msgs.value, dtype=np.float32)
  ev = np.asarray(bp_state.evidence.value, dtype=np.float32)
  return ctx, g, lp, ftov, ev


def test_golden_e2e_messages_and_map():
  """Reference tests/test_pgmax.py:416-421: 84 messages at atol 1e-6 and MAP."""
  golden = np.load(GOLDEN)
  _, bp_state, grid_vars, additional_vars = graphs.e2e_sanity_graph()
  ctx, g, lp, ftov, ev = _oracle_inputs(bp_state)
  msgs, _ = lbp.run_with_diffs(g, lp, ftov, ev, num_iters=100, damping=0.5,
                               temperature=0.0, dtype=np.float32)
  assert msgs.shape == (84,)
  np.testing.assert_allclose(msgs, golden["final_msgs"], atol=1e-6, rtol=0)
  beliefs = lbp.get_flat_beliefs(g, msgs, ev)
  from pgmax_amd.infer.inferer import unflatten_beliefs, decode_map_states
  dec = decode_map_states(unflatten_beliefs(beliefs, [grid_vars, additional_vars]))
  for grp, idx, state in zip(golden["map_group"], golden["map_index"], golden["map_state"]):
    if grp == "grid_vars":
      assert dec[grid_vars][tuple(idx)] == state
    else:
      assert dec[additional_vars][tuple(int(i) for i in idx)] == state


@pytest.mark.parametrize("mistake", ["drop_evidence", "add_incoming", "sum_not_max"])
def test_golden_fails_with_planted_mistake(monkeypatch, mistake):
  """Plausible mistakes planted in the oracle must break the golden (the check has teeth)."""
  golden = np.load(GOLDEN)
  _, bp_state, _, _ = graphs.e2e_sanity_graph()
  _, g, lp, ftov, ev = _oracle_inputs(bp_state)
  temperature = 0.0
  if mistake == "drop_evidence":
    orig = lbp.var_state_sums
    monkeypatch.setattr(lbp, "var_state_sums", lambda m, e, v: orig(m, 0 * e, v))
  elif mistake == "add_incoming":
    orig = lbp.pass_enum_fac_to_var_messages
    def wrong(vtof, *a):
      return orig(vtof, *a) + 2 * vtof
    monkeypatch.setattr(lbp, "pass_enum_fac_to_var_messages", wrong)
  else:
    temperature = 1.0
  msgs, _ = lbp.run_with_diffs(g, lp, ftov, ev, num_iters=100, damping=0.5,
                               temperature=temperature)
  assert not np.allclose(msgs, golden["final_msgs"], atol=1e-6)


def _random_tree(rng, n_vars, max_states=3):
  ns = rng.integers(2, max_states + 1, size=n_vars)
  varr = vgroup.NDVarArray(num_states=ns.astype(np.int64), shape=(n_vars,))
  fg = fgraph.FactorGraph(varr)
  for v in range(1, n_vars):
    parent = int(rng.integers(0, v))
    a, b = varr[parent], varr[v]
    configs = np.array([[i, j] for i in range(a[1]) for j in range(b[1])])
    fg.add_factors(factor.EnumFactor(
        variables=[a, b], factor_configs=configs,
        log_potentials=rng.normal(size=configs.shape[0]),
    ))
  return fg, varr, ns


@pytest.mark.parametrize("temperature", [1.0, 0.0])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_tree_bp_is_exact(seed, temperature):
  """On a tree, converged BP equals brute-force (max-)marginals in FP64."""
  rng = np.random.default_rng(seed)
  fg, varr, ns = _random_tree(rng, 6)
  bp_state = fg.bp_state
  ev = rng.normal(size=int(ns.sum()))
  _, g, lp, ftov, _ = _oracle_inputs(bp_state)
  msgs, deltas = lbp.run_with_diffs(g, lp, ftov, ev, num_iters=60, damping=0.5,
                                    temperature=temperature, dtype=np.float64)
  assert deltas[-1] < 1e-10
  beliefs = lbp.get_flat_beliefs(g, msgs, ev, dtype=np.float64)
  starts = np.insert(np.cumsum(ns), 0, 0)
  tables = brute.enum_factor_tables(g, lp.astype(np.float64))
  exact, _ = brute.exact_marginals(list(ns), list(starts[:-1]), ev, tables,
                                   temperature=temperature)
  for v in range(len(ns)):
    b = beliefs[starts[v]:starts[v + 1]]
    if temperature == 0.0:
      np.testing.assert_allclose(b - b.max(), exact[v], atol=1e-9)
    else:
      p = np.exp(b - b.max())
      np.testing.assert_allclose(p / p.sum(), exact[v], atol=1e-9)


def test_chain_transfer_matrix():
  """2-state chain: BP marginal of the last variable equals the transfer-matrix product."""
  n = 8
  rng = np.random.default_rng(5)
  varr = vgroup.NDVarArray(num_states=2, shape=(n,))
  fg = fgraph.FactorGraph(varr)
  js = rng.normal(size=n - 1)
  lpm = np.stack([np.array([[j, -j], [-j, j]]) for j in js])
  fg.add_factors(fgroup.PairwiseFactorGroup(
      variables_for_factors=[[varr[i], varr[i + 1]] for i in range(n - 1)],
      log_potential_matrix=lpm))
  h = rng.normal(size=(n, 2))
  _, g, lp, ftov, _ = _oracle_inputs(fg.bp_state)
  msgs, _ = lbp.run_with_diffs(g, lp, ftov, h.reshape(-1), num_iters=80,
                               damping=0.5, temperature=1.0, dtype=np.float64)
  beliefs = lbp.get_flat_beliefs(g, msgs, h.reshape(-1), dtype=np.float64)
  # forward transfer: alpha_{i+1}(b) = sum_a alpha_i(a) e^{h_i(a)} e^{J_i s(a)s(b)}
  alpha = np.ones(2)
  for i in range(n - 1):
    t = np.exp(np.array([[js[i], -js[i]], [-js[i], js[i]]]))
    alpha = (alpha * np.exp(h[i])) @ t
  p_last = alpha * np.exp(h[n - 1])
  p_last /= p_last.sum()
  b = beliefs[-2:]
  p = np.exp(b - b.max())
  np.testing.assert_allclose(p / p.sum(), p_last, atol=1e-10)


def _ulps(got32, want64):
  got = got32.astype(np.float64)
  sp = np.spacing(np.abs(want64.astype(np.float32))).astype(np.float64)
  return np.abs(got - want64) / sp


def test_f32_transcendentals_within_2p5_ulp():
  """oracle/f32math.py (= the kernels' pgx_math.h) against float64, <= 2.5 ulp."""
  from oracle import f32math
  rng = np.random.default_rng(11)
  x = np.concatenate([rng.uniform(-87, 88, 200000), rng.uniform(-1, 1, 100000),
                      -np.logspace(-8, 1.9, 50000), np.logspace(-8, 1.9, 50000)]).astype(np.float32)
  assert _ulps(f32math.exp(x), np.exp(x.astype(np.float64))).max() <= 2.5
  xp = np.concatenate([rng.uniform(1e-30, 1e30, 100000), rng.uniform(0.5, 2, 100000),
                       np.logspace(-37, 38, 100000)]).astype(np.float32)
  assert _ulps(f32math.log(xp), np.log(xp.astype(np.float64))).max() <= 2.5
  x1 = np.concatenate([rng.uniform(-0.999, 1, 200000), np.logspace(-9, 0, 50000),
                       -np.logspace(-9, -0.01, 50000)]).astype(np.float32)
  assert _ulps(f32math.log1p(x1), np.log1p(x1.astype(np.float64))).max() <= 2.5
  xm = np.concatenate([rng.uniform(-1, 0, 200000), -np.logspace(-9, 0, 50000),
                       rng.uniform(-20, 20, 50000)]).astype(np.float32)
  assert _ulps(f32math.expm1(xm), np.expm1(xm.astype(np.float64))).max() <= 2.5
  # special values
  sp = np.array([np.nan, np.inf, -np.inf, 0.0, -0.0], dtype=np.float32)
  np.testing.assert_array_equal(f32math.exp(sp), [np.nan, np.inf, 0, 1, 1])
  np.testing.assert_array_equal(f32math.log(np.array([0, -1, np.inf], np.float32)),
                                [-np.inf, np.nan, np.inf])
  assert f32math.log1p(np.float32(-1)) == -np.inf
  assert f32math.expm1(np.float32(-np.inf)) == -1


def test_var_sum_order():
  """Degree <= 32: plain sequential scatter-add; higher degree: chunked, within rounding of the exact sum."""
  rng = np.random.default_rng(3)
  vs_small = np.repeat(np.arange(50), 7)
  m = rng.normal(size=vs_small.shape[0]).astype(np.float32)
  ev = rng.normal(size=50).astype(np.float32)
  seq = np.zeros(50, np.float32)
  np.add.at(seq, vs_small, m)
  np.testing.assert_array_equal(lbp.var_state_sums(m, ev, vs_small), seq + ev)
  vs_big = rng.permutation(np.repeat(np.arange(40), 500))
  m = rng.normal(size=vs_big.shape[0]).astype(np.float32)
  ev = rng.normal(size=40).astype(np.float32)
  exact = np.zeros(40)
  np.add.at(exact, vs_big, m.astype(np.float64))
  got = lbp.var_state_sums(m, ev, vs_big)
  np.testing.assert_allclose(got, exact + ev, atol=500 * 6e-8 * 4)
  # the chunked order is not the plain order for long lists
  seq = np.zeros(40, np.float32)
  np.add.at(seq, vs_big, m)
  assert not np.array_equal(got, seq + ev)


def test_c_transcendentals_equal_numpy_oracle():
  """The C restatement's exp and log are the oracle's, bit for bit."""
  from oracle import cpu
  if not os.path.exists(cpu.LIB):
    cpu.build()
  rng = np.random.default_rng(11)
  x = np.concatenate([
      rng.uniform(-110, 95, 200000), rng.normal(0, 3, 200000), -np.exp(rng.uniform(-30, 5, 50000)),
      np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 88.7228317, -103.972084, -104.0, 1e-40])
  ]).astype(np.float32)
  np.testing.assert_array_equal(cpu.math("exp", x), f32math.exp(x))
  pos = np.concatenate([np.exp(rng.uniform(-90, 88, 200000)), rng.uniform(0.5, 2.0, 200000),
                        np.array([0.0, np.inf, -1.0, np.nan, 1e-42, 3.4e38])]).astype(np.float32)
  np.testing.assert_array_equal(cpu.math("log", pos), f32math.log(pos))


@pytest.mark.parametrize("temperature", [0.0, 1.0, 0.35])
def test_c_baseline_matches_numpy_oracle(temperature):
  """The timed C/OpenMP baseline computes what the NumPy FP32 oracle does, bit
  for bit, so it can serve as the oracle for full-size runs."""
  from oracle import cpu
  if not os.path.exists(cpu.LIB):
    cpu.build()
  fg, v, ev = graphs.ising_torus(side=6, seed=3)
  ctx = InfererContext(fg.bp_state)
  g = lbp.graph_from_context(ctx)
  bp = fg.bp_state
  lp = np.asarray(bp.log_potentials.value, np.float32)
  evid = np.zeros(ctx.bp_state.fg_state.num_var_states, np.float32)
  evid[:] = np.random.default_rng(1).normal(0, 0.5, size=evid.size)
  ft0 = np.zeros(g["var_states_for_edge_states"].shape[0], np.float32)
  ref, ref_d = lbp.run_with_diffs(g, lp, ft0, evid, 60, 0.5, temperature, np.float32)
  got, got_d = cpu.run_enum(g, lp, ft0, evid, 60, 0.5, temperature, threads=2)
  np.testing.assert_array_equal(got, ref)
  np.testing.assert_array_equal(got_d, ref_d)
  # batched with per-element evidence
  evb = np.stack([evid, evid[::-1].copy()])
  gb, _ = cpu.run_enum(g, lp, ft0, evb, 60, 0.5, temperature, threads=2)
  np.testing.assert_array_equal(gb[0], got)
  # independent graphs in one call, one of them with non-finite potentials
  lp_bad = lp.copy()
  lp_bad[::7] = np.float32(np.inf)
  lp_bad[3::11] = np.float32(-1e7)
  lp_bad[5] = np.float32(np.nan)
  outs, deltas = cpu.run_enum_graphs([g, g], [lp, lp_bad], [evid, evid], [ft0, ft0], 20, 0.5,
                                     temperature, threads=2)
  for out_g, d_g, lp_g in zip(outs, deltas, (lp, lp_bad)):
    r, rd = lbp.run_with_diffs(g, lp_g, ft0, evid, 20, 0.5, temperature, np.float32)
    np.testing.assert_array_equal(out_g, r)
    np.testing.assert_array_equal(d_g, rd)


@pytest.mark.parametrize("temperature", [0.0, 1.0])
def test_c_baseline_chunked_sums_match_numpy_oracle(temperature):
  """Variables with more than 32 messages (an RBM's units: 70 and 40 edges)
  and more than 32 x 32 (a star's centre: 1,100 edges) sum in runs and
  groups of runs in the C restatement too, bit for bit."""
  from oracle import cpu
  if not os.path.exists(cpu.LIB):
    cpu.build()
  fg, _, _, lp, ev = graphs.rbm(n_visible=70, n_hidden=40, batch=2, seed=4)
  g = lbp.graph_from_context(InfererContext(fg.bp_state))
  ft0 = np.zeros(g["var_states_for_edge_states"].shape[0], np.float32)
  lp32, ev32 = lp.astype(np.float32), ev.astype(np.float32)
  got, got_d = cpu.run_enum(g, lp32, ft0, ev32, 30, 0.5, temperature, threads=2)
  for b in range(2):
    ref, ref_d = lbp.run_with_diffs(g, lp32[b], ft0, ev32[b], 30, 0.5, temperature, np.float32)
    np.testing.assert_array_equal(got[b], ref)
    np.testing.assert_array_equal(got_d[b], ref_d)
  leaves = vgroup.NDVarArray(num_states=2, shape=(1100,))
  centre = vgroup.NDVarArray(num_states=3, shape=(1,))
  star = fgraph.FactorGraph([centre, leaves])
  rng = np.random.default_rng(9)
  star.add_factors(fgroup.PairwiseFactorGroup(
      [[centre[0], leaves[i]] for i in range(1100)], rng.normal(size=(1100, 3, 2))))
  ctx = InfererContext(star.bp_state)
  g = lbp.graph_from_context(ctx)
  lp = np.asarray(star.bp_state.log_potentials.value, np.float32)
  ev = rng.normal(size=ctx.bp_state.fg_state.num_var_states).astype(np.float32)
  ft0 = rng.normal(size=g["var_states_for_edge_states"].shape[0]).astype(np.float32)
  got, got_d = cpu.run_enum(g, lp, ft0, ev, 12, 0.5, temperature, threads=2)
  ref, ref_d = lbp.run_with_diffs(g, lp, ft0, ev, 12, 0.5, temperature, np.float32)
  np.testing.assert_array_equal(got, ref)
  np.testing.assert_array_equal(got_d, ref_d)


def test_chunked_segment_sum_order():
  """Plain sequential sums up to 32 values; runs of 32, then groups of 32 runs."""
  rng = np.random.default_rng(7)
  sizes = [1, 5, 32, 33, 100, 1024, 1025, 3000]
  vals = [(rng.normal(size=n) * 10.0 ** rng.integers(-3, 4, size=n)).astype(np.float32) for n in sizes]
  seg = np.concatenate([np.full(n, i) for i, n in enumerate(sizes)])
  # interleave segments so array order != segment-major order
  perm = rng.permutation(seg.size)
  flat = np.concatenate(vals)
  seg_p, flat_p = seg[perm], flat[perm]
  got = lbp.chunked_segment_sum(flat_p, seg_p, len(sizes))
  for i, n in enumerate(sizes):
    x = flat_p[seg_p == i]  # the segment's values in array order
    want = np.float32(0)
    for g in range(0, n, 32 * 32):
      gs = np.float32(0)
      for r in range(g, min(n, g + 32 * 32), 32):
        rs = np.float32(0)
        for v in x[r:r + 32]:
          rs = np.float32(rs + v)
        gs = np.float32(gs + rs)
      want = np.float32(want + gs)
    assert got[i] == want, (n, got[i], want)
    if n <= 32:
      seq = np.float32(0)
      for v in x:
        seq = np.float32(seq + v)
      assert got[i] == seq


@pytest.mark.parametrize("kind", ["or", "and", "pool"])
@pytest.mark.parametrize("temperature,atol", [(0.0, 1e-5), (1.0, 1e-5), (0.7, 1e-5), (0.3, 5e-3),
                                              (0.001, 5e-3)])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_oracle_logical_closed_forms_equal_enum_expansion(kind, temperature, atol, seed):
  """The oracle's OR / AND / Pool closed forms against their definition: the
  same factors written as EnumFactors over every valid configuration
  (reference tests/factor/test_or.py, test_and.py, test_pool.py, with their
  temperatures and tolerances). Two graphs hold the factors half as Enum,
  half closed-form, the other way round; after 5 iterations from random
  evidence and messages their beliefs agree."""
  from pgmax_amd.infer import bp_state as bps  # pylint: disable=g-import-not-at-top
  pair, (ev_p, ev_c), (ft_p, ft_c) = graphs.logical_pair(kind, seed)
  beliefs = []
  for fg, parents, children, _ in pair:
    state = fg.bp_state
    fgs = state.fg_state
    ev = bps.update_evidence(np.zeros(fgs.num_var_states), {parents: ev_p, children: ev_c}, fgs)
    updates = {parents[k]: ft_p[k] for k in range(ft_p.shape[0])}
    updates.update({children[k]: ft_c[k] for k in range(ft_c.shape[0])})
    ft = bps.update_ftov_msgs(np.zeros(fgs.total_factor_num_states), updates, fgs)
    g = lbp.graph_from_context(InfererContext(state))
    lp = np.asarray(state.log_potentials.value, np.float64)
    msgs, _ = lbp.run_with_diffs(g, lp, ft, ev, 5, 0.5, temperature, np.float64)
    beliefs.append(lbp.get_flat_beliefs(g, msgs, ev, np.float64))
  np.testing.assert_allclose(beliefs[0], beliefs[1], atol=atol)


@pytest.mark.parametrize("kind", ["or", "and", "pool"])
def test_oracle_logical_equivalence_catches_a_planted_mistake(monkeypatch, kind):
  """The check above fails when the closed form gives the second-largest
  difference to the wrong parent (each factor's argmax pointed at its first
  parent): the expansion pins the argmax term."""
  real = lbp.get_maxes_and_argmaxes

  def first_argmax(data, labels, num_labels):
    maxes, argmaxes = real(data, labels, num_labels)
    # keep the maxes, point the argmax at the first parent of each factor
    counts = np.bincount(labels, minlength=num_labels)
    first = np.concatenate([[0], np.cumsum(counts)])[:-1]
    return maxes, first[:num_labels]

  monkeypatch.setattr(lbp, "get_maxes_and_argmaxes", first_argmax)
  with pytest.raises(AssertionError):
    for seed in range(3):
      test_oracle_logical_closed_forms_equal_enum_expansion(kind, 0.0, 1e-5, seed)


def test_map_and_marginals_oracle_against_float64():
  """The oracle's MAP states are numpy's first argmax (first NaN when there is
  one); its marginals are the float64 softmax to 1e-6 and sum to 1."""
  rng = np.random.default_rng(21)
  ns = rng.integers(1, 9, size=500)
  vs0 = np.concatenate([[0], np.cumsum(ns)])
  x = rng.integers(-3, 3, size=vs0[-1]).astype(np.float32)  # many ties
  x[rng.choice(x.size, 40, replace=False)] = rng.normal(0, 30, size=40)
  x[vs0[7] + 1 if ns[7] > 1 else vs0[7]] = np.nan
  x[vs0[11]:vs0[12]] = -np.inf
  best, marg = lbp.map_and_marginals(x, vs0)
  for v in range(ns.size):
    seg = x[vs0[v]:vs0[v + 1]]
    nan = np.flatnonzero(np.isnan(seg))
    assert best[v] == (nan[0] if nan.size else int(np.argmax(seg))), v
    if nan.size or np.all(seg == -np.inf):
      assert np.all(np.isnan(marg[vs0[v]:vs0[v + 1]]))
      continue
    e = np.exp(seg.astype(np.float64) - seg.max())
    np.testing.assert_allclose(marg[vs0[v]:vs0[v + 1]], e / e.sum(), atol=1e-6, rtol=0)
    assert abs(float(marg[vs0[v]:vs0[v + 1]].astype(np.float64).sum()) - 1.0) < 1e-6


def _vjp_graph(seed):
  """A loopy graph of 3-state variables: pairwise factors on a 4-cycle plus a
  diagonal, and one 3-variable factor with a sparse config table."""
  from oracle import lbp_vjp  # pylint: disable=g-import-not-at-top
  rng = np.random.default_rng(seed)
  v = vgroup.NDVarArray(num_states=3, shape=(5,))
  fg = fgraph.FactorGraph(v)
  fg.add_factors(fgroup.PairwiseFactorGroup(
      [[v[0], v[1]], [v[1], v[2]], [v[2], v[3]], [v[3], v[0]], [v[0], v[2]]],
      rng.normal(size=(5, 3, 3))))
  cfg = np.array([[a, b, c] for a in range(3) for b in range(3) for c in range(3)
                  if (a + b + c) % 2 == 0])
  fg.add_factors(fgroup.EnumFactorGroup([[v[1], v[3], v[4]]], cfg,
                                        log_potentials=rng.normal(size=(1, cfg.shape[0]))))
  ctx = InfererContext(fg.bp_state)
  g = lbp.graph_from_context(ctx)
  args = g["args"]["enum"]
  eg = lbp_vjp.EnumGraph(g["var_states_for_edge_states"], g["edge_indices_for_edge_states"],
                         args["factor_configs_indices"], args["factor_configs_edge_states"],
                         ctx.bp_state.fg_state.num_var_states,
                         int(np.asarray(args["factor_configs_indices"]).max()) + 1)
  lp = np.asarray(fg.bp_state.log_potentials.value, np.float64)
  return eg, lp, rng


def _vjp_fd_mismatch(temperature, damping, seed=3, n=6):
  """Largest |finite difference - VJP| over every log-potential, evidence
  entry and initial message, relative to 1e-6 + 1e-5 |fd|."""
  from oracle import lbp_vjp  # pylint: disable=g-import-not-at-top
  eg, lp, rng = _vjp_graph(seed)
  ev = rng.normal(size=eg.num_vs)
  ft = rng.normal(size=eg.num_es)
  w = rng.normal(size=eg.num_es)
  grads = dict(zip(("lp", "ev", "ft"), lbp_vjp.vjp(eg, lp, ev, ft, n, damping, temperature, w)))
  base = {"lp": lp, "ev": ev, "ft": ft}

  def loss(**kw):
    x = dict(base, **kw)
    return float(w @ lbp_vjp.forward(eg, x["lp"], x["ev"], x["ft"], n, damping, temperature)[0])

  h, worst = 1e-6, 0.0
  for name, arr in base.items():
    for i in range(arr.size):
      up, dn = arr.copy(), arr.copy()
      up[i] += h
      dn[i] -= h
      fd = (loss(**{name: up}) - loss(**{name: dn})) / (2 * h)
      worst = max(worst, abs(fd - grads[name][i]) / (1e-6 + 1e-5 * abs(fd)))
  return worst


@pytest.mark.parametrize("temperature,damping", [(1.0, 0.5), (1.0, 0.0), (0.4, 0.3), (0.0, 0.5)])
def test_bp_vjp_oracle_matches_finite_differences(temperature, damping):
  """The reverse-mode oracle's product with a random cotangent equals
  central finite differences of <cotangent, run(...)> in every
  log-potential, evidence entry and initial message (FP64, 6 iterations,
  a loopy graph of pairwise and 3-variable factors)."""
  assert _vjp_fd_mismatch(temperature, damping) <= 1.0


def test_bp_vjp_oracle_catches_a_planted_mistake(monkeypatch):
  """The finite-difference check fails when the reverse pass drops the
  pull-back through the variable sums (each message feeds its variable's
  sum)."""
  from oracle import lbp_vjp  # pylint: disable=g-import-not-at-top
  monkeypatch.setattr(lbp_vjp, "var_sum_pullback", lambda g, gvt: np.zeros(g.num_vs))
  assert _vjp_fd_mismatch(1.0, 0.5) > 10.0


def _vjp_pool_graph(seed):
  """Pairwise Enum and Pool factors (1 to 5 choices; a choice that is also
  another pool's indicator) over binary variables."""
  from oracle import lbp_vjp  # pylint: disable=g-import-not-at-top
  rng = np.random.default_rng(seed)
  v = vgroup.NDVarArray(num_states=2, shape=(12,))
  fg = fgraph.FactorGraph(v)
  fg.add_factors(fgroup.PairwiseFactorGroup(
      [[v[0], v[1]], [v[2], v[3]], [v[4], v[0]]], rng.normal(size=(3, 2, 2))))
  fg.add_factors(fgroup.PoolFactorGroup(variables_for_factors=[
      [v[0], v[2], v[5], v[6]], [v[7], v[8]], [v[1], v[3], v[9], v[4], v[10], v[11]],
      [v[6], v[9], v[2]]]))
  ctx = InfererContext(fg.bp_state)
  g = lbp.graph_from_context(ctx)
  args = g["args"]["enum"]
  a = g["args"]["pool"]
  pool = [lbp_vjp.PoolType(g["msgs_range"]["pool"][0], a["pool_choices_factor_indices"],
                           a["pool_choices_msg_indices"], a["pool_indicators_edge_states"])]
  eg = lbp_vjp.EnumGraph(g["var_states_for_edge_states"], g["edge_indices_for_edge_states"],
                         args["factor_configs_indices"], args["factor_configs_edge_states"],
                         ctx.bp_state.fg_state.num_var_states,
                         int(np.asarray(args["factor_configs_indices"]).max()) + 1,
                         num_enum_es=g["msgs_range"]["enum"][1], pool=pool)
  lp = np.asarray(fg.bp_state.log_potentials.value, np.float64)
  return eg, lp, rng, g


def _vjp_logical_graph(seed):
  """Binary variables with pairwise Enum factors, OR factors of 1 to 5
  parents and AND factors of 2 and 3 parents, children that are parents
  elsewhere: (EnumGraph with its logical types, log-potentials, rng)."""
  from oracle import lbp_vjp  # pylint: disable=g-import-not-at-top
  rng = np.random.default_rng(seed)
  v = vgroup.NDVarArray(num_states=2, shape=(12,))
  fg = fgraph.FactorGraph(v)
  fg.add_factors(fgroup.PairwiseFactorGroup(
      [[v[0], v[1]], [v[1], v[2]], [v[3], v[4]], [v[5], v[0]]], rng.normal(size=(4, 2, 2))))
  fg.add_factors(fgroup.ORFactorGroup(variables_for_factors=[
      [v[0], v[2], v[3], v[6]], [v[4], v[7]], [v[1], v[5], v[8], v[9], v[2], v[10]]]))
  fg.add_factors(fgroup.ANDFactorGroup(variables_for_factors=[
      [v[6], v[7], v[11]], [v[9], v[3], v[10], v[8]]]))
  ctx = InfererContext(fg.bp_state)
  g = lbp.graph_from_context(ctx)
  args = g["args"]["enum"]
  logical = []
  for t in ("or", "and"):
    a = g["args"][t]
    logical.append(lbp_vjp.LogicalType(g["msgs_range"][t][0], a["parents_factor_indices"],
                                       a["parents_msg_indices"], a["children_edge_states"],
                                       a["edge_states_offset"]))
  eg = lbp_vjp.EnumGraph(g["var_states_for_edge_states"], g["edge_indices_for_edge_states"],
                         args["factor_configs_indices"], args["factor_configs_edge_states"],
                         ctx.bp_state.fg_state.num_var_states,
                         int(np.asarray(args["factor_configs_indices"]).max()) + 1,
                         logical=logical, num_enum_es=g["msgs_range"]["enum"][1])
  lp = np.asarray(fg.bp_state.log_potentials.value, np.float64)
  return eg, lp, rng, g


def _vjp_logical_fd_mismatch(temperature, damping, seed=4, n=5, graph=None):
  from oracle import lbp_vjp  # pylint: disable=g-import-not-at-top
  eg, lp, rng, _ = (graph or _vjp_logical_graph)(seed)
  ev = rng.normal(size=eg.num_vs)
  ft = rng.normal(size=eg.num_es)
  w = rng.normal(size=eg.num_es)
  grads = dict(zip(("lp", "ev", "ft"), lbp_vjp.vjp(eg, lp, ev, ft, n, damping, temperature, w)))
  base = {"lp": lp, "ev": ev, "ft": ft}

  def loss(**kw):
    x = dict(base, **kw)
    return float(w @ lbp_vjp.forward(eg, x["lp"], x["ev"], x["ft"], n, damping, temperature)[0])

  h, worst = 1e-6, 0.0
  for name, arr in base.items():
    for i in range(arr.size):
      up, dn = arr.copy(), arr.copy()
      up[i] += h
      dn[i] -= h
      fd = (loss(**{name: up}) - loss(**{name: dn})) / (2 * h)
      worst = max(worst, abs(fd - grads[name][i]) / (1e-6 + 1e-5 * abs(fd)))
  return worst


def test_bp_vjp_oracle_logical_forward_matches_bp_oracle():
  """The reverse-mode oracle's forward run of a graph with OR and AND
  factors is the BP oracle's (oracle/lbp.py, FP64)."""
  from oracle import lbp_vjp  # pylint: disable=g-import-not-at-top
  eg, lp, rng, g = _vjp_logical_graph(4)
  ev = rng.normal(size=eg.num_vs)
  ft = rng.normal(size=eg.num_es)
  for t in (0.0, 0.3, 1.0):
    got = lbp_vjp.forward(eg, lp, ev, ft, 7, 0.4, t)[0]
    want, _ = lbp.run_with_diffs(g, lp, ft, ev, 7, 0.4, t, np.float64)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("temperature,damping", [(1.0, 0.5), (0.3, 0.2), (0.7, 0.0), (0.0, 0.5)])
def test_bp_vjp_oracle_logical_matches_finite_differences(temperature, damping):
  """With OR and AND factors (closed forms, their top-2 and, below T = 0.5,
  the stability maxima) the reverse-mode oracle equals central finite
  differences in every log-potential, evidence entry and initial message."""
  assert _vjp_logical_fd_mismatch(temperature, damping) <= 1.0


def test_bp_vjp_oracle_pool_forward_matches_bp_oracle():
  from oracle import lbp_vjp  # pylint: disable=g-import-not-at-top
  eg, lp, rng, g = _vjp_pool_graph(5)
  ev = rng.normal(size=eg.num_vs)
  ft = rng.normal(size=eg.num_es)
  for t in (0.0, 0.3, 1.0):
    got = lbp_vjp.forward(eg, lp, ev, ft, 7, 0.4, t)[0]
    want, _ = lbp.run_with_diffs(g, lp, ft, ev, 7, 0.4, t, np.float64)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("temperature,damping", [(1.0, 0.5), (0.3, 0.2), (0.0, 0.5)])
def test_bp_vjp_oracle_pool_matches_finite_differences(temperature, damping):
  """The same with Pool factors (the argmax choice's separate
  log-sum-exp, the log-minus-exp of the others, max / min at T = 0)."""
  assert _vjp_logical_fd_mismatch(temperature, damping, seed=5, graph=_vjp_pool_graph) <= 1.0


def test_bp_vjp_oracle_pool_catches_a_planted_mistake(monkeypatch):
  """Uniform instead of softmax weights in the Pool log-sum-exps' gradient
  fails the finite-difference check."""
  from oracle import lbp_vjp  # pylint: disable=g-import-not-at-top
  orig = lbp_vjp._lse  # pylint: disable=protected-access

  def wrong(x, t):
    val, _ = orig(x, t)
    return val, np.full(x.shape, 1.0 / x.size)
  monkeypatch.setattr(lbp_vjp, "_lse", wrong)
  assert _vjp_logical_fd_mismatch(1.0, 0.5, seed=5, graph=_vjp_pool_graph) > 10.0


def test_bp_vjp_oracle_logical_catches_a_planted_mistake(monkeypatch):
  """Dropping the second-largest diff's share (the argmax parent's other
  message) fails the finite-difference check."""
  from oracle import lbp_vjp  # pylint: disable=g-import-not-at-top
  orig = lbp_vjp._top2  # pylint: disable=protected-access

  def wrong(d):
    m1, a1, m2, w1, w2 = orig(d)
    return m1, a1, m2, w1, np.zeros_like(w2)
  monkeypatch.setattr(lbp_vjp, "_top2", wrong)
  assert _vjp_logical_fd_mismatch(0.0, 0.5) > 10.0 or _vjp_logical_fd_mismatch(0.3, 0.2) > 10.0


def _or_and_graph(seed):
  """OR and AND factors over shared binary variables: factor sizes from 1
  to 1,100 parents (runs and groups of runs in the factor sums), one
  variable in 40 AND factors (chunked variable sums), children that are
  parents elsewhere."""
  rng = np.random.default_rng(seed)
  x = vgroup.NDVarArray(num_states=2, shape=(1500,))
  kids = vgroup.NDVarArray(num_states=2, shape=(60,))
  fg = fgraph.FactorGraph([x, kids])
  or_sizes = [1, 2, 3, 33, 40, 1100, 5, 64]
  ors = [[x[int(i)] for i in rng.choice(1500, n, replace=False)] + [kids[k]]
         for k, n in enumerate(or_sizes)]
  ands = []
  for k in range(len(or_sizes), 60):
    n = int(rng.integers(1, 4)) if k % 10 else 70
    par = [int(i) for i in rng.choice(np.arange(1, 1500), n, replace=False)]
    if k < len(or_sizes) + 40:
      par[0] = 0
    ands.append([x[i] for i in par] + [kids[k]])
  ands[-1] = [kids[0], kids[1], x[7], kids[59]]  # children as parents
  fg.add_factors(fgroup.ORFactorGroup(variables_for_factors=ors))
  fg.add_factors(fgroup.ANDFactorGroup(variables_for_factors=ands))
  return fg


@pytest.mark.parametrize("ties", [False, True])
def test_c_logical_max_product_matches_numpy_oracle(ties):
  """The C restatement of max-product OR / AND BP (the full-size oracle of
  the deconvolution workload) computes what the NumPy FP32 oracle does, bit
  for bit -- with values on a coarse grid, ties in the top-2 diffs, the
  maxima and the normalisation included."""
  from oracle import cpu
  if not os.path.exists(cpu.LIB):
    cpu.build()
  fg = _or_and_graph(3)
  ctx = InfererContext(fg.bp_state)
  g = lbp.graph_from_context(ctx)
  rng = np.random.default_rng(4)
  n_vs = ctx.bp_state.fg_state.num_var_states
  n_es = g["var_states_for_edge_states"].shape[0]
  ev = rng.gumbel(size=n_vs).astype(np.float32)
  ft = rng.normal(size=n_es).astype(np.float32)
  if ties:
    ev, ft = np.round(ev * 2) / 2, np.round(ft * 2) / 2
  for damping, n in ((0.5, 25), (0.0, 7)):
    ref, ref_d = lbp.run_with_diffs(g, np.zeros(0, np.float32), ft, ev, n, damping, 0.0,
                                    np.float32)
    got, got_d = cpu.run_logical_t0(g, ft, ev, n, damping, threads=3)
    np.testing.assert_array_equal(got, ref)
    np.testing.assert_array_equal(got_d, ref_d)


def test_c_logical_max_product_deconvolution():
  """The same on the deconvolution structure (AND(S, W, SW), OR(SW..., X))."""
  from oracle import cpu
  from pgmax_amd import workloads
  if not os.path.exists(cpu.LIB):
    cpu.build()
  flat, meta = workloads.deconvolution_wiring(2, 1, 10, 10, 3, 3, 3)
  g = lbp.graph_from_flat(flat)
  x = workloads.deconvolution_problem(2, 10, 10, n_feat_true=2, feat_true=3, seed=1)
  ev = workloads.deconvolution_evidence(meta, x, x_logit=-8.0)
  rng = np.random.default_rng(2)
  ev = (ev + rng.gumbel(size=ev.shape) * 0.5).astype(np.float32)
  ft = rng.normal(size=flat.num_edge_states).astype(np.float32)
  ref, ref_d = lbp.run_with_diffs(g, np.zeros(0, np.float32), ft, ev, 30, 0.5, 0.0, np.float32)
  got, got_d = cpu.run_logical_t0(g, ft, ev, 30, 0.5, threads=3)
  np.testing.assert_array_equal(got, ref)
  np.testing.assert_array_equal(got_d, ref_d)
