"""GPU parity: the HIP engine against the pinned CPU oracle and the reference golden.

Bars: messages and log-beliefs within 1e-5 max-abs of the FP32 oracle (the
reference's op order), the 84-message golden within the reference's own
1e-6, MAP decodes equal; exact bitwise equality where the arithmetic is the
same by construction (execution shapes, repeat runs, split runs).
"""

import numpy as np
import pytest

import graphs
from oracle import lbp
from pgmax_amd import _lib
from pgmax_amd import fgraph
from pgmax_amd import fgroup
from pgmax_amd import infer
from pgmax_amd import vgroup
from pgmax_amd.factor import FAC_TO_VAR_UPDATES

FAC_TO_VAR_UPDATES_KEYS = list(FAC_TO_VAR_UPDATES)

pytestmark = pytest.mark.gpu

GOLDEN = "tests/golden/e2e_sanity.npz"


def _np(x):
  return x.detach().cpu().numpy() if hasattr(x, "detach") else np.asarray(x)


def _oracle(bp, lp, ftov, ev, n, damping, temperature, dtype=np.float32):
  g = lbp.graph_from_context(bp.context)
  return lbp.run_with_diffs(g, _np(lp), _np(ftov), _np(ev), n, damping, temperature, dtype)


def _beliefs(bp, msgs, ev, dtype=np.float32):
  return lbp.get_flat_beliefs(lbp.graph_from_context(bp.context), msgs, _np(ev), dtype)


@pytest.fixture(autouse=True)
def _reset_mode(gpu):
  prev = _lib.set_option("mode", 0)
  yield
  _lib.set_option("mode", prev)


def test_golden_e2e(gpu):
  """Reference tests/test_pgmax.py:416-421 through the GPU engine."""
  golden = np.load(GOLDEN)
  _, bp_state, grid_vars, additional_vars = graphs.e2e_sanity_graph()
  bp = infer.build_inferer(bp_state, backend="bp")
  arrays = bp.init()
  out = bp.run(arrays, num_iters=100, temperature=0.0)
  msgs = _np(out.ftov_msgs)
  np.testing.assert_allclose(msgs, golden["final_msgs"], atol=1e-6, rtol=0)
  ref, _ = _oracle(bp, arrays.log_potentials, arrays.ftov_msgs, arrays.evidence, 100, 0.5, 0.0)
  np.testing.assert_array_equal(msgs, ref)  # max-product: adds and maxes only
  dec = infer.decode_map_states(bp.get_beliefs(out))
  for grp, idx, state in zip(golden["map_group"], golden["map_index"], golden["map_state"]):
    if grp == "grid_vars":
      assert int(dec[grid_vars][tuple(idx)]) == state
    else:
      assert int(dec[additional_vars][tuple(int(i) for i in idx)]) == state


@pytest.mark.parametrize("mode", [1, 2])
@pytest.mark.parametrize("temperature", [1.0, 0.0])
def test_ising_torus_1000_iters(gpu, mode, temperature):
  """Ising 50x50 torus, J ~ N(0,1), unaries N(0,0.5), damping 0.5, 1000 iterations."""
  _lib.set_option("mode", mode)
  fg, variables, evidence = graphs.ising_torus(side=50, seed=0)
  bp = infer.BP(fg.bp_state, temperature=temperature)
  arrays = bp.init(evidence_updates={variables: evidence})
  out, deltas = bp.run_with_diffs(arrays, num_iters=1000, damping=0.5)
  ref, ref_d = _oracle(bp, arrays.log_potentials, arrays.ftov_msgs, arrays.evidence,
                       1000, 0.5, temperature)
  b_gpu = _np(bp.context.get_flat_beliefs(out))
  b_ref = _beliefs(bp, ref, arrays.evidence)
  err32 = np.abs(b_gpu - b_ref).max()
  ref64, _ = _oracle(bp, arrays.log_potentials, arrays.ftov_msgs, arrays.evidence,
                     1000, 0.5, temperature, np.float64)
  err64 = np.abs(b_gpu - _beliefs(bp, ref64, arrays.evidence, np.float64)).max()
  print(f"ising T={temperature} mode={mode}: |gpu-fp32 oracle|={err32:.2e} "
        f"|gpu-fp64 oracle|={err64:.2e}")
  assert err32 <= 1e-5
  np.testing.assert_allclose(_np(deltas), ref_d, atol=1e-5)
  np.testing.assert_array_equal(_np(infer.decode_map_states(bp.get_beliefs(out))[variables]),
                                np.argmax(b_ref.reshape(50, 50, 2), -1))


@pytest.mark.parametrize("temperature", [0.0, 1.0])
def test_rbm_batch(gpu, temperature):
  """Batched RBM (one W draw per batch row), vs the oracle per batch element."""
  fg, hidden, visible, lp, ev = graphs.rbm(n_visible=60, n_hidden=20, batch=4, seed=1)
  bp = infer.BP(fg.bp_state, temperature=temperature)
  arrays = bp.init()
  arrays = infer.BPArrays(log_potentials=lp, ftov_msgs=arrays.ftov_msgs, evidence=ev)
  out, deltas = bp.run_with_diffs(arrays, num_iters=50, damping=0.5)
  msgs = _np(out.ftov_msgs)
  assert msgs.shape == (4, 20 * 60 * 4)
  for b in range(4):
    ref, ref_d = _oracle(bp, lp[b], np.zeros(msgs.shape[1]), ev[b], 50, 0.5, temperature)
    np.testing.assert_allclose(
        _np(bp.context.get_flat_beliefs(infer.BPArrays(lp[b], out.ftov_msgs[b], ev[b]))),
        _beliefs(bp, ref, ev[b]), atol=1e-5)
    np.testing.assert_allclose(_np(deltas[b]), ref_d, atol=1e-5)


@pytest.mark.parametrize("mode", [1, 2])
def test_uneven_potts_union(gpu, mode):
  """Disjoint Potts grids of different sizes and state counts (2..8) in one graph."""
  _lib.set_option("mode", mode)
  specs = [(12, 2, 0.5), (9, 5, 1.2), (16, 8, 0.3), (7, 3, 1.5), (20, 4, 0.9), (64, 2, 0.7)]
  fg, groups, ev = graphs.potts_union(specs, seed=3)
  bp = infer.BP(fg.bp_state, temperature=1.0)
  parts = np.split(ev, np.cumsum([s * s * k for s, k, _ in specs])[:-1])
  arrays = bp.init(evidence_updates={
      g: e.reshape(s, s, k) for g, e, (s, k, _) in zip(groups, parts, specs)})
  out, deltas = bp.run_with_diffs(arrays, num_iters=200, damping=0.5)
  ref, ref_d = _oracle(bp, arrays.log_potentials, arrays.ftov_msgs, arrays.evidence,
                       200, 0.5, 1.0)
  np.testing.assert_allclose(_np(bp.context.get_flat_beliefs(out)),
                             _beliefs(bp, ref, arrays.evidence), atol=1e-5)
  np.testing.assert_allclose(_np(deltas), ref_d, atol=1e-5)


@pytest.mark.parametrize("kind", ["or", "and", "pool"])
@pytest.mark.parametrize("seed", range(8))
def test_logical_factors(gpu, kind, seed):
  """OR / AND / Pool closed forms: GPU vs FP32 oracle, and vs the EnumFactor expansion.

  Temperatures and tolerances as in the reference tests/factor/test_or.py:52-87
  and test_pool.py:25-69 for the expansion comparison.
  """
  rng = np.random.default_rng(100 + seed)
  temperature = [0.0, 0.001, float(rng.uniform(0.1, 0.5)), float(rng.uniform(0.5, 1.0))][seed % 4]
  atol_equiv = 1e-5 if (seed % 4 in (0, 3) or kind == "pool") else 5e-3
  pair, (ev_p, ev_c), (ft_p, ft_c) = graphs.logical_pair(kind, seed)
  beliefs = []
  for fg, parents, children, _ in pair:
    bp = infer.BP(fg.bp_state, temperature=temperature)
    ftov_updates = {parents[k]: ft_p[k] for k in range(ft_p.shape[0])}
    ftov_updates.update({children[k]: ft_c[k] for k in range(ft_c.shape[0])})
    arrays = bp.init(evidence_updates={parents: ev_p, children: ev_c},
                     ftov_msgs_updates=ftov_updates)
    out = bp.run(arrays, num_iters=5, temperature=temperature)
    ref, _ = _oracle(bp, arrays.log_potentials, arrays.ftov_msgs, arrays.evidence, 5, 0.5,
                     temperature)
    b_gpu = _np(bp.context.get_flat_beliefs(out))
    b_ref = _beliefs(bp, ref, arrays.evidence)
    tol = 1e-5 if temperature in (0.0,) or temperature >= 0.5 else 1e-4
    np.testing.assert_allclose(b_gpu, b_ref, atol=tol)
    beliefs.append(b_gpu)
  np.testing.assert_allclose(beliefs[0], beliefs[1], atol=atol_equiv)


def test_clipping_divergent_graph(gpu):
  """Reference tests/test_clipping.py: damping 0, 100 iterations, decodes all 1."""
  fg, v, ev = graphs.divergent_complete_graph()
  bp = infer.BP(fg.bp_state)
  arrays = bp.init(evidence_updates={v: ev})
  out, _ = bp.run_with_diffs(arrays, 100, damping=0)
  decoded = _np(infer.decode_map_states(bp.get_beliefs(out))[v])
  assert np.all(decoded == 1)
  ref, _ = _oracle(bp, arrays.log_potentials, arrays.ftov_msgs, arrays.evidence, 100, 0.0, 0.0)
  np.testing.assert_array_equal(_np(out.ftov_msgs), ref)


def test_shapes_modes_repeat_and_split_runs_are_bitwise_equal(gpu):
  fg, variables, evidence = graphs.ising_torus(side=24, seed=4)
  bp = infer.BP(fg.bp_state, temperature=1.0)
  arrays = bp.init(evidence_updates={variables: evidence})
  _lib.set_option("mode", 1)
  a = _np(bp.run(arrays, 40).ftov_msgs)
  a2 = _np(bp.run(arrays, 40).ftov_msgs)
  _lib.set_option("mode", 2)
  b = _np(bp.run(arrays, 40).ftov_msgs)
  c = _np(bp.run(bp.run(arrays, 15), 25).ftov_msgs)
  np.testing.assert_array_equal(a, a2)
  np.testing.assert_array_equal(a, b)
  np.testing.assert_array_equal(a, c)


@pytest.mark.parametrize("temperature", [0.0, 1.0, 0.4])
def test_fused_iteration_is_bitwise_equal(gpu, temperature):
  """Recomputing variable sums on the factor side changes no bit (grids, Potts)."""
  specs = [(10, 2, 0.5), (9, 5, 1.2), (12, 8, 0.3), (30, 2, 0.9)]
  fg, groups, ev = graphs.potts_union(specs, seed=9)
  bp = infer.BP(fg.bp_state, temperature=temperature)
  parts = np.split(ev, np.cumsum([s * s * k for s, k, _ in specs])[:-1])
  arrays = bp.init(evidence_updates={
      g: e.reshape(s, s, k) for g, e, (s, k, _) in zip(groups, parts, specs)})
  outs = []
  for fuse in (0, 1):
    prev = _lib.set_option("fuse", fuse)
    try:
      for mode in (1, 2):
        _lib.set_option("mode", mode)
        out, d = bp.run_with_diffs(arrays, 60)
        outs.append((_np(out.ftov_msgs), _np(d)))
    finally:
      _lib.set_option("fuse", prev)
  for m, d in outs[1:]:
    np.testing.assert_array_equal(m, outs[0][0])
    np.testing.assert_array_equal(d, outs[0][1])


def test_num_iters_zero_runs_one_iteration(gpu):
  """pgmax/infer/bp.py:142-146: num_iters <= 1 runs exactly one update."""
  fg, variables, evidence = graphs.ising_torus(side=6, seed=5)
  bp = infer.BP(fg.bp_state, temperature=0.0)
  arrays = bp.init(evidence_updates={variables: evidence})
  out0, d0 = bp.run_with_diffs(arrays, 0)
  out1, d1 = bp.run_with_diffs(arrays, 1)
  assert d0.shape == (1,)
  np.testing.assert_array_equal(_np(out0.ftov_msgs), _np(out1.ftov_msgs))


@pytest.mark.parametrize("temperature", [0.0, 0.3, 1.0])
def test_registry_ops_match_oracle(gpu, temperature):
  """FAC_TO_VAR_UPDATES entries (the plugin API) on random vtof vs the oracle's functions."""
  rng = np.random.default_rng(7)
  names = ["enum", "or", "and", "pool"]
  for kind in ("or", "and", "pool"):
    pair, _, _ = graphs.logical_pair(kind, 3)
    fg = pair[0][0]
    ctx = infer.BP(fg.bp_state).context
    for t, name in zip(FAC_TO_VAR_UPDATES, names):
      m0, m1 = ctx.factor_type_to_msgs_range[t]
      if m0 == m1:
        continue
      p0, p1 = ctx.factor_type_to_potentials_range[t]
      vt = rng.normal(size=m1 - m0).astype(np.float32)
      lp = np.asarray(fg.bp_state.log_potentials.value[p0:p1], dtype=np.float32)
      args = ctx.inference_arguments[t]
      got = _np(FAC_TO_VAR_UPDATES[t](vtof_msgs=vt, log_potentials=lp,
                                      temperature=temperature, normalize=True, **args))
      if name == "enum":
        want = lbp.pass_enum_fac_to_var_messages(
            vt, args["factor_configs_indices"], args["factor_configs_edge_states"], lp,
            args["num_val_configs"], temperature)
      elif name == "pool":
        want = lbp.pass_pool_fac_to_var_messages(
            vt, args["pool_choices_factor_indices"], args["pool_choices_msg_indices"],
            args["pool_indicators_edge_states"], temperature)
      else:
        want = lbp.pass_logical_fac_to_var_messages(
            vt, args["parents_factor_indices"], args["parents_msg_indices"],
            args["children_edge_states"], args["edge_states_offset"], temperature)
      np.testing.assert_allclose(got, want, atol=1e-4 if 0 < temperature < 0.5 else 1e-5)


def test_beliefs_and_decode_batched(gpu):
  fg, hidden, visible, lp, ev = graphs.rbm(n_visible=30, n_hidden=10, batch=3, seed=2)
  bp = infer.BP(fg.bp_state, temperature=0.0)
  arrays = infer.BPArrays(lp, bp.init().ftov_msgs, ev)
  out = bp.run(arrays, 30)
  flat = bp.context.get_flat_beliefs(out)
  dec = _np(bp.context.engine.decode(flat))
  want = np.stack([np.argmax(_np(flat[b]).reshape(-1, 2), -1) for b in range(3)])
  np.testing.assert_array_equal(dec, want)
  beliefs = bp.get_beliefs(out)
  assert tuple(beliefs[hidden].shape) == (3, 10, 2)
  marg = infer.get_marginals(beliefs)
  np.testing.assert_allclose(_np(marg[visible]).sum(-1), 1.0, atol=1e-6)


def test_mixed_batch_strides(gpu):
  """Shared log-potentials and initial messages (stride 0), per-element
  evidence: each element equals its own oracle run (jax.vmap with
  in_axes=(None, None, 0) semantics)."""
  fg, _, _, lp, ev = graphs.rbm(n_visible=40, n_hidden=12, batch=3, seed=5)
  bp = infer.BP(fg.bp_state, temperature=1.0)
  rng = np.random.default_rng(6)
  ft0 = rng.normal(size=bp.init().ftov_msgs.shape[-1]).astype(np.float32)
  eng = bp.context.engine
  out, deltas = eng.run(lp[0], ft0, ev, 30, 0.5, 1.0)
  assert tuple(out.shape) == (3, ft0.size)
  for b in range(3):
    ref, ref_d = _oracle(bp, lp[0], ft0, ev[b], 30, 0.5, 1.0)
    np.testing.assert_array_equal(_np(out[b]), ref)
    np.testing.assert_array_equal(_np(deltas[b]), ref_d)


def _rbm_runs(fg, lp, ev, iters, temperature):
  """Per-phase runs with the factor-side variable sums on (fuse 1) and off,
  returning each run's (messages, deltas, kernels launched)."""
  bp = infer.BP(fg.bp_state, temperature=temperature)
  eng = bp.context.engine
  ft0 = np.zeros(eng.graph.num_es, np.float32)
  outs = []
  prev_mode = _lib.set_option("mode", 2)
  prev_fuse = _lib.set_option("fuse", 1)
  try:
    for fuse in (1, 0):
      _lib.set_option("fuse", fuse)
      _lib.profile_kernels()
      _lib.set_option("profile", 2)
      try:
        out, d = eng.run(lp, ft0, ev, iters, 0.5, temperature)
        outs.append((_np(out), _np(d), _lib.profile_kernels()))
      finally:
        _lib.set_option("profile", 0)
        _lib.profile_read()
  finally:
    _lib.set_option("mode", prev_mode)
    _lib.set_option("fuse", prev_fuse)
  return bp, ft0, outs


@pytest.mark.parametrize("n_visible,n_hidden,visible_major,temperature", [
    (70, 40, False, 0.0), (70, 40, False, 1.0), (70, 40, True, 0.4),
    (100, 64, False, 0.0), (33, 1100, False, 1.0)])
def test_bipartite_iteration_matches_oracle(gpu, n_visible, n_hidden, visible_major,
                                            temperature):
  """Complete bipartite binary graphs (RBMs whose units all have more than 32
  edges) run one k_bipartite launch per iteration: bitwise equal to the
  per-phase iteration and to the FP32 oracle, for ragged tiles, both factor
  orders and rows of more than 32 chunks (1,100 edges)."""
  batch = 2
  fg, _, _, lp, ev = graphs.rbm(n_visible=n_visible, n_hidden=n_hidden, batch=batch, seed=11,
                                std=0.5, visible_major=visible_major)
  lp = lp.astype(np.float32)
  ev = ev.astype(np.float32)
  iters = 25
  bp, ft0, outs = _rbm_runs(fg, lp, ev, iters, temperature)
  (m1, d1, k1), (m0, d0, k0) = outs
  assert k1["bipartite"][1] == iters
  assert "bipartite" not in k0 and "fac_phase_enum2x2" in k0
  np.testing.assert_array_equal(m1, m0)
  np.testing.assert_array_equal(d1, d0)
  for b in range(batch):
    ref, ref_d = _oracle(bp, lp[b], ft0, ev[b], iters, 0.5, temperature)
    np.testing.assert_array_equal(m1[b], ref)
    np.testing.assert_array_equal(d1[b], ref_d)


@pytest.mark.parametrize("n_visible,n_hidden,visible_major,temperature,tile,block", [
    (70, 40, False, 0.0, 32, 256), (70, 40, False, 1.0, 64, 256), (100, 64, False, 0.4, 128, 512),
    (95, 63, False, 1.0, 32, 128), (33, 1100, False, 1.0, 64, 512), (70, 40, True, 1.0, 32, 256)])
def test_bipartite_unary_form_matches_oracle(gpu, n_visible, n_hidden, visible_major, temperature,
                                             tile, block):
  """The reference's RBM form (benchmark/rbm_lib.py:140-168): biases as unary
  Enum factors added before the pairwise factors, so each unit's first edge
  is its unary factor. The bipartite iteration runs them as row 0 and column
  0 of its tile grid, which keeps each unit's 32-edge runs inside a tile:
  bitwise equal to the per-phase iteration and the FP32 oracle, for ragged
  tiles of every shape, units of more than 32 chunks, non-finite initial
  messages, potentials and evidence. With the pairwise factors running over
  visible units first, the unary factors' order (hidden first) no longer
  matches the rows, and the per-phase kernels run, still bitwise."""
  fg, _, _, lp, ev = graphs.rbm(n_visible=n_visible, n_hidden=n_hidden, batch=2, seed=12,
                                std=0.5, visible_major=visible_major, unary_factors=True)
  lp = lp.astype(np.float32)
  ev = np.random.default_rng(13).normal(0, 0.3, size=ev.shape).astype(np.float32)
  nu = n_visible + n_hidden
  lp[1, 3] = np.inf  # a unary factor's table
  lp[0, 2 * nu - 1] = -1e7
  lp[1, 2 * nu + 4 * 7 + 3] = np.nan  # a pairwise table
  ev[1, 9] = np.nan
  bp = infer.BP(fg.bp_state, temperature=temperature)
  eng = bp.context.engine
  rng = np.random.default_rng(14)
  ft0 = rng.normal(0, 1, size=eng.graph.num_es).astype(np.float32)
  ft0[[1, 2 * nu + 5]] = np.inf
  ft0[[6, 2 * nu + 40]] = -1e38
  outs = []
  prev = {k: _lib.set_option(k, 0) for k in ("mode", "fuse", "tile", "block")}
  try:
    _lib.set_option("mode", 2)
    _lib.set_option("tile", tile)
    _lib.set_option("block", block)
    for fuse in (1, 0):
      _lib.set_option("fuse", fuse)
      _lib.profile_kernels()
      _lib.set_option("profile", 2)
      try:
        out, d = eng.run(lp, ft0, ev, 9, 0.5, temperature)
        outs.append((_np(out), _np(d), _lib.profile_kernels()))
      finally:
        _lib.set_option("profile", 0)
        _lib.profile_read()
  finally:
    for k, v in prev.items():
      _lib.set_option(k, v)
  (m1, d1, k1), (m0, d0, k0) = outs
  if visible_major:
    assert "bipartite" not in k1
  else:
    assert k1["bipartite"][1] == 9
  assert "bipartite" not in k0
  np.testing.assert_array_equal(m1, m0)
  np.testing.assert_array_equal(d1, d0)
  for b in range(2):
    ref, ref_d = _oracle(bp, lp[b], ft0, ev[b], 9, 0.5, temperature)
    np.testing.assert_array_equal(m1[b], ref)
    np.testing.assert_array_equal(d1[b], ref_d)


@pytest.mark.parametrize("temperature", [0.0, 1.0])
@pytest.mark.parametrize("iters", [1, 3])
def test_bipartite_nonfinite_inputs(gpu, temperature, iters):
  """The bipartite iteration stores one float per binary message. Initial
  messages with +inf (entry normalisation leaves the pair (NaN, -1e32)) and
  -1e38 (clipped to -1e32), and NaN evidence, decode to what the per-phase
  iteration and the FP32 oracle produce, NaNs included."""
  fg, _, _, lp, ev = graphs.rbm(n_visible=70, n_hidden=40, batch=2, seed=17, std=0.5)
  bp = infer.BP(fg.bp_state, temperature=temperature)
  eng = bp.context.engine
  lp = lp.astype(np.float32)
  ev = ev.astype(np.float32)
  ev[1, 7] = np.nan
  rng = np.random.default_rng(18)
  ft0 = rng.normal(0, 1, size=eng.graph.num_es).astype(np.float32)
  ft0[[4, 101, 2002]] = np.inf
  ft0[[9, 555]] = -1e38
  outs = []
  prev_mode = _lib.set_option("mode", 2)
  prev_fuse = _lib.set_option("fuse", 1)
  try:
    for fuse in (1, 0):
      _lib.set_option("fuse", fuse)
      out, d = eng.run(lp, ft0, ev, iters, 0.5, temperature)
      outs.append((_np(out), _np(d)))
  finally:
    _lib.set_option("mode", prev_mode)
    _lib.set_option("fuse", prev_fuse)
  (m1, d1), (m0, d0) = outs
  assert np.isnan(m1).any()
  np.testing.assert_array_equal(m1, m0)
  np.testing.assert_array_equal(d1, d0)
  for b in range(2):
    ref, ref_d = _oracle(bp, lp[b], ft0, ev[b], iters, 0.5, temperature)
    np.testing.assert_array_equal(m1[b], ref)
    np.testing.assert_array_equal(d1[b], ref_d)


@pytest.mark.parametrize("temperature", [0.0, 1.0])
@pytest.mark.parametrize("fuse", [0, 1])
def test_packed_messages_per_phase(gpu, temperature, fuse):
  """Per-phase runs of a binary grid (2x2 factors, 4 edges per variable) keep
  one float per message ("pack" 1): bitwise equal to two floats ("pack" 0)
  and to the FP32 oracle, with finite inputs and with +inf / -1e38 initial
  messages and NaN evidence."""
  fg, v, ev1 = graphs.ising_torus(side=20, seed=31)
  bp = infer.BP(fg.bp_state, temperature=temperature)
  eng = bp.context.engine
  g = eng.graph
  rng = np.random.default_rng(32)
  lp = np.asarray(bp.init().log_potentials.cpu().numpy())
  ev = rng.normal(0, 0.5, size=(3, g.num_vs)).astype(np.float32)
  ev[2, 11] = np.nan
  ft0 = np.tile(rng.normal(0, 1, size=g.num_es).astype(np.float32), (3, 1))
  ft0[1, [6, 300]] = np.inf
  ft0[1, [13, 777]] = -1e38
  outs = []
  prev = {k: _lib.set_option(k, 0) for k in ("mode", "fuse", "pack")}
  try:
    _lib.set_option("mode", 2)
    _lib.set_option("fuse", fuse)
    for pack in (1, 0):
      _lib.set_option("pack", pack)
      out, d = eng.run(lp, ft0, ev, 9, 0.5, temperature)
      outs.append((_np(out), _np(d)))
  finally:
    for k, val in prev.items():
      _lib.set_option(k, val)
  (m1, d1), (m0, d0) = outs
  np.testing.assert_array_equal(m1, m0)
  np.testing.assert_array_equal(d1, d0)
  assert np.isnan(m1[2]).any() and not np.isnan(m1[0]).any()
  for b in range(3):
    ref, ref_d = _oracle(bp, lp, ft0[b], ev[b], 9, 0.5, temperature)
    np.testing.assert_array_equal(m1[b], ref)
    np.testing.assert_array_equal(d1[b], ref_d)


@pytest.mark.parametrize("temperature", [0.0, 1.0])
def test_bipartite_interleaved_units(gpu, temperature):
  """Row and column units interleaved in one variable group (rows at even
  ids, columns at odd ids and then the tail): the per-unit tables are not
  affine, so the bipartite kernel loads them; still bitwise equal to the
  per-phase iteration and the FP32 oracle."""
  v = vgroup.NDVarArray(num_states=2, shape=(110,))
  fg = fgraph.FactorGraph(variable_groups=[v])
  rows = list(range(0, 80, 2))
  cols = list(range(1, 80, 2)) + list(range(80, 110))
  vff = [[v[a], v[c]] for a in rows for c in cols]
  fg.add_factors(fgroup.PairwiseFactorGroup(variables_for_factors=vff))
  rng = np.random.default_rng(41)
  lp = rng.normal(0, 0.5, size=(2, len(vff) * 4)).astype(np.float32)
  ev = rng.normal(0, 0.5, size=(2, 220)).astype(np.float32)
  bp, ft0, outs = _rbm_runs(fg, lp, ev, 15, temperature)
  (m1, d1, k1), (m0, d0, _) = outs
  assert k1["bipartite"][1] == 15
  np.testing.assert_array_equal(m1, m0)
  np.testing.assert_array_equal(d1, d0)
  for b in range(2):
    ref, ref_d = _oracle(bp, lp[b], ft0, ev[b], 15, 0.5, temperature)
    np.testing.assert_array_equal(m1[b], ref)
    np.testing.assert_array_equal(d1[b], ref_d)


@pytest.mark.parametrize("temperature", [0.0, 1.0])
def test_autotuned_bipartite_run_matches_oracle(gpu, temperature):
  """pgx_autotune over the shapes of a complete bipartite graph (per-phase,
  task loop, bipartite tiles of 32/64/128 columns); the run it selects
  equals the FP32 oracle bit for bit."""
  fg, _, _, lp, ev = graphs.rbm(n_visible=130, n_hidden=70, batch=3, seed=51, std=0.5)
  bp = infer.BP(fg.bp_state, temperature=temperature)
  eng = bp.context.engine
  lp = lp.astype(np.float32)
  ev = ev.astype(np.float32)
  ft0 = np.zeros(eng.graph.num_es, np.float32)
  chosen = eng.autotune(lp, ft0, ev, 0.5, temperature, probe_iters=3)
  assert set(chosen) >= {"mode", "block", "fuse", "vec", "tile"}
  out, d = eng.run(lp, ft0, ev, 12, 0.5, temperature)
  for b in range(3):
    ref, ref_d = _oracle(bp, lp[b], ft0, ev[b], 12, 0.5, temperature)
    np.testing.assert_array_equal(_np(out[b]), ref)
    np.testing.assert_array_equal(_np(d[b]), ref_d)


def test_bipartite_path_needs_the_exact_structure(gpu):
  """An RBM with one extra visible-visible factor is not complete bipartite:
  the per-phase kernels run and still match the oracle bit for bit."""
  hidden = vgroup.NDVarArray(num_states=2, shape=(40,))
  visible = vgroup.NDVarArray(num_states=2, shape=(50,))
  fg = fgraph.FactorGraph(variable_groups=[hidden, visible])
  vff = [[hidden[i], visible[j]] for i in range(40) for j in range(50)]
  vff.append([visible[0], visible[1]])
  fg.add_factors(fgroup.PairwiseFactorGroup(variables_for_factors=vff))
  rng = np.random.default_rng(3)
  lp = rng.normal(0, 0.5, size=(2, len(vff) * 4)).astype(np.float32)
  ev = rng.normal(0, 0.5, size=(2, 180)).astype(np.float32)
  bp, ft0, outs = _rbm_runs(fg, lp, ev, 20, 0.0)
  for m, d, k in outs:
    assert "bipartite" not in k
    for b in range(2):
      ref, ref_d = _oracle(bp, lp[b], ft0, ev[b], 20, 0.5, 0.0)
      np.testing.assert_array_equal(m[b], ref)
      np.testing.assert_array_equal(d[b], ref_d)


@pytest.mark.parametrize("fuse,temperature", [(0, 0.0), (1, 1.0), (0, 0.4)])
def test_many_workgroup_deltas_match_oracle(gpu, fuse, temperature):
  """A graph of >= 256 workgroups per iteration folds its per-launch deltas
  into slots reduced after the run: deltas and messages still equal the FP32
  oracle bit for bit (130x130 torus, 33,800 factors, batch 2)."""
  fg, v, ev1 = graphs.ising_torus(side=130, seed=12)
  bp = infer.BP(fg.bp_state, temperature=temperature)
  eng = bp.context.engine
  g = eng.graph
  rng = np.random.default_rng(13)
  lp = np.asarray(bp.init().log_potentials.cpu().numpy())
  evb = rng.normal(0, 0.5, size=(2, g.num_vs)).astype(np.float32)
  ft0 = rng.normal(0, 0.3, size=g.num_es).astype(np.float32)
  prev_mode = _lib.set_option("mode", 2)
  prev_fuse = _lib.set_option("fuse", fuse)
  try:
    out, d = eng.run(lp, ft0, evb, 8, 0.5, temperature)
  finally:
    _lib.set_option("mode", prev_mode)
    _lib.set_option("fuse", prev_fuse)
  for b in range(2):
    ref, ref_d = _oracle(bp, lp, ft0, evb[b], 8, 0.5, temperature)
    np.testing.assert_array_equal(_np(out[b]), ref)
    np.testing.assert_array_equal(_np(d[b]), ref_d)


@pytest.mark.parametrize("mode,fuse", [(1, 0), (2, 0), (2, 1)])
@pytest.mark.parametrize("temperature", [0.0, 1.0, 0.4])
def test_square_pairwise_tables_match_oracle(gpu, mode, fuse, temperature):
  """Square pairwise tables of 3..8 states run exact-size code (whole edges,
  table rows and variable sums in vector accesses): messages and deltas equal
  the FP32 oracle bit for bit, with random potentials, evidence and initial
  messages (grids of 2..8 states in one graph, batch 2)."""
  specs = [(6, 3, 0.8), (5, 4, 1.1), (7, 5, 0.6), (4, 6, 1.3), (5, 7, 0.4), (6, 8, 0.9),
           (5, 2, 0.7)]
  fg, _, _ = graphs.potts_union(specs, seed=11)
  bp = infer.BP(fg.bp_state, temperature=temperature)
  eng = bp.context.engine
  g = eng.graph
  rng = np.random.default_rng(17)
  n_lp = bp.init().log_potentials.shape[0]
  lp = rng.normal(0, 1, size=(2, n_lp)).astype(np.float32)
  evb = rng.normal(0, 0.5, size=(2, g.num_vs)).astype(np.float32)
  ft0 = rng.normal(0, 0.3, size=g.num_es).astype(np.float32)
  prev_mode = _lib.set_option("mode", mode)
  prev_fuse = _lib.set_option("fuse", fuse)
  try:
    out, d = eng.run(lp, ft0, evb, 12, 0.5, temperature)
  finally:
    _lib.set_option("mode", prev_mode)
    _lib.set_option("fuse", prev_fuse)
  for b in range(2):
    ref, ref_d = _oracle(bp, lp[b], ft0, evb[b], 12, 0.5, temperature)
    np.testing.assert_array_equal(_np(out[b]), ref)
    np.testing.assert_array_equal(_np(d[b]), ref_d)


def _general_enum_graph(kind):
  """Graphs of general Enum factors: "tri" 3-variable 5-state factors with all
  125 configs on a 6x6 torus; "sparse" pairwise 40-state factors with 300
  random configs (80 edge states, more than a wavefront); "big" 3-variable
  11-state factors (1,331 configs, beyond the wavefront LDS: thread path)."""
  rng = np.random.default_rng(21)
  if kind == "sparse":
    v = vgroup.NDVarArray(num_states=40, shape=(12,))
    fg = fgraph.FactorGraph(variable_groups=v)
    cfg = np.unique(rng.integers(0, 40, size=(400, 2)), axis=0)[:300]
    vff = [[v[i], v[(i + 1) % 12]] for i in range(12)]
  else:
    s = 5 if kind == "tri" else 11
    side = 6 if kind == "tri" else 3
    v = vgroup.NDVarArray(num_states=s, shape=(side, side))
    fg = fgraph.FactorGraph(variable_groups=v)
    cfg = np.array([[a, b, c] for a in range(s) for b in range(s) for c in range(s)])
    vff = [[v[i, j], v[i, (j + 1) % side], v[(i + 1) % side, j]]
           for i in range(side) for j in range(side)]
  fg.add_factors(fgroup.EnumFactorGroup(variables_for_factors=vff, factor_configs=cfg))
  return fg


@pytest.mark.parametrize("kind", ["tri", "sparse", "big"])
@pytest.mark.parametrize("temperature", [0.0, 1.0, 0.4])
def test_general_enum_wavefront_path(gpu, kind, temperature):
  """General Enum factors in the per-phase shape (one wavefront per factor,
  LDS-staged summaries and per-edge-state config lists) equal the task-loop
  shape (one thread per factor) and the FP32 oracle bit for bit."""
  fg = _general_enum_graph(kind)
  bp = infer.BP(fg.bp_state, temperature=temperature)
  eng = bp.context.engine
  g = eng.graph
  rng = np.random.default_rng(22)
  lp = rng.normal(0, 1, size=(2, g.num_lp)).astype(np.float32)
  ev = rng.normal(0, 1, size=(2, g.num_vs)).astype(np.float32)
  ft0 = rng.normal(0, 0.5, size=g.num_es).astype(np.float32)
  outs = []
  prev = _lib.set_option("mode", 1)
  try:
    for mode in (1, 2):
      _lib.set_option("mode", mode)
      out, d = eng.run(lp, ft0, ev, 12, 0.5, temperature)
      outs.append((_np(out), _np(d)))
  finally:
    _lib.set_option("mode", prev)
  np.testing.assert_array_equal(outs[0][0], outs[1][0])
  np.testing.assert_array_equal(outs[0][1], outs[1][1])
  for b in range(2):
    ref, ref_d = _oracle(bp, lp[b], ft0, ev[b], 12, 0.5, temperature)
    np.testing.assert_array_equal(outs[1][0][b], ref)
    np.testing.assert_array_equal(outs[1][1][b], ref_d)


@pytest.mark.parametrize("temperature", [0.0, 1.0])
def test_cooperative_shape_is_bitwise_equal(gpu, temperature):
  """One cooperative launch for all iterations (mode 3) == per-phase launches."""
  fg, v, ev = graphs.ising_torus(side=20, seed=4)
  bp = infer.BP(fg.bp_state, temperature=temperature)
  eng = bp.context.engine
  g = eng.graph
  rng = np.random.default_rng(8)
  lp = np.asarray(bp.init().log_potentials.cpu().numpy())
  evb = rng.normal(0, 0.5, size=(3, g.num_vs)).astype(np.float32)
  ft0 = np.zeros(g.num_es, np.float32)
  outs = []
  for mode in (2, 3):
    _lib.set_option("mode", mode)
    out, d = eng.run(lp, ft0, evb, 60, 0.5, temperature)
    outs.append((_np(out), _np(d)))
  np.testing.assert_array_equal(outs[0][0], outs[1][0])
  np.testing.assert_array_equal(outs[0][1], outs[1][1])


def test_graph_replays_match_eager_runs(gpu):
  """With the "graph" option, pgx_run captures its launches once and replays
  them; every replay (with eager work and synchronisation in between) equals
  the eager run, bit for bit."""
  import torch
  fg, v, ev = graphs.ising_torus(side=12, seed=9)
  bp = infer.BP(fg.bp_state, temperature=1.0)
  g = bp.context.engine.graph
  lib = _lib.load()
  rng = np.random.default_rng(10)
  lp = torch.as_tensor(np.asarray(bp.init().log_potentials.cpu().numpy()), device="cuda")
  evs = [torch.as_tensor(rng.normal(0, 0.5, size=g.num_vs).astype(np.float32), device="cuda")
         for _ in range(3)]
  ev_buf = torch.empty_like(evs[0])
  ft0 = torch.zeros(g.num_es, device="cuda")
  out = torch.empty(g.num_es, device="cuda")
  deltas = torch.empty(40, device="cuda")
  stream = torch.cuda.Stream()

  def run():
    with torch.cuda.stream(stream):
      _lib.check(lib.pgx_run(g.handle, _lib.ptr(lp), 0, _lib.ptr(ev_buf), 0, _lib.ptr(ft0), 0,
                             _lib.ptr(out), _lib.ptr(deltas), 1, 40, 0.5, 1.0,
                             int(stream.cuda_stream)), "pgx_run")
    stream.synchronize()
    return out.cpu().numpy().copy(), deltas.cpu().numpy().copy()

  eager = []
  for e in evs:
    ev_buf.copy_(e)
    eager.append(run())
  prev = _lib.set_option("graph", 1)
  try:
    for rep in range(3):
      for k, e in enumerate(evs):
        ev_buf.copy_(e)
        torch.cuda.synchronize()
        got = run()
        np.testing.assert_array_equal(got[0], eager[k][0])
        np.testing.assert_array_equal(got[1], eager[k][1])
        _ = (torch.randn(1000, device="cuda") * 2).sum().item()  # eager work between replays
  finally:
    _lib.set_option("graph", prev)


def test_batch_beyond_one_grid_dimension(gpu):
  """A batch of 70,000 elements (more than one launch's grid dimension holds)
  runs in parts and matches the elements run alone."""
  import torch
  fg, v, _ = graphs.ising_torus(side=3, seed=2)
  bp = infer.BP(fg.bp_state, temperature=1.0)
  eng = bp.context.engine
  g = eng.graph
  B = 70000
  rng = np.random.default_rng(3)
  ev = torch.as_tensor(rng.normal(0, 0.5, size=(B, g.num_vs)).astype(np.float32), device="cuda")
  lp = bp.init().log_potentials
  ft0 = torch.zeros(g.num_es, device="cuda")
  out, d = eng.run(lp, ft0, ev, 10, 0.5, 1.0)
  for b in (0, 65534, 65535, 69999):
    o1, d1 = eng.run(lp, ft0, ev[b], 10, 0.5, 1.0)
    np.testing.assert_array_equal(_np(out[b]), _np(o1))
    np.testing.assert_array_equal(_np(d[b]), _np(d1))
  bel = eng.beliefs(ev, out)
  np.testing.assert_array_equal(_np(bel[69999]), _np(eng.beliefs(ev[69999], out[69999])))


@pytest.mark.parametrize("temperature", [0.0, 0.3, 1.0])
@pytest.mark.parametrize("normalize", [True, False])
def test_registry_ops_ties_and_unnormalized(gpu, temperature, normalize):
  """OR / AND / Pool updates (the plugin API) on inputs full of ties (every
  message drawn from {-1, 0, 1}, so most factors have several largest
  differences) with and without normalisation, against the oracle's closed
  forms (last argmax among ties, update_utils.py:25-64), bit for bit: same
  operation order and the same FP32 transcendentals."""
  rng = np.random.default_rng(11)
  names = {FAC_TO_VAR_UPDATES_KEYS[1]: "or", FAC_TO_VAR_UPDATES_KEYS[2]: "and",
           FAC_TO_VAR_UPDATES_KEYS[3]: "pool"}
  for kind in ("or", "and", "pool"):
    for seed in (5, 6):
      pair, _, _ = graphs.logical_pair(kind, seed)
      fg = pair[1][0]
      ctx = infer.BP(fg.bp_state).context
      for t, name in names.items():
        m0, m1 = ctx.factor_type_to_msgs_range[t]
        if m0 == m1:
          continue
        vt = rng.integers(-1, 2, size=m1 - m0).astype(np.float32)
        args = ctx.inference_arguments[t]
        got = _np(FAC_TO_VAR_UPDATES[t](vtof_msgs=vt, log_potentials=np.zeros(0, np.float32),
                                        temperature=temperature, normalize=normalize, **args))
        if name == "pool":
          want = lbp.pass_pool_fac_to_var_messages(
              vt, args["pool_choices_factor_indices"], args["pool_choices_msg_indices"],
              args["pool_indicators_edge_states"], temperature, normalize=normalize)
        else:
          want = lbp.pass_logical_fac_to_var_messages(
              vt, args["parents_factor_indices"], args["parents_msg_indices"],
              args["children_edge_states"], args["edge_states_offset"], temperature,
              normalize=normalize)
        err = np.nanmax(np.abs(np.where(np.isfinite(want), got - want, 0)))
        print(f"{kind} seed {seed} T={temperature} normalize={normalize}: max err {err:.3e}")
        np.testing.assert_array_equal(got, want)


def test_decode_takes_the_first_max_among_ties(gpu):
  """MAP decoding on beliefs with exact ties picks the first state with the
  largest belief, as jnp.argmax does (pgmax/infer/bp_state.py decode)."""
  vd = vgroup.VarDict(variable_names=tuple(range(40)), num_states=np.array([2, 3, 5, 4] * 10))
  fg = fgraph.FactorGraph(vd)
  fg.add_factors(fgroup.EnumFactorGroup(variables_for_factors=[[vd[i]] for i in range(40)][:10:4],
                                        factor_configs=np.arange(2)[:, None]))
  bp = infer.BP(fg.bp_state, temperature=0.0)
  rng = np.random.default_rng(2)
  nvs = int(vd.num_states.sum())
  flat = rng.integers(0, 2, size=(3, nvs)).astype(np.float32)
  flat[1] = 0.0  # every state tied
  flat[2, ::7] = np.inf
  import torch  # pylint: disable=g-import-not-at-top
  dec = _np(bp.context.engine.decode(torch.as_tensor(flat, device="cuda")))
  bounds = np.r_[0, np.cumsum(vd.num_states)]
  want = np.array([[int(np.argmax(f[bounds[i]:bounds[i + 1]])) for i in range(40)] for f in flat])
  np.testing.assert_array_equal(dec, want)
  assert np.all(dec[1] == 0)


@pytest.mark.parametrize("temperature", [0.0, 1.0])
def test_fused_beliefs_map_marginals(gpu, temperature):
  """pgx_beliefs_decode: beliefs equal pgx_get_beliefs', MAP states and
  marginals equal the oracle's (first max, softmax in state order) bit for
  bit, batched, on uneven state counts with ties and non-finite beliefs; the
  per-group results equal get_beliefs / decode_map_states and sum to 1."""
  import torch  # pylint: disable=g-import-not-at-top
  specs = [(6, 2, 0.5), (5, 5, 1.2), (4, 8, 0.3), (7, 3, 1.5)]
  fg, groups, ev = graphs.potts_union(specs, seed=8)
  bp = infer.BP(fg.bp_state, temperature=temperature)
  arrays = bp.init()
  rng = np.random.default_rng(4)
  ev_b = np.stack([ev, np.round(ev), ev]).astype(np.float32)  # row 1: many ties
  ev_b[2, ::13] = np.inf
  ev_b[2, 5] = np.nan
  ev_b[2, 7] = -np.inf
  arrays = infer.BPArrays(arrays.log_potentials, arrays.ftov_msgs, torch.as_tensor(ev_b, device="cuda"))
  out = bp.run(arrays, 20)
  ctx = bp.context
  flat, states, marg = ctx.engine.beliefs_decode(out.evidence, out.ftov_msgs)
  np.testing.assert_array_equal(_np(flat), _np(ctx.get_flat_beliefs(out)))
  np.testing.assert_array_equal(_np(states), _np(ctx.engine.decode(flat)))
  vs0 = np.concatenate([[0], np.cumsum(np.concatenate(
      [np.asarray(g.num_states).reshape(-1) for g in groups]))])
  for b in range(3):
    best, m_ref = lbp.map_and_marginals(_np(flat[b]), vs0)
    np.testing.assert_array_equal(_np(states[b]), best)
    np.testing.assert_array_equal(_np(marg[b]), m_ref)
  beliefs, map_states, marginals = ctx.get_beliefs_map_marginals(out)
  ref_dec = infer.decode_map_states(bp.get_beliefs(out))
  for g in groups:
    np.testing.assert_array_equal(_np(beliefs[g]), _np(bp.get_beliefs(out)[g]))
    np.testing.assert_array_equal(_np(map_states[g])[:2], _np(ref_dec[g])[:2])
    sums = _np(marginals[g])[:2].astype(np.float64).sum(-1)
    np.testing.assert_allclose(sums, 1.0, atol=1e-6)


@pytest.mark.parametrize("kind", ["or", "and"])
@pytest.mark.parametrize("sizes", [(33, 40, 64, 100, 180, 255, 256, 2, 1),
                                   (200, 257, 300, 700, 5)])
@pytest.mark.parametrize("temperature", [0.0, 0.3, 1.0])
def test_many_parent_logical_factors(gpu, kind, sizes, temperature):
  """OR / AND factors of up to 256 parents run one wavefront per factor in
  the per-phase shape, larger ones (blocks of up to 1,024) eight lanes per
  factor; the one-thread-per-factor task loop runs both sequentially. All
  three agree bit for bit (same sums in the same chunked order, same top-2
  with ties to the larger index) and match the FP32 oracle."""
  rng = np.random.default_rng(sum(sizes) + int(10 * temperature))
  npar = np.asarray(sizes)
  parents = vgroup.NDVarArray(num_states=2, shape=(int(npar.sum()),))
  children = vgroup.NDVarArray(num_states=2, shape=(len(sizes),))
  fg = fgraph.FactorGraph(variable_groups=[parents, children])
  cum = np.insert(np.cumsum(npar), 0, 0)
  group = {"or": fgroup.ORFactorGroup, "and": fgroup.ANDFactorGroup}[kind]
  fg.add_factors(group(variables_for_factors=[
      [parents[k] for k in range(cum[f], cum[f + 1])] + [children[f]] for f in range(len(sizes))]))
  bp = infer.BP(fg.bp_state, temperature=temperature)
  ftov = {parents[k]: rng.normal(size=2) for k in range(int(npar.sum()))}
  ftov.update({children[f]: rng.normal(size=2) for f in range(len(sizes))})
  arrays = bp.init(evidence_updates={parents: rng.gumbel(size=(int(npar.sum()), 2)),
                                     children: rng.gumbel(size=(len(sizes), 2))},
                   ftov_msgs_updates=ftov)
  outs = {}
  for mode in (1, 2):
    _lib.set_option("mode", mode)
    outs[mode] = _np(bp.run(arrays, num_iters=6, temperature=temperature).ftov_msgs)
  ref, _ = _oracle(bp, arrays.log_potentials, arrays.ftov_msgs, arrays.evidence, 6, 0.5,
                   temperature)
  np.testing.assert_array_equal(outs[2], outs[1])
  nan = np.isnan(ref)
  np.testing.assert_array_equal(np.isnan(outs[2]), nan)
  np.testing.assert_allclose(outs[2][~nan], ref[~nan], atol=1e-5)
