"""Batched uneven graphs: layout checks on the CPU, parity and per-graph convergence on the GPU."""

import ctypes
import dataclasses
import itertools

import numpy as np
import pytest

import graphs
from pgmax_amd import _lib
from pgmax_amd import batch
from pgmax_amd import factor
from pgmax_amd import fgraph
from pgmax_amd import fgroup
from pgmax_amd import vgroup
from pgmax_amd import infer
from pgmax_amd.infer.inferer import InfererContext


def _flat(fg):
  return batch.FlatWiring.from_context(InfererContext(fg.bp_state))


def _assert_same(a: batch.FlatWiring, b: batch.FlatWiring):
  for name in ("var_states_for_edge_states", "edge_indices_for_edge_states",
               "factor_indices_for_edge_states", "evidence_to_vars"):
    np.testing.assert_array_equal(getattr(a, name), getattr(b, name), err_msg=name)
  for x, y in zip(a.enum_rows(), b.enum_rows()):
    np.testing.assert_array_equal(x, y)
  assert a.msgs_range == b.msgs_range and a.potentials_range == b.potentials_range
  assert a.num_var_states == b.num_var_states
  for t in (batch.OR, batch.AND, batch.POOL):
    for x, y in zip(a.logical[t][:3], b.logical[t][:3]):
      np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("side,states,torus", [(5, 3, False), (4, 2, True), (6, 7, False)])
def test_grid_wiring_matches_factor_graph(side, states, torus):
  if torus:
    fg, _, _ = graphs.ising_torus(side=side)
  else:
    fg, _, _ = graphs.potts_union([(side, states, 0.5)])
  _assert_same(batch.grid_wiring(side, side, states if not torus else 2, torus=torus), _flat(fg))


def test_union_of_grids_matches_factor_graph():
  specs = [(5, 3, 0.5), (4, 2, 1.0), (6, 4, 0.2)]
  fg, _, _ = graphs.potts_union(specs)
  union, layout = batch.concatenate([batch.grid_wiring(s, s, k) for s, k, _ in specs])
  _assert_same(union, _flat(fg))
  assert layout.num_es == union.num_edge_states


def _info(flat: batch.FlatWiring):
  w = flat.to_engine_wiring()
  info = (ctypes.c_int64 * 8)()
  _lib.check(_lib.load().pgx_wiring_check(ctypes.byref(w.struct), info), "pgx_wiring_check")
  return [int(v) for v in info]


@pytest.mark.parametrize("states", [2, 3, 7])
def test_full_tables_match_explicit_rows(states):
  """Implicit full Enum tables list the same rows and compile to the same plan."""
  full = batch.grid_wiring(5, 6, states, full_tables=True)
  explicit = batch.grid_wiring(5, 6, states)
  assert full.full_tables and not explicit.full_tables
  _assert_same(full, explicit)
  assert _info(full) == _info(explicit)


def test_full_tables_of_higher_order_factors():
  """A 3-variable full table compiles like its explicit row-major rows."""
  fg = fgraph.FactorGraph([vgroup.NDVarArray(num_states=3, shape=(4,)),
                           vgroup.NDVarArray(num_states=2, shape=(2,))])
  v3, v2 = fg.variable_groups
  configs = np.array(list(itertools.product(range(3), range(2), range(3))))
  fg.add_factors(factor.EnumFactor(variables=[v3[0], v2[1], v3[2]], factor_configs=configs,
                                   log_potentials=np.zeros(len(configs))))
  fg.add_factors(factor.EnumFactor(variables=[v3[1], v3[3]],
                                   factor_configs=np.array(list(itertools.product(range(3), range(3)))),
                                   log_potentials=np.zeros(9)))
  explicit = _flat(fg)
  full = dataclasses.replace(explicit, enum_ci=None, enum_ce=None)
  _assert_same(full, explicit)
  assert _info(full) == _info(explicit)


def test_full_tables_reject_wrong_potential_count():
  full = batch.grid_wiring(3, 3, 2, full_tables=True)
  bad = dataclasses.replace(full, num_log_potentials=full.num_log_potentials - 1,
                            potentials_range=[(0, full.num_log_potentials - 1)] +
                            [(full.num_log_potentials - 1,) * 2] * 3)
  with pytest.raises(ValueError, match="full-table"):
    _info(bad)


def test_union_keeps_full_tables_and_mixes_with_rows():
  a = batch.grid_wiring(4, 5, 3, full_tables=True)
  b = batch.grid_wiring(3, 3, 2, full_tables=True)
  c = batch.grid_wiring(3, 4, 4)
  u_full, _ = batch.concatenate([a, b])
  assert u_full.full_tables
  u_ref, _ = batch.concatenate([batch.grid_wiring(4, 5, 3), batch.grid_wiring(3, 3, 2)])
  _assert_same(u_full, u_ref)
  u_mixed, _ = batch.concatenate([a, c])
  assert not u_mixed.full_tables
  u_ref2, _ = batch.concatenate([batch.grid_wiring(4, 5, 3), c])
  _assert_same(u_mixed, u_ref2)


def test_union_of_mixed_factor_types_matches_factor_graph():
  """Enum + OR graphs and Enum + Pool graphs: the union is type-major."""
  pair_or, _, _ = graphs.logical_pair("or", 1)
  pair_pool, _, _ = graphs.logical_pair("pool", 2)
  g1, g2 = pair_or[0], pair_pool[1]
  union, _ = batch.concatenate([_flat(g1[0]), _flat(g2[0])])
  fg = fgraph.FactorGraph([g1[1], g1[2], g2[1], g2[2]])
  for src in (g1[0], g2[0]):
    for groups in src.factor_groups.values():
      for grp in groups:
        fg.add_factors(grp)
  _assert_same(union, _flat(fg))


# ---------------------------------------------------------------------------
# GPU
# ---------------------------------------------------------------------------


def _specs(n, seed):
  rng = np.random.default_rng(seed)
  return [(int(rng.integers(4, 24)), int(rng.integers(2, 9)), float(rng.uniform(0.2, 1.5)))
          for _ in range(n)]


@pytest.mark.gpu
@pytest.mark.parametrize("temperature", [0.0, 1.0])
def test_batched_equals_single_graph_runs(gpu, temperature):
  specs = _specs(12, 4)
  rng = np.random.default_rng(5)
  wirings = [batch.grid_wiring(s, s, k) for s, k, _ in specs]
  lps = [np.tile((b * np.eye(k)).reshape(-1), w.num_log_potentials // (k * k))
         for (s, k, b), w in zip(specs, wirings)]
  evs = [rng.normal(size=w.num_var_states) for w in wirings]
  bb = batch.BatchedBP(wirings, temperature=temperature)
  out, _ = bb.run(bb.pack_log_potentials(lps), None, bb.pack_evidence(evs), 40)
  per_graph = bb.unpack_ftov(out)
  for g, (s, k, b) in enumerate(specs):
    fg, groups, _ = graphs.potts_union([(s, k, b)])
    bp = infer.BP(fg.bp_state, temperature=temperature)
    arrays = infer.BPArrays(lps[g], bp.init().ftov_msgs, evs[g])
    ref = bp.run(arrays, 40).ftov_msgs.cpu().numpy()
    np.testing.assert_array_equal(per_graph[g].cpu().numpy(), ref)


@pytest.mark.gpu
def test_full_tables_run_like_explicit_rows(gpu):
  specs = _specs(6, 8)
  rng = np.random.default_rng(9)
  full = [batch.grid_wiring(s, s, k, full_tables=True) for s, k, _ in specs]
  expl = [batch.grid_wiring(s, s, k) for s, k, _ in specs]
  lps = [rng.normal(size=w.num_log_potentials) for w in full]
  evs = [rng.normal(size=w.num_var_states) for w in full]
  outs = []
  for ws in (full, expl):
    bb = batch.BatchedBP(ws, temperature=1.0)
    out, deltas = bb.run(bb.pack_log_potentials(lps), None, bb.pack_evidence(evs), 30)
    outs.append((out.cpu().numpy(), deltas.cpu().numpy()))
  np.testing.assert_array_equal(outs[0][0], outs[1][0])
  np.testing.assert_array_equal(outs[0][1], outs[1][1])


@pytest.mark.gpu
def test_run_converge_per_graph(gpu):
  """Each graph stops at its first delta <= tol; messages equal a fixed-count run."""
  specs = _specs(10, 6)
  rng = np.random.default_rng(7)
  wirings = [batch.grid_wiring(s, s, k) for s, k, _ in specs]
  lps = [np.tile((b * np.eye(k)).reshape(-1), w.num_log_potentials // (k * k))
         for (s, k, b), w in zip(specs, wirings)]
  evs = [rng.normal(size=w.num_var_states) for w in wirings]
  bb = batch.BatchedBP(wirings, temperature=1.0)
  lp, ev = bb.pack_log_potentials(lps), bb.pack_evidence(evs)
  out, deltas, iters = bb.run_converge(lp, None, ev, max_iters=300, tol=1e-4)
  iters = iters.cpu().numpy()
  deltas = deltas.cpu().numpy()
  per_graph = bb.unpack_ftov(out)
  for g, (s, k, b) in enumerate(specs):
    fg, _, _ = graphs.potts_union([(s, k, b)])
    bp = infer.BP(fg.bp_state, temperature=1.0)
    arrays = infer.BPArrays(lps[g], bp.init().ftov_msgs, evs[g])
    ref_out, ref_d = bp.run_with_diffs(arrays, 300)
    ref_d = ref_d.cpu().numpy()
    below = np.nonzero(ref_d <= 1e-4)[0]
    want = int(below[0]) + 1 if below.size else 300
    assert iters[g] == want
    np.testing.assert_array_equal(deltas[g, :want], ref_d[:want])
    assert np.all(deltas[g, want:] == 0)
    fixed = bp.run(arrays, want).ftov_msgs.cpu().numpy()
    np.testing.assert_array_equal(per_graph[g].cpu().numpy(), fixed)


@pytest.mark.gpu
def test_autotune_keeps_results(gpu):
  """The tuned shape is one of the valid shapes and leaves results bit-identical."""
  specs = _specs(8, 21)
  rng = np.random.default_rng(22)
  ws = [batch.grid_wiring(s, s, k, full_tables=True) for s, k, _ in specs]
  lps = [rng.normal(size=w.num_log_potentials) for w in ws]
  evs = [rng.normal(size=w.num_var_states) for w in ws]
  bb = batch.BatchedBP(ws, temperature=1.0)
  lp, ev = bb.pack_log_potentials(lps), bb.pack_evidence(evs)
  before, d0 = bb.run(lp, None, ev, 25)
  cfg = bb.autotune(lp, None, ev, probe_iters=5)
  assert cfg["mode"] in (1, 2) and cfg["vec"] in (0, 3)
  after, d1 = bb.run(lp, None, ev, 25)
  assert np.array_equal(before.cpu().numpy(), after.cpu().numpy())
  assert np.array_equal(d0.cpu().numpy(), d1.cpu().numpy())
  cfg_c = bb.autotune(lp, None, ev, converge=True, probe_iters=5)
  assert cfg_c["mode"] == 1


def test_deconvolution_wiring_matches_factor_graph():
  """The direct builder equals FactorGraph + AND/OR factor groups built as in
  the reference's deconvolution notebook (cells 12-14), on a small instance."""
  import collections
  n_img, n_chan, H, Wd, n_feat, fh, fw = 2, 1, 6, 5, 2, 3, 2
  flat, _ = graphs.deconvolution_wiring(n_img, n_chan, H, Wd, n_feat, fh, fw)
  sh, sw = H - fh + 1, Wd - fw + 1
  W = vgroup.NDVarArray(num_states=2, shape=(n_chan, n_feat, fh, fw))
  S = vgroup.NDVarArray(num_states=2, shape=(n_img, n_feat, sh, sw))
  SW = vgroup.NDVarArray(num_states=2, shape=(n_img, n_chan, H, Wd, n_feat, fh, fw))
  X = vgroup.NDVarArray(num_states=2, shape=(n_img, n_chan, H, Wd))
  fg = fgraph.FactorGraph(variable_groups=[S, W, SW, X])
  and_vars, or_dict = [], collections.defaultdict(list)
  for i in range(n_img):
    for c in range(n_chan):
      for a in range(sh):
        for b in range(sw):
          for f in range(n_feat):
            for p in range(fh):
              for q in range(fw):
                swv = SW[i, c, a + p, b + q, f, p, q]
                and_vars.append([S[i, f, a, b], W[c, f, p, q], swv])
                or_dict[X[i, c, a + p, b + q]].append(swv)
  fg.add_factors(fgroup.ANDFactorGroup(and_vars))
  fg.add_factors(fgroup.ORFactorGroup([list(v) + [k] for k, v in or_dict.items()]))
  _assert_same(flat, _flat(fg))
  assert _info(flat)[5] == len(and_vars) + len(or_dict)


def _deconv_oracle_graph(flat):
  return {"var_states_for_edge_states": flat.var_states_for_edge_states,
       "edge_indices_for_edge_states": flat.edge_indices_for_edge_states,
       "factor_indices_for_edge_states": flat.factor_indices_for_edge_states,
       "num_edges": int(flat.edge_indices_for_edge_states[-1]) + 1,
       "msgs_range": dict(zip(("enum", "or", "and", "pool"), flat.msgs_range)),
       "potentials_range": dict(zip(("enum", "or", "and", "pool"), flat.potentials_range)),
       "args": {"enum": {"factor_configs_indices": flat.enum_ci,
                         "factor_configs_edge_states": flat.enum_ce},
                "or": dict(zip(("parents_factor_indices", "parents_msg_indices",
                                "children_edge_states", "edge_states_offset"), flat.logical[batch.OR])),
                "and": dict(zip(("parents_factor_indices", "parents_msg_indices",
                                 "children_edge_states", "edge_states_offset"), flat.logical[batch.AND])),
                "pool": {"pool_choices_factor_indices": np.zeros(0, np.int64),
                         "pool_choices_msg_indices": np.zeros(0, np.int64),
                         "pool_indicators_edge_states": np.zeros(0, np.int64)}}}


# (images, channels, height, width, features, feature height, feature width):
# the second has OR factors of up to 64 parents (lane groups) and W variables
# of 1,156 edges (group sums of chunk sums)
DECONV_CASES = [(2, 1, 10, 10, 3, 3, 3), (4, 1, 20, 20, 4, 4, 4)]


@pytest.mark.gpu
@pytest.mark.parametrize("case", DECONV_CASES)
@pytest.mark.parametrize("temperature", [0.0, 0.3, 1.0])
@pytest.mark.parametrize("mode", [1, 2])
def test_deconvolution_matches_oracle(gpu, case, temperature, mode):
  """OR/AND factors with the deconvolution structure, engine vs the FP32 oracle."""
  from oracle import lbp
  from pgmax_amd import _lib, engine
  flat, meta = graphs.deconvolution_wiring(*case)
  n_img, _, h, w = case[:4]
  X = graphs.deconvolution_problem(n_img, h, w, n_feat_true=2, feat_true=3, seed=1)
  ev = graphs.deconvolution_evidence(meta, X, x_logit=-8.0)
  rng = np.random.default_rng(2)
  ev = (ev + rng.gumbel(size=ev.shape) * 0.5).astype(np.float32)
  eng = engine.Engine.from_wiring(flat.to_engine_wiring())
  ft0 = rng.normal(size=flat.num_edge_states).astype(np.float32)
  prev = _lib.set_option("mode", mode)
  try:
    out, d = eng.run(np.zeros(0, np.float32), ft0, ev, 25, 0.5, temperature)
  finally:
    _lib.set_option("mode", prev)
  ref, ref_d = lbp.run_with_diffs(_deconv_oracle_graph(flat), np.zeros(0, np.float32), ft0, ev,
                                  25, 0.5, temperature, np.float32)
  got = out.cpu().numpy()
  # the reference's T > 0 logical updates can produce inf - inf = NaN
  # (logminusexp of equal sums); the NaN positions must agree
  nan = np.isnan(ref)
  np.testing.assert_array_equal(np.isnan(got), nan)
  err = float(np.abs(got[~nan] - ref[~nan]).max(initial=0.0))
  print(f"deconvolution {case} T={temperature} mode {mode}: max |engine - oracle| = {err:.3e} "
        f"({int(nan.sum())} NaN in both), bitwise equal {np.array_equal(got, ref, equal_nan=True)}")
  assert err <= 1e-5
  np.testing.assert_allclose(d.cpu().numpy(), ref_d, atol=1e-5)


def test_many_variable_states_with_heavy_variables():
  """Descriptors hold per-variable sizes, not global offsets: a graph with
  more than 2^23 variable states and a 40-edge variable compiles."""
  nv = (1 << 22) + 10  # 2 states each: more than 2^23 variable states
  pairs = np.stack([np.zeros(40, np.int64), np.arange(1, 41)], 1)
  w = batch.pairwise_wiring(nv, 2, pairs, full_tables=True)
  info = _info(w)
  assert info[1] == 2 * nv and info[4] == nv and info[5] == 40
