"""Host API: variable groups, factor groups and FactorGraph (CPU only).

The error cases and messages follow the reference's own tests
(tests/vgroup/test_vgroup.py, tests/fgroup/test_fgroup.py and
tests/fgraph/test_fgraph.py of the reference); the rest checks this
package's array forms of ``variables_for_factors``, its variable index and
its duplicate detection against the per-factor forms.
"""

import dataclasses
import re

import numpy as np
import pytest

from pgmax_amd import factor
from pgmax_amd import fgraph
from pgmax_amd import fgroup
from pgmax_amd import infer
from pgmax_amd import vgroup
from pgmax_amd import workloads
from pgmax_amd.fgraph import fgraph as fgraph_mod


# ---- VarDict (reference tests/vgroup/test_vgroup.py:26-88) ----------------

def test_var_dict_errors_and_round_trips():
  with pytest.raises(ValueError, match=re.escape("Expected num_states shape (3,). Got (4,).")):
    vgroup.VarDict(variable_names=(0, 1, 2), num_states=np.full((4,), 2))
  with pytest.raises(ValueError, match=re.escape(
      "num_states should be an integer or a NumPy array of dtype int")):
    vgroup.VarDict(variable_names=(0, 1, 2), num_states=np.full((3,), 2, dtype=np.float32))
  vd = vgroup.VarDict(variable_names=(0, 1, 2), num_states=15)
  with pytest.raises(ValueError, match="data is referring to a non-existent variable 3"):
    vd.flatten({3: np.zeros(10)})
  with pytest.raises(ValueError, match=re.escape(
      "Variable 2 expects a data array of shape (15,) or (1,). Got (10,).")):
    vd.flatten({2: np.zeros(10)})
  with pytest.raises(ValueError, match="Can only unflatten 1D array. Got a 2D array."):
    vd.unflatten(np.zeros((10, 20)), True)
  with pytest.raises(ValueError, match=re.escape(
      "flat_data should be shape (num_variable_states(=45),). Got (100,)")):
    vd.unflatten(np.zeros(100), True)
  with pytest.raises(ValueError, match=re.escape(
      "flat_data should be shape (num_variables(=3),). Got (100,)")):
    vd.unflatten(np.zeros(100), False)
  out = vd.unflatten(np.zeros(3), False)
  assert set(out) == {0, 1, 2} and all(float(v) == 0.0 for v in out.values())
  # per-state round trip, and leading batch axes kept
  names = ("a", "b", "c")
  vd2 = vgroup.VarDict(variable_names=names, num_states=np.array([2, 3, 1]))
  data = {n: np.arange(k, dtype=float) + 10 * i for i, (n, k) in enumerate(zip(names, (2, 3, 1)))}
  flat = vd2.flatten(data)
  np.testing.assert_array_equal(flat, [0, 1, 10, 11, 12, 20])
  back = vd2.unflatten(flat, True)
  for n in names:
    np.testing.assert_array_equal(back[n], data[n])
  batched = vd2.unflatten(np.stack([flat, -flat]), True)
  np.testing.assert_array_equal(batched["b"], [[10, 11, 12], [-10, -11, -12]])
  # variables: group hash + position, state count per name
  assert vd2["b"] == (hash(vd2) + 1, 3)
  with pytest.raises(ValueError, match="Variable zz is not in VarDict"):
    _ = vd2["zz"]


# ---- factor groups (reference tests/fgroup/test_fgroup.py) ---------------

def test_single_factor_group_errors_and_order():
  with pytest.raises(ValueError, match="Cannot create a FactorGroup with no Factor."):
    fgroup.ORFactorGroup(variables_for_factors=[])
  a = vgroup.NDVarArray(num_states=2, shape=(10,))
  b = vgroup.NDVarArray(num_states=2, shape=(10,))
  v0, v1 = (a[0], b[0]), (a[1], b[1])
  g0 = fgroup.ORFactorGroup(variables_for_factors=[v0])
  with pytest.raises(ValueError, match="SingleFactorGroup should only contain one factor. Got 2"):
    fgroup.SingleFactorGroup(variables_for_factors=[v0, v1], single_factor=g0)
  g1 = fgroup.ORFactorGroup(variables_for_factors=[v1])
  assert g0 < g1 and not g1 < g0  # groups order by creation


def test_enum_factor_group_errors_flatten_unflatten():
  vg = vgroup.NDVarArray(shape=(2, 2), num_states=3)
  vff = [[vg[0, 0], vg[0, 1], vg[1, 1]], [vg[0, 1], vg[1, 0], vg[1, 1]]]
  with pytest.raises(ValueError, match=re.escape(
      "Expected log potentials shape: (1,) or (2, 1). Got (3, 2)")):
    fgroup.EnumFactorGroup(variables_for_factors=vff, factor_configs=np.zeros((1, 3), dtype=int),
                           log_potentials=np.zeros((3, 2)))
  with pytest.raises(ValueError, match=re.escape("Potentials should be floats")):
    fgroup.EnumFactorGroup(variables_for_factors=vff, factor_configs=np.zeros((1, 3), dtype=int),
                           log_potentials=np.zeros((2, 1), dtype=int))
  g = fgroup.EnumFactorGroup(variables_for_factors=vff, factor_configs=np.zeros((1, 3), dtype=int))
  name = [vg[0, 0], vg[1, 1]]
  with pytest.raises(ValueError, match=re.escape(
      f"The queried factor connected to the set of variables {frozenset(name)} is not"
      " present in the factor group.")):
    _ = g[name]
  assert g[[vg[0, 1], vg[1, 0], vg[1, 1]]] is g.factors[1]
  with pytest.raises(ValueError, match=re.escape(
      "data should be of shape (2, 1) or (2, 9) or (1,). Got (4, 5).")):
    g.flatten(np.zeros((4, 5)))
  np.testing.assert_array_equal(g.flatten(np.ones(1)), np.ones(2))
  np.testing.assert_array_equal(g.flatten(np.ones((2, 9))), np.ones(18))
  with pytest.raises(ValueError, match=re.escape("Can only unflatten 1D array. Got a 3D array.")):
    g.unflatten(np.ones((1, 2, 3)))
  with pytest.raises(ValueError, match=re.escape(
      "flat_data should be compatible with shape (2, 1) or (2, 9). Got (30,)")):
    g.unflatten(np.zeros(30))
  np.testing.assert_array_equal(g.unflatten(np.arange(2)), [[0], [1]])
  np.testing.assert_array_equal(g.unflatten(np.ones(18)), np.ones((2, 9)))


def test_pairwise_factor_group_errors_flatten_unflatten():
  vg = vgroup.NDVarArray(shape=(2, 2), num_states=3)
  with pytest.raises(ValueError, match=re.escape("log_potential_matrix should be either a 2D array")):
    fgroup.PairwiseFactorGroup([[vg[0, 0], vg[1, 1]]], np.zeros((1,), dtype=float))
  with pytest.raises(ValueError, match=re.escape("Potential matrix should be floats")):
    fgroup.PairwiseFactorGroup([[vg[0, 0], vg[1, 1]]], np.zeros((3, 3), dtype=int))
  with pytest.raises(ValueError, match=re.escape(
      "Expected log_potential_matrix for 1 factors. Got log_potential_matrix for 2 factors.")):
    fgroup.PairwiseFactorGroup([[vg[0, 0], vg[1, 1]]], np.zeros((2, 3, 3), dtype=float))
  with pytest.raises(ValueError, match=re.escape(
      "All pairwise factors should connect to exactly 2 variables. Got a factor connecting to 3"
      " variables")):
    fgroup.PairwiseFactorGroup([[vg[0, 0], vg[1, 1], vg[0, 1]]], np.zeros((3, 3), dtype=float))
  name = [vg[0, 0], vg[1, 1]]
  with pytest.raises(ValueError, match=re.escape(f"The specified pairwise factor {name}")):
    fgroup.PairwiseFactorGroup([name], np.zeros((4, 4), dtype=float))
  g = fgroup.PairwiseFactorGroup([[vg[0, 0], vg[1, 1]], [vg[1, 0], vg[0, 1]]])
  with pytest.raises(ValueError, match=re.escape(
      "data should be of shape (2, 3, 3) or (2, 6) or (3, 3). Got (4, 4).")):
    g.flatten(np.zeros((4, 4)))
  np.testing.assert_array_equal(g.flatten(np.zeros((3, 3))), np.zeros(18))
  np.testing.assert_array_equal(g.flatten(np.zeros((2, 6))), np.zeros(12))
  with pytest.raises(ValueError, match="Can only unflatten 1D array. Got a 2D array"):
    g.unflatten(np.zeros((10, 20)))
  np.testing.assert_array_equal(g.unflatten(np.zeros(18)), np.zeros((2, 3, 3)))
  np.testing.assert_array_equal(g.unflatten(np.zeros(12)), np.zeros((2, 6)))
  with pytest.raises(ValueError, match=re.escape(
      "flat_data should be compatible with shape (2, 3, 3) or (2, 6). Got (10,).")):
    g.unflatten(np.zeros(10))


# ---- FactorGraph (reference tests/fgraph/test_fgraph.py) -----------------

def test_factor_graph_duplicate_messages():
  vg = vgroup.VarDict(variable_names=(0,), num_states=15)
  fg = fgraph.FactorGraph(vg)
  wrong = factor.EnumFactor(variables=[vg[0], vg[0]],
                            factor_configs=np.array([[i, i] for i in range(15)]),
                            log_potentials=np.zeros(15))
  with pytest.raises(ValueError, match=re.escape(
      f"A Factor of type {factor.EnumFactor} involving variables"
      f" {[(hash(vg), 15), (hash(vg), 15)]} contains variables duplicates.")):
    fg.add_factors(wrong)
  fg.add_factors(factor.EnumFactor(variables=[vg[0]], factor_configs=np.arange(15)[:, None],
                                   log_potentials=np.zeros(15)))
  group = fgroup.EnumFactorGroup(variables_for_factors=[[vg[0]]],
                                 factor_configs=np.arange(15)[:, None],
                                 log_potentials=np.zeros(15))
  with pytest.raises(ValueError, match=re.escape(
      f"A Factor of type {factor.EnumFactor} involving variables"
      f" {frozenset([(hash(vg), 15)])} already exists.")):
    fg.add_factors(group)


def test_bp_state_pieces_errors():
  vg = vgroup.VarDict(variable_names=(0,), num_states=15)
  fg0 = fgraph.FactorGraph(vg)
  f = factor.EnumFactor(variables=[vg[0]], factor_configs=np.arange(15)[:, None],
                        log_potentials=np.zeros(15))
  fg0.add_factors(f)
  fg1 = fgraph.FactorGraph(vg)
  fg1.add_factors(f)
  with pytest.raises(ValueError, match=(
      "log_potentials, ftov_msgs and evidence should be derived from the same fg_state")):
    infer.BPState(log_potentials=fg0.bp_state.log_potentials,
                  ftov_msgs=fg1.bp_state.ftov_msgs, evidence=fg1.bp_state.evidence)

  fg = fgraph.FactorGraph(vg)
  group = fgroup.EnumFactorGroup(variables_for_factors=[[vg[0]]],
                                 factor_configs=np.arange(10)[:, None])
  fg.add_factors(group)
  with pytest.raises(ValueError, match=re.escape(
      "Expected log potentials shape (10,) for factor group.")):
    fg.bp_state.log_potentials[group] = np.zeros((1, 15))
  with pytest.raises(ValueError, match=re.escape("Invalid FactorGroup queried to access log potentials.")):
    _ = fg.bp_state.log_potentials[vg[0]]
  with pytest.raises(ValueError, match=re.escape("Expected log potentials shape (10,). Got (15,)")):
    infer.LogPotentials(fg_state=fg.fg_state, value=np.zeros(15))
  with pytest.raises(ValueError, match=re.escape(
      "Provided variable or factor type is not in the FactorGraph")):
    fg.bp_state.ftov_msgs[0] = np.ones(10)
  with pytest.raises(ValueError, match=re.escape(
      f"Expected ftov_msgs shape (15,) for variable ({hash(vg)}, 15). Got incompatible shape (10,).")):
    fg.bp_state.ftov_msgs[vg[0]] = np.ones(10)
  with pytest.raises(ValueError, match=re.escape("Expected messages shape (15,). Got (10,)")):
    infer.FToVMessages(fg_state=fg.fg_state, value=np.zeros(10))

  vg2 = vgroup.VarDict(variable_names=(0, 1), num_states=15)
  fg2 = fgraph.FactorGraph(vg2)
  fg2.add_factors(fgroup.EnumFactorGroup(variables_for_factors=[[vg2[0], vg2[1]]],
                                         factor_configs=np.arange(10)[:, None] + np.zeros((1, 2))))
  ev = infer.Evidence(fg_state=fg2.fg_state, value=np.zeros(30))
  with pytest.raises(ValueError, match=re.escape("Expected evidence shape (30,). Got (20,).")):
    infer.Evidence(fg_state=fg2.fg_state, value=np.zeros(20))
  with pytest.raises(ValueError, match=re.escape(
      f"Expected evidence shape (15,) for variable {vg2[0]}. Got incompatible shape (10,).")):
    infer.bp_state.update_evidence(ev.value, {vg2[0]: np.zeros(10)}, fg2.fg_state)
  other = vgroup.VarDict(variable_names=(0,), num_states=15)
  with pytest.raises(ValueError, match=re.escape(
      "Got evidence for a variable or a VarGroup not in the FactorGraph!")):
    infer.bp_state.update_evidence(ev.value, {other[0]: np.zeros(15)}, fg2.fg_state)


# ---- array forms, variable index, duplicate detection ---------------------

def _flat_state(fg):
  st = fg.fg_state
  return st, {t: st.wiring[t].get_inference_arguments() for t in st.wiring}


def _same_state(fg_a, fg_b):
  sa, ia = _flat_state(fg_a)
  sb, ib = _flat_state(fg_b)
  assert sa.num_var_states == sb.num_var_states
  assert sa.total_factor_num_states == sb.total_factor_num_states
  assert dict(sa.factor_type_to_msgs_range) == dict(sb.factor_type_to_msgs_range)
  np.testing.assert_array_equal(sa.log_potentials, sb.log_potentials)
  np.testing.assert_array_equal(sa.evidence_to_vars, sb.evidence_to_vars)
  for t in sa.wiring:
    np.testing.assert_array_equal(sa.wiring[t].var_states_for_edges, sb.wiring[t].var_states_for_edges)
    for k, v in ia[t].items():
      np.testing.assert_array_equal(np.asarray(v), np.asarray(ib[t][k]), err_msg=f"{t} {k}")


def test_array_and_ragged_forms_compile_like_lists():
  rng = np.random.default_rng(3)
  vg = vgroup.NDVarArray(num_states=2, shape=(40,))
  ns = vgroup.NDVarArray(num_states=np.array([2, 3, 4, 3, 2, 5]), shape=(6,))
  pairs = [rng.choice(40, 3, replace=False) for _ in range(25)]
  ragged = [rng.choice(40, int(k), replace=False) for k in rng.integers(2, 7, size=30)]
  enum_vars = [[0, 1], [2, 3], [4, 5], [1, 2]]
  cfg = np.array([[a, b] for a in range(2) for b in range(2)])

  def build(form):
    fg = fgraph.FactorGraph([vg, ns])
    if form == "list":
      fg.add_factors(fgroup.ANDFactorGroup([[vg[int(i)] for i in p] for p in pairs]))
      fg.add_factors(fgroup.ORFactorGroup([[vg[int(i)] for i in r] for r in ragged]))
      fg.add_factors(fgroup.EnumFactorGroup([[ns[i] for i in e] for e in enum_vars], cfg,
                                            log_potentials=np.arange(16.0).reshape(4, 4)))
    else:
      arr = np.array([[(hash(vg) + int(i), 2) for i in p] for p in pairs], np.int64)
      fg.add_factors(fgroup.ANDFactorGroup(arr))
      flat = np.array([(hash(vg) + int(i), 2) for r in ragged for i in r], np.int64)
      fg.add_factors(fgroup.ORFactorGroup(fgroup.RaggedVariables(
          pairs=flat, sizes=np.array([len(r) for r in ragged]))))
      earr = np.array([[(hash(ns) + i, int(ns.num_states[i])) for i in e] for e in enum_vars])
      fg.add_factors(fgroup.EnumFactorGroup(earr, cfg, log_potentials=np.arange(16.0).reshape(4, 4)))
    return fg

  a, b = build("list"), build("array")
  _same_state(a, b)
  # factor objects come out the same either way
  fa = a.factors[factor.ORFactor]
  fb = b.factors[factor.ORFactor]
  assert [list(map(tuple, f.variables)) for f in fa] == [[tuple(map(int, v)) for v in f.variables]
                                                         for f in fb]


def test_var_index_mapping_and_bulk_lookup():
  g1 = vgroup.NDVarArray(num_states=np.array([[2, 3], [4, 1]]), shape=(2, 2))
  g2 = vgroup.VarDict(variable_names=("x", "y"), num_states=5)
  idx = fgraph_mod.VarIndex([g1, g2])
  expected = {}
  start = 0
  for var in g1.variables + g2.variables:
    expected[(int(var[0]), int(var[1]))] = start
    start += int(var[1])
  assert dict(idx) == expected and len(idx) == 6 and idx.num_var_states == start
  assert g2["y"] in idx and idx[g2["y"]] == expected[g2["y"]]
  assert (hash(g2) + 1, 4) not in idx  # right hash, wrong state count
  assert 0 not in idx and g1 not in idx and (1, 2, 3) not in idx
  hashes = np.array([v[0] for v in expected])
  states = np.array([v[1] for v in expected])
  np.testing.assert_array_equal(idx.lookup(hashes[::-1], states[::-1]),
                                np.array(list(expected.values()))[::-1])
  with pytest.raises(KeyError):
    idx.lookup(np.array([hash(g2) + 7]), np.array([5]))


@dataclasses.dataclass(frozen=True, eq=False)
class _FixedHashes(vgroup.VarGroup):
  hashes: tuple = ()

  @property
  def variable_hashes(self):
    return np.array(self.hashes, dtype=np.int64)


def test_var_index_scattered_hashes_and_shared_hash():
  g = _FixedHashes(num_states=np.array([2, 2, 3, 2]), hashes=(50, 7, 8, 100))
  idx = fgraph_mod.VarIndex([g])
  np.testing.assert_array_equal(idx.lookup(np.array([100, 8, 7, 50]), np.array([2, 3, 2, 2])),
                                [7, 4, 2, 0])
  clash = _FixedHashes(num_states=np.array([2]), hashes=(8,))
  with pytest.raises(ValueError, match="share a hash"):
    fgraph_mod.VarIndex([g, clash])


def test_duplicates_found_in_bulk_in_factor_order():
  vg = vgroup.NDVarArray(num_states=2, shape=(30,))
  h = hash(vg)
  fg = fgraph.FactorGraph(vg)
  fg.add_factors(fgroup.ANDFactorGroup(np.array([[(h + i, 2), (h + i + 1, 2), (h + i + 2, 2)]
                                                 for i in range(20)])))
  # same set in another order, inside an array-form group: reported at factor 3
  # (expected sets are built in the factor's own order: a frozenset prints in
  # an order that can depend on insertion order)
  arr = np.array([[(h + 25, 2), (h + 26, 2), (h + 27, 2)]] * 3 + [[(h + 7, 2), (h + 5, 2), (h + 6, 2)]])
  arr[1, 0, 0] = h + 28
  arr[2, 0, 0] = h + 29
  with pytest.raises(ValueError, match=re.escape(
      f"involving variables {frozenset([(h + 7, 2), (h + 5, 2), (h + 6, 2)])} already exists")):
    fg.add_factors(fgroup.ANDFactorGroup(arr))
  # a repeat inside the group: the second occurrence is reported; a repeated
  # variable in an earlier factor wins over it
  rep = np.array([[(h + 0, 2), (h + 29, 2)], [(h + 1, 2), (h + 28, 2)], [(h + 29, 2), (h + 0, 2)]])
  with pytest.raises(ValueError, match=re.escape(
      f"involving variables {frozenset([(h + 29, 2), (h, 2)])} already exists")):
    fg.add_factors(fgroup.ORFactorGroup(rep))
  rep[1, 1, 0] = h + 1
  with pytest.raises(ValueError, match=re.escape(
      f"involving variables {[(h + 1, 2), (h + 1, 2)]} contains variables duplicates")):
    fg.add_factors(fgroup.ORFactorGroup(rep))
  # a refused group leaves nothing behind; the same sets are fine for OR
  fg.add_factors(fgroup.ORFactorGroup(arr))
  assert len(fg.factor_groups[factor.ANDFactor]) == 1
  assert len(fg.factor_groups[factor.ORFactor]) == 1


def test_duplicates_confirmed_exactly_under_fingerprint_collisions(monkeypatch):
  """With every variable key forced equal, all fingerprints of one arity
  collide: distinct sets must still pass and a real repeat still raise."""
  monkeypatch.setattr(fgraph_mod, "_variable_keys",
                      lambda hashes, num_states: np.zeros(hashes.shape, np.uint64))
  vg = vgroup.NDVarArray(num_states=2, shape=(12,))
  fg = fgraph.FactorGraph(vg)
  fg.add_factors(fgroup.ORFactorGroup([[vg[i], vg[i + 1]] for i in range(10)]))
  fg.add_factors(fgroup.ORFactorGroup([[vg[0], vg[i]] for i in range(2, 12)]))
  with pytest.raises(ValueError, match="already exists"):
    fg.add_factors(fgroup.ORFactorGroup([[vg[0], vg[11]], [vg[5], vg[4]]]))
  with pytest.raises(ValueError, match="contains variables duplicates"):
    fg.add_factors(fgroup.ORFactorGroup([[vg[3], vg[9]], [vg[3], vg[3]]]))


def test_single_factor_adds_and_cache_versions():
  vg = vgroup.NDVarArray(num_states=3, shape=(50,))
  fg = fgraph.FactorGraph(vg)
  for i in range(49):
    fg.add_factors(factor.EnumFactor(variables=[vg[i], vg[i + 1]],
                                     factor_configs=np.array([[0, 0], [1, 2]]),
                                     log_potentials=np.array([0.5, float(i)])))
  s1 = fg.fg_state
  assert fg.fg_state is s1 and s1.total_factor_num_states == 49 * 6
  with pytest.raises(ValueError, match="already exists"):
    fg.add_factors(factor.EnumFactor(variables=[vg[1], vg[0]],
                                     factor_configs=np.array([[0, 0]]),
                                     log_potentials=np.zeros(1)))
  assert fg.fg_state is s1  # a refused add changes nothing
  fg.add_factors(fgroup.PairwiseFactorGroup([[vg[0], vg[2]]]))
  s2 = fg.fg_state
  assert s2 is not s1 and s2.total_factor_num_states == 49 * 6 + 6
  np.testing.assert_array_equal(s2.log_potentials[:98:2], 0.5)
  np.testing.assert_array_equal(s2.log_potentials[1:98:2], np.arange(49.0))


def test_deconvolution_factor_graph_equals_direct_builder():
  """Array-form AND / ragged OR groups through FactorGraph compile to the
  same flat wiring as the direct builder, and as the per-factor build."""
  import sys, os  # pylint: disable=g-import-not-at-top,multiple-imports
  sys.path.insert(0, os.path.dirname(__file__))
  import test_batch  # pylint: disable=g-import-not-at-top
  for case in ((2, 1, 6, 5, 2, 3, 2), (3, 2, 7, 6, 2, 2, 3)):
    flat, _ = workloads.deconvolution_wiring(*case)
    fg, _ = workloads.deconvolution_factor_graph(*case)
    test_batch._assert_same(flat, test_batch._flat(fg))
