"""Host-side checks that need no GPU: the C ABI library, wiring validation and
the layout produced by the host contract (pgmax_amd.vgroup/fgroup/fgraph).
"""

import ctypes
import os
import re

import numpy as np
import pytest

import graphs
from pgmax_amd import _lib
from pgmax_amd import engine
from pgmax_amd import factor
from pgmax_amd import fgraph
from pgmax_amd import fgroup
from pgmax_amd import vgroup
from pgmax_amd.infer.inferer import InfererContext

HEADER = os.path.join(os.path.dirname(__file__), "..", "include", "pgx.h")


def _declared_functions():
  text = open(HEADER).read()
  return sorted(set(re.findall(r"^\s*(?:int|int64_t|const char\*)\s+(pgx_\w+)\(", text, re.M)))


def test_library_exports_every_declared_symbol():
  lib = _lib.load()
  names = _declared_functions()
  assert len(names) >= 10
  for name in names:
    assert hasattr(lib, name), name
    assert name in _lib.SYMBOLS, f"{name} has no ctypes prototype"


def _check(ctx):
  w = engine.wiring_from_context(ctx)
  info = (ctypes.c_int64 * 8)()
  rc = _lib.load().pgx_wiring_check(ctypes.byref(w.struct), info)
  _lib.check(rc, "pgx_wiring_check")
  return [int(v) for v in info]


def test_wiring_check_golden_graph():
  _, bp_state, _, _ = graphs.e2e_sanity_graph()
  ctx = InfererContext(bp_state)
  es, vs, c, e, nv, nf, ntask, nblk = _check(ctx)
  assert (es, vs, nv, nf) == (84, 36, 12, 10)
  assert e == 28 and ntask == 1 and nblk >= 2


def test_wiring_check_finds_components():
  """Three disjoint Potts grids compile to three independent tasks."""
  rng = np.random.default_rng(0)
  groups, fg_groups = [], []
  for side, s in ((5, 3), (4, 2), (6, 4)):
    v = vgroup.NDVarArray(num_states=s, shape=(side, side))
    groups.append(v)
    vff = [[v[i, j], v[i + 1, j]] for i in range(side - 1) for j in range(side)]
    vff += [[v[i, j], v[i, j + 1]] for i in range(side) for j in range(side - 1)]
    fg_groups.append(fgroup.PairwiseFactorGroup(
        variables_for_factors=vff, log_potential_matrix=rng.normal(size=(s, s))))
  fg = fgraph.FactorGraph(groups)
  fg.add_factors(fg_groups)
  info = _check(InfererContext(fg.bp_state))
  assert info[6] == 1  # tiny components are merged into one task


def test_wiring_check_rejects_bad_wiring():
  fg, _, _ = graphs.ising_torus(side=4)
  ctx = InfererContext(fg.bp_state)
  w = engine.wiring_from_context(ctx)
  bad = np.array(w.keep[0], copy=True)
  bad[3] = bad[2]  # an edge whose states repeat a variable state
  w.struct.var_states_for_edge_states = bad.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
  rc = _lib.load().pgx_wiring_check(ctypes.byref(w.struct), None)
  with pytest.raises(ValueError, match="edge"):
    _lib.check(rc, "pgx_wiring_check")


def test_pairwise_configs_row_major():
  """pgmax/fgroup/enum.py:303-310: config a*S1+b is (a, b)."""
  v = vgroup.NDVarArray(num_states=np.array([3, 2]), shape=(2,))
  g = fgroup.PairwiseFactorGroup(variables_for_factors=[[v[0], v[1]]])
  np.testing.assert_array_equal(
      g.factor_configs, [[0, 0], [0, 1], [1, 0], [1, 1], [2, 0], [2, 1]])


def test_one_group_equals_single_factors():
  """Reference tests/fgroup/test_wiring.py: a group and its single factors wire identically."""
  v = vgroup.NDVarArray(num_states=3, shape=(4,))
  configs = np.array([[0, 0, 1], [1, 2, 0], [2, 2, 2]])
  vff = [[v[0], v[1], v[2]], [v[1], v[2], v[3]]]
  fg1 = fgraph.FactorGraph(v)
  fg1.add_factors(fgroup.EnumFactorGroup(variables_for_factors=vff, factor_configs=configs))
  fg2 = fgraph.FactorGraph(v)
  for vs in vff:
    fg2.add_factors(factor.EnumFactor(variables=vs, factor_configs=configs,
                                      log_potentials=np.zeros(3)))
  w1 = fg1.wiring[factor.EnumFactor]
  w2 = fg2.wiring[factor.EnumFactor]
  np.testing.assert_array_equal(w1.var_states_for_edges, w2.var_states_for_edges)
  np.testing.assert_array_equal(w1.factor_configs_edge_states, w2.factor_configs_edge_states)
  # spot-check the layout against the definition: row (f*C + c)*V + j
  f, c, j = 1, 2, 1
  row = (f * 3 + c) * 3 + j
  assert w1.factor_configs_edge_states[row, 0] == f * 3 + c
  assert w1.factor_configs_edge_states[row, 1] == (f * 3 + j) * 3 + configs[c, j]


def test_logical_wiring_positions():
  """pgmax/factor/logical.py:457-488: parents then child; relevant state 0 (OR) / 1 (AND)."""
  v = vgroup.NDVarArray(num_states=2, shape=(5,))
  for grp, rel in ((fgroup.ORFactorGroup, 0), (fgroup.ANDFactorGroup, 1)):
    fg = fgraph.FactorGraph(v)
    fg.add_factors(grp(variables_for_factors=[[v[0], v[1], v[2]], [v[3], v[4]]]))
    w = fg.wiring[factor.ORFactor if rel == 0 else factor.ANDFactor]
    np.testing.assert_array_equal(w.parents_edge_states,
                                  [[0, 0 + rel], [0, 2 + rel], [1, 6 + rel]])
    np.testing.assert_array_equal(w.children_edge_states, [4 + rel, 8 + rel])


def test_ftov_update_spreads_over_edges():
  """pgmax/infer/bp_state.py:208-228: a variable's message is split over its edges."""
  fg, variables, _ = graphs.ising_torus(side=3)
  bp_state = fg.bp_state
  bp_state.ftov_msgs[variables[1, 1]] = np.array([1.0, 3.0])
  ctx = InfererContext(bp_state)
  vs0 = bp_state.fg_state.vars_to_starts[variables[1, 1]]
  idx = np.nonzero(ctx.var_states_for_edge_states == vs0)[0]
  assert idx.shape[0] == 4
  for i in idx:
    np.testing.assert_allclose(bp_state.ftov_msgs.value[i:i + 2], [0.25, 0.75])


def test_evidence_vardict_and_errors():
  vd = vgroup.VarDict(variable_names=("a", "b"), num_states=np.array([2, 3]))
  fg = fgraph.FactorGraph(vd)
  fg.add_factors(factor.EnumFactor(variables=[vd["a"], vd["b"]],
                                   factor_configs=np.array([[0, 0], [1, 2]]),
                                   log_potentials=np.zeros(2)))
  bp_state = fg.bp_state
  bp_state.evidence[vd] = {"a": np.array([1.0, 2.0]), "b": np.array([3.0, 4.0, 5.0])}
  np.testing.assert_allclose(bp_state.evidence.value, [1, 2, 3, 4, 5])
  with pytest.raises(ValueError):
    bp_state.evidence[vd["b"]] = np.zeros(2)
  with pytest.raises(ValueError):
    fg.add_factors(factor.EnumFactor(variables=[vd["a"], vd["b"]],
                                     factor_configs=np.array([[0, 0]]),
                                     log_potentials=np.zeros(1)))


def test_run_without_gpu_fails_loudly():
  import torch  # pylint: disable=g-import-not-at-top
  if torch.cuda.is_available():
    pytest.skip("a GPU is visible")
  from pgmax_amd import infer  # pylint: disable=g-import-not-at-top
  fg, _, _ = graphs.ising_torus(side=3)
  bp = infer.BP(fg.bp_state, temperature=1.0)
  with pytest.raises(RuntimeError, match="GPU"):
    bp.run(bp.init(), num_iters=2)


class _FakeHandle:
  """Stands in for a native object in the cache tests (no GPU needed)."""


def test_op_cache_keys_on_contents_and_evicts():
  """The plugin-API cache keys wiring arrays by their values: a new array
  with the same shape (possibly at a recycled id) but other values gets its
  own entry, equal values share one, and the LRU bound closes evicted
  native handles."""
  from pgmax_amd import _lib  # pylint: disable=g-import-not-at-top
  from pgmax_amd import engine  # pylint: disable=g-import-not-at-top
  a = np.arange(12, dtype=np.int32)
  key_a = engine._content_key(a)  # pylint: disable=protected-access
  del a
  b = np.arange(12, dtype=np.int32)[::-1].copy()   # same shape, other values
  assert engine._content_key(b) != key_a  # pylint: disable=protected-access
  assert engine._content_key(np.arange(12, dtype=np.int64)) == key_a  # pylint: disable=protected-access

  closed = []

  class Handle(_lib.NativeHandle):
    def close(self):
      closed.append(self)

  cache = engine._OpCache(max_entries=2)  # pylint: disable=protected-access
  h = [Handle() for _ in range(3)]
  for i, x in enumerate(h):
    cache.put(("k", i), x)
  assert len(cache) == 2 and closed == [h[0]]
  assert cache.get(("k", 0)) is None and cache.get(("k", 2)) is h[2]
  cache.put(("k", 3), _FakeHandle())   # plain objects are dropped without close
  assert closed == [h[0], h[1]]
  cache.clear()
  assert closed == [h[0], h[1], h[2]] and len(cache) == 0


def test_native_handle_closes_once_and_skips_finalising_interpreter():
  """close() releases a handle exactly once; the atexit registry forgets it."""
  from pgmax_amd import _lib  # pylint: disable=g-import-not-at-top
  calls = []

  class FakeLib:
    def fake_destroy(self, h):
      calls.append(h)
      return 0

  class Handle(_lib.NativeHandle):
    _destroy_symbol = "fake_destroy"

  real_load = _lib.load
  _lib.load = FakeLib
  try:
    h = Handle()
    h._adopt(7)  # pylint: disable=protected-access
    assert h in set(_lib._LIVE)  # pylint: disable=protected-access
    with h:
      pass
    h.close()
    assert calls == [7] and h not in set(_lib._LIVE)  # pylint: disable=protected-access
  finally:
    _lib.load = real_load
