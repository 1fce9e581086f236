"""Factor graphs and their flat layout (the interface of pgmax/fgraph/fgraph.py).

The flat layout is what the HIP engine consumes (pgmax/fgraph/fgraph.py:111-124
and :185-232 fix it): variable states numbered in variable-group order;
messages and log-potentials grouped by factor type in registry order (Enum,
OR, AND, Pool), then by the order the factor groups were added.

Everything scales with array operations, so groups of millions of factors
compile in seconds:

* ``VarIndex`` resolves (hash, num_states) variables to their first state
  with one ``searchsorted`` over runs of consecutive hashes (a variable
  group's hashes are its base hash plus 0, 1, 2, ...).
* ``_FactorSets`` keeps, per factor type, a 64-bit fingerprint of every
  factor's variable set; a new group is checked against them (and against
  itself) in bulk, and only fingerprint matches are compared exactly.
* Derived state (offsets, wiring, potentials, ``fg_state``) is cached per
  version of the graph; adding factors starts a new version.
"""

import collections
import dataclasses
import itertools
import types
from typing import Any, Dict, Iterator, List, Mapping, Optional, OrderedDict, Sequence, Tuple, Type, Union

import numpy as np

from pgmax_amd import factor
from pgmax_amd import fgroup
from pgmax_amd import vgroup
from pgmax_amd.factor import FAC_TO_VAR_UPDATES


class VarIndex(Mapping):
  """``(hash, num_states) -> first state`` over a graph's variable groups.

  Reads as the reference's ``vars_to_starts`` dict
  (pgmax/fgraph/fgraph.py:111-124) and adds ``lookup`` for arrays of
  variables. Variables are numbered in group order; ``starts[i]`` is the
  first state of variable ``i`` and ``num_states[i]`` its state count.
  """

  def __init__(self, variable_groups: Sequence[vgroup.VarGroup]):
    hashes = [np.asarray(g.variable_hashes, dtype=np.int64).reshape(-1) for g in variable_groups]
    states = [np.asarray(g.num_states, dtype=np.int64).reshape(-1) for g in variable_groups]
    self.hashes = np.concatenate(hashes) if hashes else np.zeros(0, np.int64)
    self.num_states = np.concatenate(states) if states else np.zeros(0, np.int64)
    bounds = np.concatenate([[0], np.cumsum(self.num_states)]).astype(np.int64)
    self.starts = bounds[:-1]
    self.num_var_states = int(bounds[-1])
    n = self.hashes.shape[0]
    # maximal runs h, h + 1, h + 2, ... in variable order, sorted by first hash
    breaks = np.flatnonzero(np.diff(self.hashes) != 1) + 1 if n else np.zeros(0, np.int64)
    first = np.concatenate([[0], breaks]).astype(np.int64) if n else np.zeros(0, np.int64)
    length = np.diff(np.append(first, n))
    order = np.argsort(self.hashes[first], kind="stable")
    self._run_hash = self.hashes[first][order]
    self._run_var = first[order]
    self._run_len = length[order]
    if np.any(self._run_hash[:-1] + self._run_len[:-1] > self._run_hash[1:]):
      raise ValueError("Two variables of the FactorGraph's variable groups share a hash.")

  def index(self, hashes: np.ndarray, num_states: np.ndarray) -> np.ndarray:
    """Variable numbers of (hash, num_states) arrays; -1 where absent."""
    hashes = np.asarray(hashes, dtype=np.int64)
    num_states = np.asarray(num_states, dtype=np.int64)
    if self._run_hash.shape[0] == 0:
      return np.full(hashes.shape, -1, np.int64)
    run = np.searchsorted(self._run_hash, hashes, side="right") - 1
    run_c = np.maximum(run, 0)
    offset = hashes - self._run_hash[run_c]
    inside = (run >= 0) & (offset < self._run_len[run_c])
    var = np.where(inside, self._run_var[run_c] + offset, 0)
    found = inside & (self.num_states[var] == num_states)
    return np.where(found, var, -1)

  def lookup(self, hashes: np.ndarray, num_states: np.ndarray) -> np.ndarray:
    """First states of (hash, num_states) arrays; KeyError names an absent one."""
    var = self.index(hashes, num_states)
    missing = np.flatnonzero(var < 0)
    if missing.size:
      k = int(missing[0])
      raise KeyError((int(np.asarray(hashes).reshape(-1)[k]),
                      int(np.asarray(num_states).reshape(-1)[k])))
    return self.starts[var]

  @staticmethod
  def _as_variable(key: Any) -> Optional[Tuple[int, int]]:
    if not isinstance(key, tuple) or len(key) != 2:
      return None
    try:
      return int(key[0]), int(key[1])
    except (TypeError, ValueError):
      return None

  def __getitem__(self, key: Any) -> int:
    var = self._as_variable(key)
    if var is None:
      raise KeyError(key)
    i = int(self.index(np.array([var[0]]), np.array([var[1]]))[0])
    if i < 0:
      raise KeyError(key)
    return int(self.starts[i])

  def __contains__(self, key: Any) -> bool:
    var = self._as_variable(key)
    return var is not None and int(
        self.index(np.array([var[0]]), np.array([var[1]]))[0]) >= 0

  def __iter__(self) -> Iterator[Tuple[int, int]]:
    return zip(self.hashes.tolist(), self.num_states.tolist())

  def __len__(self) -> int:
    return int(self.hashes.shape[0])


# ---------------------------------------------------------------------------
# Variable-set fingerprints
# ---------------------------------------------------------------------------

_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_GOLD = np.uint64(0x9E3779B97F4A7C15)
_POLY = np.uint64(0xD6E8FEB86659FD93)


def _mix(z: np.ndarray) -> np.ndarray:
  """A 64-bit finaliser (xor-shift-multiply), elementwise on uint64."""
  z = (z ^ (z >> np.uint64(30))) * _M1
  z = (z ^ (z >> np.uint64(27))) * _M2
  return z ^ (z >> np.uint64(31))


def _variable_keys(hashes: np.ndarray, num_states: np.ndarray) -> np.ndarray:
  with np.errstate(over="ignore"):
    h = hashes.astype(np.uint64)
    return _mix(h * _GOLD + _mix(num_states.astype(np.uint64) + _GOLD))


def _sort_rows(rows: np.ndarray) -> np.ndarray:
  """Each row sorted; short rows by a compare-exchange network over columns
  (odd-even transposition), which beats a per-row sort by far."""
  a = rows.shape[1]
  if a > 8:
    return np.sort(rows, axis=1)
  cols = [rows[:, j] for j in range(a)]
  for step in range(a):
    for j in range(step % 2, a - 1, 2):
      lo, hi = np.minimum(cols[j], cols[j + 1]), np.maximum(cols[j], cols[j + 1])
      cols[j], cols[j + 1] = lo, hi
  return np.stack(cols, axis=1) if a else rows


class _SetPrints:
  """Per-factor canonical (sorted) variable keys and their fingerprints."""

  def __init__(self, edges: "fgroup.FactorEdges"):
    keys = _variable_keys(edges.hashes, edges.num_states)
    sizes = edges.sizes
    n_f = edges.num_factors
    max_size = int(sizes.max()) if n_f else 0
    with np.errstate(over="ignore"):
      powers = np.cumprod(np.full(max(max_size, 1), _POLY, dtype=np.uint64))
    if edges.arity is not None:
      a = edges.arity
      rows = _sort_rows(keys.reshape(n_f, a))
      repeated = (rows[:, 1:] == rows[:, :-1]).any(axis=1) if a > 1 else np.zeros(n_f, bool)
      with np.errstate(over="ignore"):
        prints = (rows * powers[:a]).sum(axis=1, dtype=np.uint64)
    else:
      factor_of = np.repeat(np.arange(n_f, dtype=np.int64), sizes)
      order = np.lexsort((keys, factor_of))
      sorted_keys = keys[order]
      same = (sorted_keys[1:] == sorted_keys[:-1]) & (factor_of[1:] == factor_of[:-1])
      repeated = np.zeros(n_f, bool)
      repeated[factor_of[1:][same]] = True
      within = np.arange(edges.num_edges, dtype=np.int64) - edges.offsets[:-1][factor_of]
      with np.errstate(over="ignore"):
        running = np.concatenate([[np.uint64(0)],
                                  np.cumsum(sorted_keys * powers[within], dtype=np.uint64)])
        prints = running[edges.offsets[1:]] - running[edges.offsets[:-1]]
    with np.errstate(over="ignore"):
      self.prints = _mix(prints + sizes.astype(np.uint64) * _GOLD)
    # a fingerprint-level repeat inside a factor is confirmed exactly later
    self.maybe_repeated = repeated


def _merge(a, b):
  """Union of two (sorted prints, refs) pairs, sorted by print."""
  if a[0].shape[0] == 0:
    return b
  prints = np.concatenate([a[0], b[0]])
  order = np.argsort(prints, kind="stable")  # two sorted runs: a linear merge
  return prints[order], np.concatenate([a[1], b[1]])[order]


class _FactorSets:
  """Variable sets of every factor of one type added so far.

  Fingerprints live in two sorted arrays: a large one and a small one that
  takes new groups and is merged into the large one when it grows, so that
  adding one factor at a time stays cheap. ``refs`` encode (group number,
  factor) as ``group << 40 | factor``.
  """

  _MERGE = 1 << 16

  def __init__(self):
    self.groups: List[fgroup.FactorGroup] = []
    empty = (np.zeros(0, np.uint64), np.zeros(0, np.int64))
    self.big, self.small = empty, empty

  def _variable_set(self, ref: int) -> frozenset:
    g = self.groups[ref >> 40]
    return frozenset(g.factor_edges.factor(ref & ((1 << 40) - 1)))

  def _existing_match(self, fp: np.ndarray) -> np.ndarray:
    """Factors of the new group whose fingerprint is already registered."""
    hit = np.zeros(fp.shape, bool)
    for prints, _ in (self.big, self.small):
      if prints.shape[0]:
        pos = np.minimum(np.searchsorted(prints, fp), prints.shape[0] - 1)
        hit |= prints[pos] == fp
    return hit

  def _registered_sets(self, fp_value: np.uint64) -> List[frozenset]:
    out = []
    for prints, refs in (self.big, self.small):
      a, z = np.searchsorted(prints, fp_value), np.searchsorted(prints, fp_value, "right")
      out.extend(self._variable_set(int(r)) for r in refs[a:z])
    return out

  def admit(self, group: "fgroup.FactorGroup", factor_type: Any) -> None:
    """Register a group's factors; ValueError on the first factor that has
    a repeated variable or whose variable set a factor of the same type
    already has (pgmax/fgraph/fgraph.py:137-183 checks in this order)."""
    edges = group.factor_edges
    sp = _SetPrints(edges)
    fp = sp.prints
    n_f = edges.num_factors
    # candidates, confirmed exactly in factor order by _raise_first:
    # (a) a factor with a repeated variable key
    inner = np.flatnonzero(sp.maybe_repeated)
    # (b) a factor whose fingerprint an earlier factor of the group has
    order = np.argsort(fp, kind="stable")
    equal_prev = np.zeros(n_f, bool)
    equal_prev[1:] = fp[order][1:] == fp[order][:-1]
    repeats = order[equal_prev]
    # (c) a factor whose fingerprint an earlier group has
    hits = np.flatnonzero(self._existing_match(fp))
    checks = np.union1d(np.union1d(inner, repeats), hits)
    if checks.size:
      self._raise_first(group, factor_type, checks, sp.maybe_repeated, fp, order)
    gid = len(self.groups)
    self.groups.append(group)
    refs = (np.int64(gid) << np.int64(40)) | order.astype(np.int64)
    self.small = _merge(self.small, (fp[order], refs))
    if self.small[0].shape[0] > max(self._MERGE, self.big[0].shape[0] // 8):
      self.big = _merge(self.big, self.small)
      self.small = (np.zeros(0, np.uint64), np.zeros(0, np.int64))

  def _raise_first(self, group, factor_type, checks, maybe_repeated, fp, order):
    """Confirms candidate factors exactly, in factor order; raises at the
    first real one (fingerprint collisions are passed over)."""
    edges = group.factor_edges
    rank = np.empty_like(order)
    rank[order] = np.arange(order.shape[0])
    for f in checks.tolist():
      variables = edges.factor(f)
      as_set = frozenset(variables)
      if maybe_repeated[f] and len(as_set) != len(variables):
        raise ValueError(
            f"A Factor of type {factor_type} involving variables"
            f" {group.given_factor(f)} contains variables duplicates."
        )
      # earlier factors of the group with the same fingerprint precede f in
      # the stable order
      r = int(rank[f])
      while r > 0 and fp[order[r - 1]] == fp[f]:
        r -= 1
        if frozenset(edges.factor(int(order[r]))) == as_set:
          self._raise_exists(group, factor_type, f)
      if as_set in self._registered_sets(fp[f]):
        self._raise_exists(group, factor_type, f)

  @staticmethod
  def _raise_exists(group, factor_type, f):
    raise ValueError(
        f"A Factor of type {factor_type} involving variables"
        f" {frozenset(group.given_factor(f))} already exists. Please merge the"
        " corresponding factors."
    )


# ---------------------------------------------------------------------------
# FactorGraph
# ---------------------------------------------------------------------------


@dataclasses.dataclass(frozen=True, eq=False)
class FactorGraphState:
  """A compiled factor graph (pgmax/fgraph/fgraph.py:34-78)."""

  variable_groups: Sequence[vgroup.VarGroup]
  vars_to_starts: Mapping[Tuple[int, int], int]
  num_var_states: int
  total_factor_num_states: int
  factor_type_to_msgs_range: OrderedDict[Type[factor.Factor], Tuple[int, int]]
  factor_type_to_potentials_range: OrderedDict[Type[factor.Factor], Tuple[int, int]]
  factor_group_to_potentials_starts: OrderedDict[fgroup.FactorGroup, int]
  log_potentials: np.ndarray
  evidence_to_vars: np.ndarray
  wiring: OrderedDict[Type[factor.Factor], factor.Wiring]

  def __post_init__(self):
    for name in self.__dataclass_fields__:
      value = getattr(self, name)
      if isinstance(value, np.ndarray):
        value.flags.writeable = False
      elif isinstance(value, dict):
        object.__setattr__(self, name, types.MappingProxyType(value))


@dataclasses.dataclass(frozen=True)
class _Layout:
  """Where each factor type and group sits in the flat message and
  potential arrays (pgmax/fgraph/fgraph.py:185-232)."""

  msgs_range: OrderedDict
  potentials_range: OrderedDict
  group_msgs_start: OrderedDict
  group_potentials_start: OrderedDict
  total_states: int
  total_configs: int


@dataclasses.dataclass(eq=False)
class FactorGraph:
  """Variable groups plus factor groups (pgmax/fgraph/fgraph.py:81-371)."""

  variable_groups: Union[vgroup.VarGroup, Sequence[vgroup.VarGroup]]

  def __post_init__(self):
    if isinstance(self.variable_groups, vgroup.VarGroup):
      self.variable_groups = [self.variable_groups]
    self._index = VarIndex(self.variable_groups)
    self._groups: OrderedDict[Type[factor.Factor], List[fgroup.FactorGroup]] = (
        collections.OrderedDict((t, []) for t in FAC_TO_VAR_UPDATES))
    self._sets = {t: _FactorSets() for t in FAC_TO_VAR_UPDATES}
    self._version = 0
    self._memo: Dict[str, Tuple[int, Any]] = {}

  def _derived(self, name: str, build) -> Any:
    """``build()`` once per version of the graph."""
    hit = self._memo.get(name)
    if hit is None or hit[0] != self._version:
      hit = (self._version, build())
      self._memo[name] = hit
    return hit[1]

  # ---- building -----------------------------------------------------------
  def add_factors(
      self,
      factors: Union[
          factor.Factor,
          fgroup.FactorGroup,
          Sequence[Union[factor.Factor, fgroup.FactorGroup]],
      ],
  ) -> None:
    """Add a factor, a factor group, or a list of them (pgmax/fgraph/fgraph.py:137-183).

    Raises:
      ValueError: a factor has a repeated variable, or a factor of the same
        type over the same variable set is already in the graph (the group
        is then not added).
    """
    if isinstance(factors, list):
      for item in factors:
        self.add_factors(item)
      return
    if isinstance(factors, fgroup.FactorGroup):
      group = factors
    elif isinstance(factors, factor.Factor):
      group = fgroup.SingleFactorGroup(
          variables_for_factors=[factors.variables], single_factor=factors
      )
    else:
      raise ValueError(f"Cannot add {type(factors)} to a FactorGraph")
    factor_type = group.factor_type
    if factor_type not in self._groups:
      raise ValueError(
          f"Factor type {factor_type} has no registered factor to variable update."
      )
    self._sets[factor_type].admit(group, factor_type)
    self._groups[factor_type].append(group)
    self._version += 1

  # ---- derived state ------------------------------------------------------
  def _layout(self) -> _Layout:
    def build():
      msgs, pots = collections.OrderedDict(), collections.OrderedDict()
      g_msgs, g_pots = collections.OrderedDict(), collections.OrderedDict()
      n_states = n_configs = 0
      for factor_type, groups in self._groups.items():
        s0, c0 = n_states, n_configs
        for group in groups:
          g_msgs[group], g_pots[group] = n_states, n_configs
          n_states += int(group.factor_edges_num_states.sum())
          n_configs += int(group.factor_group_log_potentials.shape[0])
        msgs[factor_type], pots[factor_type] = (s0, n_states), (c0, n_configs)
      return _Layout(msgs, pots, g_msgs, g_pots, n_states, n_configs)
    return self._derived("layout", build)

  def compute_offsets(self) -> None:
    """Computes the message and potential offsets (pgmax/fgraph/fgraph.py:185-232)."""
    self._layout()

  @property
  def wiring(self) -> OrderedDict[Type[factor.Factor], factor.Wiring]:
    """Per-type wiring (pgmax/fgraph/fgraph.py:234-256)."""
    def build():
      return collections.OrderedDict(
          (t, t.concatenate_wirings([g.compile_wiring(self._index) for g in groups]))
          for t, groups in self._groups.items())
    return self._derived("wiring", build)

  @property
  def log_potentials(self) -> OrderedDict[Type[factor.Factor], np.ndarray]:
    """Per-type flat log-potentials (pgmax/fgraph/fgraph.py:258-283)."""
    def build():
      return collections.OrderedDict(
          (t, np.concatenate([np.asarray(g.factor_group_log_potentials) for g in groups])
           if groups else np.empty((0,)))
          for t, groups in self._groups.items())
    return self._derived("log_potentials", build)

  @property
  def factors(self) -> OrderedDict[Type[factor.Factor], Tuple[factor.Factor, ...]]:
    """Factor objects per type, built on request (pgmax/fgraph/fgraph.py:285-308)."""
    def build():
      return collections.OrderedDict(
          (t, tuple(itertools.chain.from_iterable(g.factors for g in groups)))
          for t, groups in self._groups.items())
    return self._derived("factors", build)

  @property
  def factor_groups(self) -> OrderedDict[Type[factor.Factor], List[fgroup.FactorGroup]]:
    return self._groups

  @property
  def evidence_to_vars(self) -> np.ndarray:
    """Variable number of every evidence entry (pgmax/fgraph/fgraph.py:317-326)."""
    return np.repeat(np.arange(len(self._index), dtype=np.int64), self._index.num_states)

  @property
  def fg_state(self) -> FactorGraphState:
    """The compiled graph (pgmax/fgraph/fgraph.py:328-356)."""
    def build():
      layout = self._layout()
      lps = self.log_potentials
      return FactorGraphState(
          variable_groups=self.variable_groups,
          vars_to_starts=self._index,
          num_var_states=self._index.num_var_states,
          total_factor_num_states=layout.total_states,
          factor_type_to_msgs_range=collections.OrderedDict(layout.msgs_range),
          factor_type_to_potentials_range=collections.OrderedDict(layout.potentials_range),
          factor_group_to_potentials_starts=collections.OrderedDict(
              layout.group_potentials_start),
          log_potentials=np.concatenate([lps[t] for t in lps]),
          evidence_to_vars=self.evidence_to_vars,
          wiring=self.wiring,
      )
    return self._derived("fg_state", build)

  @property
  def bp_state(self) -> Any:
    """Initial BP state of the compiled graph (pgmax/fgraph/fgraph.py:358-371)."""
    from pgmax_amd.infer import bp_state  # pylint: disable=g-import-not-at-top
    state = self.fg_state
    return bp_state.BPState(
        log_potentials=bp_state.LogPotentials(fg_state=state),
        ftov_msgs=bp_state.FToVMessages(fg_state=state),
        evidence=bp_state.Evidence(fg_state=state),
    )
