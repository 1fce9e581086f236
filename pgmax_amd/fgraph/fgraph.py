"""Factor graph construction and flat-layout compilation (pgmax/fgraph/fgraph.py).

The flat layout is what the HIP engine consumes: variable states numbered in
variable-group order, messages and log-potentials grouped by factor type in
registry order (Enum, OR, AND, Pool) and then by factor-group insertion order
(pgmax/fgraph/fgraph.py:185-232).
"""

import collections
import copy
import dataclasses
import functools
import types
from typing import Any, Dict, FrozenSet, List, Mapping, OrderedDict, Sequence, Set, Tuple, Type, Union

import numpy as np

from pgmax_amd import factor
from pgmax_amd import fgroup
from pgmax_amd import vgroup
from pgmax_amd.factor import FAC_TO_VAR_UPDATES
from pgmax_amd.utils import cached_property


class VarsToStarts(dict):
  """``{(hash, num_states): first_state}`` with a bulk lookup.

  Behaves as the reference's plain dict (pgmax/fgraph/fgraph.py:111-124) and
  adds ``lookup_many`` so wiring compilation can resolve millions of
  variables with one ``searchsorted`` instead of a Python loop.
  """

  def _build_index(self):
    keys = np.array(list(self.keys()), dtype=np.int64).reshape(-1, 2)
    vals = np.fromiter(self.values(), dtype=np.int64, count=len(self))
    order = np.argsort(keys[:, 0], kind="stable")
    self._sorted_hashes = keys[order, 0]
    self._sorted_states = keys[order, 1]
    self._sorted_starts = vals[order]
    self._index_size = len(self)

  def lookup_many(self, hashes: np.ndarray, num_states: np.ndarray) -> np.ndarray:
    if getattr(self, "_index_size", -1) != len(self):
      self._build_index()
    pos = np.searchsorted(self._sorted_hashes, hashes)
    pos_c = np.minimum(pos, max(len(self._sorted_hashes) - 1, 0))
    ok = (
        (pos < len(self._sorted_hashes))
        & (self._sorted_hashes[pos_c] == hashes)
        & (self._sorted_states[pos_c] == num_states)
    )
    if not np.all(ok):
      bad = int(np.nonzero(~ok)[0][0])
      raise KeyError((hashes[bad], num_states[bad]))
    return self._sorted_starts[pos_c]


@dataclasses.dataclass(frozen=True, eq=False)
class FactorGraphState:
  """Compiled factor graph (pgmax/fgraph/fgraph.py:34-78)."""

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
    for field in self.__dataclass_fields__:
      value = getattr(self, field)
      if isinstance(value, np.ndarray):
        value.flags.writeable = False
      if isinstance(value, Mapping) and not isinstance(value, VarsToStarts):
        object.__setattr__(self, field, types.MappingProxyType(value))


@dataclasses.dataclass
class FactorGraph:
  """A factor graph built from variable groups and factor groups (pgmax/fgraph/fgraph.py:81-371)."""

  variable_groups: Union[vgroup.VarGroup, Sequence[vgroup.VarGroup]]

  def __post_init__(self):
    if isinstance(self.variable_groups, vgroup.VarGroup):
      self.variable_groups = [self.variable_groups]
    self._factor_types_to_groups: OrderedDict[
        Type[factor.Factor], List[fgroup.FactorGroup]
    ] = collections.OrderedDict([(t, []) for t in FAC_TO_VAR_UPDATES])
    self._factor_types_to_variables_for_factors: OrderedDict[
        Type[factor.Factor], Set[FrozenSet[Any]]
    ] = collections.OrderedDict([(t, set()) for t in FAC_TO_VAR_UPDATES])
    # Variable states are numbered group by group (pgmax/fgraph/fgraph.py:111-124).
    self._vars_to_starts = VarsToStarts()
    total = 0
    for variable_group in self.variable_groups:
      vg_num_states = variable_group.num_states.flatten()
      vg_cumsum = np.insert(np.cumsum(vg_num_states), 0, 0)
      self._vars_to_starts.update(
          zip(variable_group.variables, total + vg_cumsum[:-1])
      )
      total += int(vg_cumsum[-1])
    self._num_var_states = total

  def __hash__(self) -> int:
    return hash(tuple(
        g for groups in self._factor_types_to_groups.values() for g in groups
    ))

  def add_factors(
      self,
      factors: Union[
          factor.Factor,
          fgroup.FactorGroup,
          Sequence[Union[factor.Factor, fgroup.FactorGroup]],
      ],
  ) -> None:
    """Add factors or factor groups (pgmax/fgraph/fgraph.py:137-183).

    Raises:
      ValueError: a factor has duplicate variables, or a factor of the same
        type over the same variable set already exists.
    """
    if isinstance(factors, list):
      for this_factor in factors:
        self.add_factors(this_factor)
      return None
    if isinstance(factors, fgroup.FactorGroup):
      factor_group = factors
    elif isinstance(factors, factor.Factor):
      factor_group = fgroup.SingleFactorGroup(
          variables_for_factors=[factors.variables], single_factor=factors
      )
    else:
      raise ValueError(f"Cannot add {type(factors)} to a FactorGraph")
    factor_type = factor_group.factor_type
    seen = self._factor_types_to_variables_for_factors[factor_type]
    for var_names_for_factor in factor_group.variables_for_factors:
      var_names = frozenset(var_names_for_factor)
      if len(var_names_for_factor) != len(var_names):
        raise ValueError(
            f"A Factor of type {factor_type} involving variables"
            f" {var_names_for_factor} contains variables duplicates."
        )
      if var_names in seen:
        raise ValueError(
            f"A Factor of type {factor_type} involving variables"
            f" {var_names} already exists. Please merge the corresponding"
            " factors."
        )
      seen.add(var_names)
    self._factor_types_to_groups[factor_type].append(factor_group)

  @functools.lru_cache()
  def compute_offsets(self) -> None:
    """Message and potential offsets per type and group (pgmax/fgraph/fgraph.py:185-232)."""
    self._factor_type_to_msgs_range = collections.OrderedDict()
    self._factor_group_to_msgs_starts = collections.OrderedDict()
    self._factor_type_to_potentials_range = collections.OrderedDict()
    self._factor_group_to_potentials_starts = collections.OrderedDict()
    n_states = 0
    n_configs = 0
    for factor_type, groups in self._factor_types_to_groups.items():
      type_states_start = n_states
      type_configs_start = n_configs
      for factor_group in groups:
        self._factor_group_to_msgs_starts[factor_group] = n_states
        self._factor_group_to_potentials_starts[factor_group] = n_configs
        n_states += int(factor_group.factor_edges_num_states.sum())
        n_configs += int(factor_group.factor_group_log_potentials.shape[0])
      self._factor_type_to_msgs_range[factor_type] = (type_states_start, n_states)
      self._factor_type_to_potentials_range[factor_type] = (
          type_configs_start, n_configs,
      )
    self._total_factor_num_states = n_states
    self._total_factor_num_configs = n_configs

  @cached_property
  def wiring(self) -> OrderedDict[Type[factor.Factor], factor.Wiring]:
    """Per-type wiring (pgmax/fgraph/fgraph.py:234-256)."""
    wiring = collections.OrderedDict()
    for factor_type, groups in self._factor_types_to_groups.items():
      wiring[factor_type] = factor_type.concatenate_wirings(
          [g.compile_wiring(self._vars_to_starts) for g in groups]
      )
    return wiring

  @cached_property
  def log_potentials(self) -> OrderedDict[Type[factor.Factor], np.ndarray]:
    """Per-type flat log-potentials (pgmax/fgraph/fgraph.py:258-283)."""
    out = collections.OrderedDict()
    for factor_type, groups in self._factor_types_to_groups.items():
      if not groups:
        out[factor_type] = np.empty((0,))
      else:
        out[factor_type] = np.concatenate(
            [np.asarray(g.factor_group_log_potentials) for g in groups]
        )
    return out

  @cached_property
  def factors(self) -> OrderedDict[Type[factor.Factor], Tuple[factor.Factor, ...]]:
    """pgmax/fgraph/fgraph.py:285-308."""
    print(
        "Factors have not been added to the factor graph yet, this may take a"
        " while..."
    )
    out = collections.OrderedDict()
    for factor_type, groups in self._factor_types_to_groups.items():
      out[factor_type] = tuple(f for g in groups for f in g.factors)
    return out

  @property
  def factor_groups(self) -> OrderedDict[Type[factor.Factor], List[fgroup.FactorGroup]]:
    return self._factor_types_to_groups

  @property
  def evidence_to_vars(self) -> np.ndarray:
    """Variable index of each evidence entry (pgmax/fgraph/fgraph.py:317-326)."""
    starts = np.fromiter(self._vars_to_starts.values(), dtype=np.int64)
    bounds = np.append(starts, self._num_var_states)
    return np.repeat(np.arange(starts.shape[0], dtype=np.int64), np.diff(bounds))

  @cached_property
  def fg_state(self) -> FactorGraphState:
    """pgmax/fgraph/fgraph.py:328-356."""
    self.compute_offsets()
    log_potentials = np.concatenate(
        [self.log_potentials[t] for t in self.log_potentials.keys()]
    )
    return FactorGraphState(
        variable_groups=self.variable_groups,
        vars_to_starts=self._vars_to_starts,
        num_var_states=self._num_var_states,
        total_factor_num_states=self._total_factor_num_states,
        factor_type_to_msgs_range=copy.copy(self._factor_type_to_msgs_range),
        factor_type_to_potentials_range=copy.copy(
            self._factor_type_to_potentials_range
        ),
        factor_group_to_potentials_starts=copy.copy(
            self._factor_group_to_potentials_starts
        ),
        log_potentials=log_potentials,
        evidence_to_vars=self.evidence_to_vars,
        wiring=self.wiring,
    )

  @property
  def bp_state(self) -> Any:
    """pgmax/fgraph/fgraph.py:358-371."""
    self.compute_offsets()
    from pgmax_amd.infer import bp_state  # pylint: disable=g-import-not-at-top
    return bp_state.BPState(
        log_potentials=bp_state.LogPotentials(fg_state=self.fg_state),
        ftov_msgs=bp_state.FToVMessages(fg_state=self.fg_state),
        evidence=bp_state.Evidence(fg_state=self.fg_state),
    )
