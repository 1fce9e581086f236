"""Logical OR / AND factors (pgmax/factor/logical.py).

Wiring compilation is the vectorised form of pgmax/factor/logical.py:208-291
and :457-488. The registered update calls ``pgx_fac_to_var``
(csrc/pgx_engine.hip, per-factor code in csrc/pgx_device.h) whose closed forms
follow pgmax/factor/logical.py:491-779.
"""

import dataclasses
from typing import Any, Dict, Hashable, List, Mapping, Sequence, Tuple
import warnings

import numpy as np

from pgmax_amd.factor import factor

# Below this temperature the closed forms add lower bounds for stability
# (pgmax/factor/logical.py:31-33).
TEMPERATURE_STABILITY_THRE = 0.5


@dataclasses.dataclass(frozen=True, eq=False)
class LogicalWiring(factor.Wiring):
  """Wiring for LogicalFactors (pgmax/factor/logical.py:36-104).

  Attributes:
    parents_edge_states: (num_parents, 2): factor index, message index of the
      parent's state addressed without offset.
    children_edge_states: (num_factors,): message index of the child's state
      addressed without offset.
    edge_states_offset: +1 for OR, -1 for AND.
  """

  parents_edge_states: np.ndarray
  children_edge_states: np.ndarray
  edge_states_offset: int

  def get_inference_arguments(self) -> Dict[str, Any]:
    """Arguments of the update function (pgmax/factor/logical.py:76-104)."""
    if self.children_edge_states.shape[0] > 0:
      idx = self.parents_edge_states[:, 0]
      n = self.children_edge_states.shape[0]
      if factor.count_distinct(idx) != n:
        raise ValueError(
            f"The LogicalWiring must have {n} different LogicalFactor indices"
        )
      if idx.max() >= n:
        raise ValueError(f"The highest LogicalFactor index must be {n - 1}")
      if self.edge_states_offset not in (1, -1):
        raise ValueError(
            "The LogicalWiring's edge_states_offset must be 1 (for OR) and -1"
            f" (for AND), but is {self.edge_states_offset}"
        )
    return {
        "parents_factor_indices": self.parents_edge_states[..., 0],
        "parents_msg_indices": self.parents_edge_states[..., 1],
        "children_edge_states": self.children_edge_states,
        "edge_states_offset": self.edge_states_offset,
    }


@dataclasses.dataclass(frozen=True, eq=False)
class LogicalFactor(factor.Factor):
  """OR/AND factor over (p1..pn, c) (pgmax/factor/logical.py:107-137)."""

  log_potentials: np.ndarray = dataclasses.field(
      init=False, default_factory=lambda: np.empty((0,))
  )
  edge_states_offset: int = dataclasses.field(init=False)

  def __post_init__(self):
    if len(self.variables) < 2:
      raise ValueError(
          "A LogicalFactor requires at least one parent variable and one child"
          " variable"
      )
    if not np.all([variable[1] == 2 for variable in self.variables]):
      raise ValueError("All the variables in a LogicalFactor should be binary")

  @staticmethod
  def concatenate_wirings(wirings: Sequence[LogicalWiring]) -> LogicalWiring:
    """Concatenate LogicalWirings (pgmax/factor/logical.py:139-206)."""
    if not wirings:
      return LogicalWiring(
          var_states_for_edges=np.empty((0, 3), dtype=np.int64),
          parents_edge_states=np.empty((0, 2), dtype=np.int64),
          children_edge_states=np.empty((0,), dtype=np.int64),
          edge_states_offset=1,
      )
    nf_cumsum = np.insert(
        np.cumsum([w.parents_edge_states[-1, 0] + 1 for w in wirings]), 0, 0
    )[:-1]
    vsfe = factor.concatenate_var_states_for_edges(
        [w.var_states_for_edges for w in wirings]
    )
    es_cumsum = np.insert(
        np.cumsum([w.var_states_for_edges.shape[0] for w in wirings]), 0, 0
    )[:-1]
    parents, children = [], []
    for ww, w in enumerate(wirings):
      offsets = np.array([[nf_cumsum[ww], es_cumsum[ww]]], dtype=np.int64)
      parents.append(w.parents_edge_states + offsets)
      children.append(w.children_edge_states + offsets[:, 1])
    return LogicalWiring(
        var_states_for_edges=vsfe,
        parents_edge_states=np.concatenate(parents, axis=0),
        children_edge_states=np.concatenate(children, axis=0),
        edge_states_offset=wirings[0].edge_states_offset,
    )

  @staticmethod
  def compile_wiring(
      factor_edges_num_states: np.ndarray,
      variables_for_factors: Sequence[List[Tuple[int, int]]],
      factor_sizes: np.ndarray,
      vars_to_starts: Mapping[Tuple[int, int], int],
      edge_states_offset: int,
  ) -> LogicalWiring:
    """Compile LogicalWiring (pgmax/factor/logical.py:208-291, :457-488).

    Factor ``f`` owns edge states ``[s_f, s_f + 2 * size_f)`` with
    ``s_f = cumsum(2 * sizes)``; parent ``k`` is addressed at
    ``s_f + relevant + 2k`` and the child at ``s_f + 2 n_f + relevant`` with
    ``relevant = (1 - offset) // 2``.
    """
    factor_sizes = np.asarray(factor_sizes, dtype=np.int64)
    # the checks LogicalFactor makes per factor, over the whole group at once
    if np.any(factor_sizes < 2):
      raise ValueError(
          "A LogicalFactor requires at least one parent variable and one child"
          " variable"
      )
    if np.any(np.asarray(factor_edges_num_states) != 2):
      raise ValueError("All the variables in a LogicalFactor should be binary")
    first_var_state_by_edges, factor_indices = factor.lookup_var_starts(
        variables_for_factors, vars_to_starts
    )
    n_edges = first_var_state_by_edges.shape[0]
    edge_cumsum = np.arange(0, 2 * n_edges + 2, 2, dtype=np.int64)
    var_states_for_edges = factor.compile_var_states_for_edges(
        edge_cumsum, first_var_state_by_edges, factor_indices
    )
    num_parents = factor_sizes - 1
    relevant_state = (-edge_states_offset + 1) // 2
    fstart = np.insert(np.cumsum(2 * factor_sizes), 0, 0)
    num_f = factor_sizes.shape[0]
    pf = np.repeat(np.arange(num_f, dtype=np.int64), num_parents)
    pcum = np.insert(np.cumsum(num_parents), 0, 0)
    k = np.arange(pcum[-1], dtype=np.int64) - pcum[:-1][pf]
    parents_edge_states = np.empty((int(pcum[-1]), 2), dtype=np.int64)
    parents_edge_states[:, 0] = pf
    parents_edge_states[:, 1] = fstart[:-1][pf] + relevant_state + 2 * k
    children_edge_states = fstart[:-1] + 2 * num_parents + relevant_state
    return LogicalWiring(
        var_states_for_edges=var_states_for_edges,
        parents_edge_states=parents_edge_states,
        children_edge_states=children_edge_states.astype(np.int64),
        edge_states_offset=edge_states_offset,
    )

  @staticmethod
  def compute_energy(
      edge_states_one_hot_decoding: np.ndarray,
      parents_factor_indices: np.ndarray,
      parents_msg_indices: np.ndarray,
      children_edge_states: np.ndarray,
      edge_states_offset: int,
      log_potentials=None,
  ) -> float:
    """Energy contribution of LogicalFactors (pgmax/factor/logical.py:293-358)."""
    n = children_edge_states.shape[0]
    parents_decoded = np.ones((n,))
    np.multiply.at(
        parents_decoded,
        parents_factor_indices,
        edge_states_one_hot_decoding[parents_msg_indices],
    )
    children_decoded = edge_states_one_hot_decoding[children_edge_states]
    return np.inf if np.any(parents_decoded != children_decoded) else 0.0


@dataclasses.dataclass(frozen=True, eq=False)
class ORFactor(LogicalFactor):
  """c = OR(p1..pn) (pgmax/factor/logical.py:361-406)."""

  edge_states_offset: int = dataclasses.field(init=False, default=1)

  @staticmethod
  def compute_factor_energy(
      variables: List[Hashable], vars_to_map_states: Dict[Hashable, Any], **kwargs
  ) -> float:
    decoded = np.array([vars_to_map_states[var] for var in variables])
    if np.any(decoded[:-1]) != decoded[-1]:
      warnings.warn(
          f"Invalid decoding for OR factor {variables} "
          f"with parents set to {decoded[:-1]} "
          f"and child set to {decoded[-1]}!"
      )
      return np.inf
    return 0.0


@dataclasses.dataclass(frozen=True, eq=False)
class ANDFactor(LogicalFactor):
  """c = AND(p1..pn) (pgmax/factor/logical.py:409-454)."""

  edge_states_offset: int = dataclasses.field(init=False, default=-1)

  @staticmethod
  def compute_factor_energy(
      variables: List[Hashable], vars_to_map_states: Dict[Hashable, Any], **kwargs
  ) -> float:
    decoded = np.array([vars_to_map_states[var] for var in variables])
    if np.all(decoded[:-1]) != decoded[-1]:
      warnings.warn(
          f"Invalid decoding for AND factor {variables} "
          f"with parents set to {decoded[:-1]} "
          f"and child set to {decoded[-1]}!"
      )
      return np.inf
    return 0.0


def pass_logical_fac_to_var_messages(
    vtof_msgs,
    parents_factor_indices,
    parents_msg_indices,
    children_edge_states,
    edge_states_offset: int,
    temperature: float,
    normalize: bool,
    log_potentials=None,
):
  """Logical factor -> variable messages on the GPU (pgmax/factor/logical.py:491-779)."""
  from pgmax_amd import _lib  # pylint: disable=g-import-not-at-top
  return _lib.logical_fac_to_var(
      vtof_msgs, parents_factor_indices, parents_msg_indices,
      children_edge_states, edge_states_offset, temperature, normalize,
  )
