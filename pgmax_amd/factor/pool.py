"""Pool factors (pgmax/factor/pool.py).

A pool factor over (c1..cn, i) allows all-zero, or i = 1 with exactly one
choice set (pgmax/factor/pool.py:86-99). Its wiring is the logical wiring
with offset +1 (pgmax/factor/pool.py:116-180). The registered update calls
``pgx_fac_to_var`` (csrc/pgx_engine.hip) following pgmax/factor/pool.py:275-474.
"""

import dataclasses
from typing import Any, Dict, Hashable, List, Mapping, Sequence, Tuple
import warnings

import numpy as np

from pgmax_amd.factor import factor
from pgmax_amd.factor import logical


@dataclasses.dataclass(frozen=True, eq=False)
class PoolWiring(factor.Wiring):
  """Wiring for PoolFactors (pgmax/factor/pool.py:31-83)."""

  pool_choices_edge_states: np.ndarray
  pool_indicators_edge_states: np.ndarray

  def get_inference_arguments(self) -> Dict[str, Any]:
    if self.pool_choices_edge_states.shape[0] > 0:
      idx = self.pool_choices_edge_states[:, 0]
      n = self.pool_indicators_edge_states.shape[0]
      if factor.count_distinct(idx) != n:
        raise ValueError(
            f"The PoolWiring must have {n} different PoolFactor indices"
        )
      if idx.max() >= n:
        raise ValueError(f"The highest PoolFactor index must be {n - 1}")
    return {
        "pool_choices_factor_indices": self.pool_choices_edge_states[..., 0],
        "pool_choices_msg_indices": self.pool_choices_edge_states[..., 1],
        "pool_indicators_edge_states": self.pool_indicators_edge_states,
    }


@dataclasses.dataclass(frozen=True, eq=False)
class PoolFactor(factor.Factor):
  """Pool factor (pgmax/factor/pool.py:86-114)."""

  log_potentials: np.ndarray = dataclasses.field(
      init=False, default_factory=lambda: np.empty((0,))
  )

  def __post_init__(self):
    if len(self.variables) < 2:
      raise ValueError(
          "A PoolFactor requires at least one pool choice and one pool "
          "indicator."
      )
    if not np.all([variable[1] == 2 for variable in self.variables]):
      raise ValueError("All the variables in a PoolFactor should all be binary")

  @staticmethod
  def concatenate_wirings(wirings: Sequence[PoolWiring]) -> PoolWiring:
    """Concatenate via the logical wiring (pgmax/factor/pool.py:116-149)."""
    if not wirings:
      return PoolWiring(
          var_states_for_edges=np.empty((0, 3), dtype=np.int64),
          pool_choices_edge_states=np.empty((0, 2), dtype=np.int64),
          pool_indicators_edge_states=np.empty((0,), dtype=np.int64),
      )
    lw = logical.LogicalFactor.concatenate_wirings([
        logical.LogicalWiring(
            var_states_for_edges=w.var_states_for_edges,
            parents_edge_states=w.pool_choices_edge_states,
            children_edge_states=w.pool_indicators_edge_states,
            edge_states_offset=1,
        )
        for w in wirings
    ])
    return PoolWiring(
        var_states_for_edges=lw.var_states_for_edges,
        pool_choices_edge_states=lw.parents_edge_states,
        pool_indicators_edge_states=lw.children_edge_states,
    )

  @staticmethod
  def compile_wiring(
      factor_edges_num_states: np.ndarray,
      variables_for_factors: Sequence[List[Tuple[int, int]]],
      factor_sizes: np.ndarray,
      vars_to_starts: Mapping[Tuple[int, int], int],
  ) -> PoolWiring:
    """pgmax/factor/pool.py:151-180."""
    lw = logical.LogicalFactor.compile_wiring(
        factor_edges_num_states=factor_edges_num_states,
        variables_for_factors=variables_for_factors,
        factor_sizes=factor_sizes,
        vars_to_starts=vars_to_starts,
        edge_states_offset=1,
    )
    return PoolWiring(
        var_states_for_edges=lw.var_states_for_edges,
        pool_choices_edge_states=lw.parents_edge_states,
        pool_indicators_edge_states=lw.children_edge_states,
    )

  @staticmethod
  def compute_energy(
      edge_states_one_hot_decoding: np.ndarray,
      pool_choices_factor_indices: np.ndarray,
      pool_choices_msg_indices: np.ndarray,
      pool_indicators_edge_states: np.ndarray,
      log_potentials=None,
  ) -> float:
    """pgmax/factor/pool.py:182-239."""
    n = pool_indicators_edge_states.shape[0]
    choices = np.zeros((n,))
    np.add.at(
        choices,
        pool_choices_factor_indices,
        edge_states_one_hot_decoding[pool_choices_msg_indices + 1],
    )
    indicators = edge_states_one_hot_decoding[pool_indicators_edge_states + 1]
    return np.inf if np.any(choices != indicators) else 0.0

  @staticmethod
  def compute_factor_energy(
      variables: List[Hashable], vars_to_map_states: Dict[Hashable, Any], **kwargs
  ) -> float:
    """pgmax/factor/pool.py:241-272."""
    decoded = np.array([vars_to_map_states[var] for var in variables])
    if np.sum(decoded[:-1]) != decoded[-1]:
      warnings.warn(
          f"Invalid decoding for Pool factor {variables} "
          f"with pool choices set to {decoded[:-1]} "
          f"and pool indicators set to {decoded[-1]}!"
      )
      return np.inf
    return 0.0


def pass_pool_fac_to_var_messages(
    vtof_msgs,
    pool_choices_factor_indices,
    pool_choices_msg_indices,
    pool_indicators_edge_states,
    temperature: float,
    normalize: bool,
    log_potentials=None,
):
  """Pool factor -> variable messages on the GPU (pgmax/factor/pool.py:275-474)."""
  from pgmax_amd import _lib  # pylint: disable=g-import-not-at-top
  return _lib.pool_fac_to_var(
      vtof_msgs, pool_choices_factor_indices, pool_choices_msg_indices,
      pool_indicators_edge_states, temperature, normalize,
  )
