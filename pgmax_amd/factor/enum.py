"""Enumeration factors (pgmax/factor/enum.py).

Wiring compilation is the vectorised form of pgmax/factor/enum.py:191-270 and
:363-394. The factor-to-variable update registered for EnumFactor calls
``pgx_fac_to_var`` / ``pgx_enum_rows_fac_to_var`` (csrc/pgx_engine.hip); its
semantics are pgmax/factor/enum.py:397-475.
"""

import dataclasses
from typing import Any, Dict, Hashable, List, Mapping, Sequence, Tuple
import warnings

import numpy as np

from pgmax_amd.factor import factor
from pgmax_amd.utils import NEG_INF


@dataclasses.dataclass(frozen=True, eq=False)
class EnumWiring(factor.Wiring):
  """Wiring for EnumFactors (pgmax/factor/enum.py:33-72).

  Attributes:
    factor_configs_edge_states: int array (num_rows, 2); column 0 is the
      global config index and column 1 the edge-state index of each
      (config, variable) row.
  """

  factor_configs_edge_states: np.ndarray

  def get_inference_arguments(self) -> Dict[str, Any]:
    """Arguments of the update function (pgmax/factor/enum.py:54-72)."""
    if self.factor_configs_edge_states.shape[0] == 0:
      num_val_configs = 0
    else:
      num_val_configs = int(self.factor_configs_edge_states[-1, 0]) + 1
    object.__setattr__(self, "num_val_configs", num_val_configs)
    if self.var_states_for_edges.shape[0] == 0:
      num_factors = 0
    else:
      num_factors = int(self.var_states_for_edges[-1, 2]) + 1
    object.__setattr__(self, "num_factors", num_factors)
    return {
        "factor_configs_indices": self.factor_configs_edge_states[..., 0],
        "factor_configs_edge_states": self.factor_configs_edge_states[..., 1],
        "num_val_configs": self.num_val_configs,
        "num_factors": self.num_factors,
    }


@dataclasses.dataclass(frozen=True, eq=False)
class EnumFactor(factor.Factor):
  """An enumeration factor (pgmax/factor/enum.py:75-135)."""

  factor_configs: np.ndarray
  log_potentials: np.ndarray

  def __post_init__(self):
    self.factor_configs.flags.writeable = False
    if not np.issubdtype(self.factor_configs.dtype, np.integer):
      raise ValueError(
          f"Configurations should be integers. Got {self.factor_configs.dtype}."
      )
    if not np.issubdtype(self.log_potentials.dtype, np.floating):
      raise ValueError(
          f"Potential should be floats. Got {self.log_potentials.dtype}."
      )
    if self.factor_configs.ndim != 2:
      raise ValueError(
          "factor_configs should be a 2D array containing a list of valid"
          " configurations for EnumFactor. Got a factor_configs array of shape"
          f" {self.factor_configs.shape}."
      )
    if len(self.variables) != self.factor_configs.shape[1]:
      raise ValueError(
          f"Number of variables {len(self.variables)} doesn't match given"
          f" configurations {self.factor_configs.shape}"
      )
    if self.log_potentials.shape != (self.factor_configs.shape[0],):
      raise ValueError(
          "Expected log potentials of shape"
          f" {(self.factor_configs.shape[0],)} for"
          f" ({self.factor_configs.shape[0]}) valid configurations. Got log"
          f" potentials of shape {self.log_potentials.shape}."
      )
    vars_num_states = np.array([variable[1] for variable in self.variables])
    if not np.logical_and(
        self.factor_configs >= 0, self.factor_configs < vars_num_states[None]
    ).all():
      raise ValueError("Invalid configurations for given variables")

  @staticmethod
  def concatenate_wirings(wirings: Sequence[EnumWiring]) -> EnumWiring:
    """Concatenate EnumWirings (pgmax/factor/enum.py:137-189)."""
    if not wirings:
      return EnumWiring(
          var_states_for_edges=np.empty((0, 3), dtype=np.int64),
          factor_configs_edge_states=np.empty((0, 2), dtype=np.int64),
      )
    vsfe = factor.concatenate_var_states_for_edges(
        [w.var_states_for_edges for w in wirings]
    )
    configs_cumsum = np.insert(
        np.cumsum([w.factor_configs_edge_states[-1, 0] + 1 for w in wirings]),
        0, 0,
    )[:-1]
    es_cumsum = np.insert(
        np.cumsum([w.var_states_for_edges.shape[0] for w in wirings]), 0, 0
    )[:-1]
    fces = [
        w.factor_configs_edge_states
        + np.array([[configs_cumsum[ii], es_cumsum[ii]]], dtype=np.int64)
        for ii, w in enumerate(wirings)
    ]
    return EnumWiring(
        var_states_for_edges=vsfe,
        factor_configs_edge_states=np.concatenate(fces, axis=0),
    )

  @staticmethod
  def compile_wiring(
      factor_edges_num_states: np.ndarray,
      variables_for_factors: Sequence[List[Tuple[int, int]]],
      factor_configs: np.ndarray,
      vars_to_starts: Mapping[Tuple[int, int], int],
      num_factors: int,
  ) -> EnumWiring:
    """Compile the wiring of a group of EnumFactors (pgmax/factor/enum.py:191-270).

    Row ``(f * C + c) * V + j`` of ``factor_configs_edge_states`` holds config
    ``f * C + c`` and edge state ``edge_start[f * V + j] + configs[c, j]``,
    as the numba loop at pgmax/factor/enum.py:363-394 writes it.
    """
    first_var_state_by_edges, factor_indices = factor.lookup_var_starts(
        variables_for_factors, vars_to_starts
    )
    edge_cumsum = np.insert(np.cumsum(factor_edges_num_states), 0, 0)
    var_states_for_edges = factor.compile_var_states_for_edges(
        edge_cumsum, first_var_state_by_edges, factor_indices
    )
    num_configs, num_variables = factor_configs.shape
    if factor_edges_num_states.shape != (num_factors * num_variables,):
      raise ValueError(
          "Expected factor_edges_num_states shape is"
          f" {(num_factors * num_variables,)}. Got"
          f" {factor_edges_num_states.shape}."
      )
    fces = np.empty((num_factors * num_configs * num_variables, 2), dtype=np.int64)
    fces[:, 0] = np.repeat(
        np.arange(num_factors * num_configs, dtype=np.int64), num_variables
    )
    starts = edge_cumsum[:-1].reshape(num_factors, num_variables)
    fces[:, 1] = (
        starts[:, None, :] + np.asarray(factor_configs, dtype=np.int64)[None]
    ).reshape(-1)
    return EnumWiring(
        var_states_for_edges=var_states_for_edges,
        factor_configs_edge_states=fces,
    )

  @staticmethod
  def compute_energy(
      edge_states_one_hot_decoding: np.ndarray,
      log_potentials: np.ndarray,
      factor_configs_indices: np.ndarray,
      factor_configs_edge_states: np.ndarray,
      num_val_configs: int,
      num_factors: int,
  ) -> float:
    """Energy contribution of EnumFactors (pgmax/factor/enum.py:272-323)."""
    decoded = np.ones((num_val_configs,), dtype=bool)
    np.logical_and.at(
        decoded,
        factor_configs_indices,
        edge_states_one_hot_decoding[factor_configs_edge_states].astype(bool),
    )
    lp = np.asarray(log_potentials, dtype=np.float64)
    clipped = np.where(
        np.logical_and(np.isinf(lp), ~decoded), -NEG_INF * np.sign(lp), lp
    )
    if decoded.sum() != num_factors:
      return np.inf
    return -np.sum(clipped, where=decoded)

  @staticmethod
  def compute_factor_energy(
      variables: List[Hashable],
      vars_to_map_states: Dict[Hashable, Any],
      factor_configs: np.ndarray,
      log_potentials: np.ndarray,
  ) -> float:
    """Energy of one EnumFactor (pgmax/factor/enum.py:325-360)."""
    lookup = {tuple(c): ii for ii, c in enumerate(factor_configs)}
    decoded = tuple(int(vars_to_map_states[var]) for var in variables)
    if decoded not in lookup:
      warnings.warn(
          f"Invalid decoding for Enum factor {variables} "
          f"with variables set to {decoded}!"
      )
      return float(np.inf)
    return float(-log_potentials[lookup[decoded]])


def pass_enum_fac_to_var_messages(
    vtof_msgs,
    factor_configs_indices,
    factor_configs_edge_states,
    log_potentials,
    num_val_configs: int,
    num_factors: int,
    temperature: float,
    normalize: bool,
):
  """EnumFactor -> variable messages on the GPU (pgmax/factor/enum.py:397-475).

  Inputs are device tensors; the work is one launch of the HIP kernel
  ``pgx_enum_fac_to_var``.
  """
  from pgmax_amd import _lib  # pylint: disable=g-import-not-at-top
  return _lib.enum_fac_to_var(
      vtof_msgs, factor_configs_indices, factor_configs_edge_states,
      log_potentials, num_val_configs, temperature,
  )
