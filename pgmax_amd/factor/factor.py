"""Base classes for factors and wirings (pgmax/factor/factor.py:27-225).

The reference fills its wiring arrays with numba loops
(pgmax/factor/factor.py:197-225); here the same arrays are produced with
vectorised NumPy (np.repeat / cumsum), which gives identical int64 values.
"""

import dataclasses
import itertools
from typing import Any, Dict, Hashable, List, Mapping, Sequence, Tuple

import numpy as np


@dataclasses.dataclass(frozen=True, eq=False)
class Wiring:
  """Wiring shared by all factor types (pgmax/factor/factor.py:27-57).

  Attributes:
    var_states_for_edges: int array (num_edge_states, 3); column 0 is the
      global variable-state index, column 1 the global edge index, column 2
      the global factor index of each edge state.
  """

  var_states_for_edges: np.ndarray

  def __post_init__(self):
    for field in self.__dataclass_fields__:
      if isinstance(getattr(self, field), np.ndarray):
        getattr(self, field).flags.writeable = False

  def get_inference_arguments(self):
    raise NotImplementedError(
        "Please subclass the Wiring class and override this method."
    )


@dataclasses.dataclass(frozen=True, eq=False)
class Factor:
  """A factor (pgmax/factor/factor.py:60-129).

  ``accepts_factor_edges``: the type's ``compile_wiring`` takes a factor
  group's flat edge table as ``variables_for_factors`` (the package's own
  types do); a type without it gets variable lists.
  """

  variables: List[Tuple[int, int]]
  log_potentials: np.ndarray
  accepts_factor_edges = True

  def __post_init__(self):
    if not hasattr(self, "compile_wiring"):
      raise NotImplementedError(
          "Please implement compile_wiring in for your factor"
      )

  @staticmethod
  def concatenate_wirings(wirings: Sequence[Wiring]) -> Wiring:
    raise NotImplementedError(
        "Please subclass the Wiring class and override this method."
    )

  @staticmethod
  def compute_energy(wiring, edge_states_one_hot_decoding, log_potentials):
    raise NotImplementedError(
        "Please subclass the Factor class and override this method."
    )

  @staticmethod
  def compute_factor_energy(
      variables: List[Hashable], vars_to_map_states: Dict[Hashable, Any]
  ):
    raise NotImplementedError(
        "Please subclass the Factor class and override this method"
    )


def concatenate_var_states_for_edges(
    list_var_states_for_edges: Sequence[np.ndarray],
) -> np.ndarray:
  """Concatenate wirings, offsetting edge and factor ids.

  Follows pgmax/factor/factor.py:132-192: empty wirings are dropped, then the
  edge column is offset by the running edge count and the factor column by
  the running factor count.
  """
  if list_var_states_for_edges is None:
    raise ValueError("list_var_states_for_edges cannot be None")
  non_empty = []
  for vsfe in list_var_states_for_edges:
    if vsfe is None:
      raise ValueError("var_states_for_edges cannot be None")
    if vsfe.shape[0] > 0:
      non_empty.append(vsfe)
  if not non_empty:
    return np.empty((0, 3), dtype=np.int64)
  out = np.empty((sum(v.shape[0] for v in non_empty), 3), dtype=np.int64)
  row = edge0 = factor0 = 0
  for vsfe in non_empty:
    n = vsfe.shape[0]
    block = out[row:row + n]
    block[:] = vsfe
    if edge0:
      block[:, 1] += edge0
    if factor0:
      block[:, 2] += factor0
    edge0 += int(vsfe[-1, 1]) + 1
    factor0 += int(vsfe[-1, 2]) + 1
    row += n
  return out


def compile_var_states_for_edges(
    num_edges_states_cumsum: np.ndarray,
    first_var_state_by_edges: np.ndarray,
    factor_indices: np.ndarray,
) -> np.ndarray:
  """Vectorised form of pgmax/factor/factor.py:197-225.

  For edge ``e`` with states ``[c[e], c[e+1])``: column 0 is
  ``first_var_state_by_edges[e] + arange``, column 1 is ``e`` and column 2 is
  ``factor_indices[e]``.
  """
  cumsum = np.asarray(num_edges_states_cumsum, dtype=np.int64)
  num_edges = cumsum.shape[0] - 1
  counts = np.diff(cumsum)
  edge_of_es = np.repeat(np.arange(num_edges, dtype=np.int64), counts)
  within = np.arange(cumsum[-1], dtype=np.int64) - cumsum[:-1][edge_of_es]
  out = np.empty((int(cumsum[-1]), 3), dtype=np.int64)
  out[:, 0] = np.asarray(first_var_state_by_edges, dtype=np.int64)[edge_of_es] + within
  out[:, 1] = edge_of_es
  out[:, 2] = np.asarray(factor_indices, dtype=np.int64)[edge_of_es]
  return out


def count_distinct(idx: np.ndarray) -> int:
  """Number of distinct values; linear when ``idx`` is sorted, as the factor
  indices of a compiled wiring are."""
  step = np.diff(idx)
  if np.all(step >= 0):
    return 1 + int(np.count_nonzero(step)) if idx.size else 0
  return int(np.unique(idx).shape[0])


def lookup_var_starts(
    variables_for_factors: Sequence[Sequence[Tuple[int, int]]],
    vars_to_starts: Mapping[Tuple[int, int], int],
) -> Tuple[np.ndarray, np.ndarray]:
  """First variable state of every (factor, variable) edge, in edge order.

  Returns ``(first_var_state_by_edges, factor_indices)`` as built by the
  loops at pgmax/factor/enum.py:226-233 and pgmax/factor/logical.py:241-248.
  ``variables_for_factors`` is a factor group's edge table (flat ``hashes``,
  ``num_states`` and per-factor ``offsets``) or a sequence of variable
  lists. A ``vars_to_starts`` with a vectorised ``lookup`` (the
  FactorGraph's variable index) resolves every edge in one call; any other
  Mapping is queried per variable.
  """
  if hasattr(variables_for_factors, "offsets"):
    edges = variables_for_factors
    hashes, num_states = edges.hashes, edges.num_states
    sizes = np.diff(edges.offsets)
  else:
    sizes = np.fromiter((len(v) for v in variables_for_factors), dtype=np.int64)
    flat = list(itertools.chain.from_iterable(variables_for_factors))
    if not hasattr(vars_to_starts, "lookup"):
      starts = np.fromiter(
          (vars_to_starts[v] for v in flat), dtype=np.int64, count=len(flat)
      )
      return starts, np.repeat(np.arange(sizes.shape[0], dtype=np.int64), sizes)
    arr = np.array(flat, dtype=np.int64).reshape(-1, 2)
    hashes, num_states = arr[:, 0], arr[:, 1]
  factor_indices = np.repeat(np.arange(sizes.shape[0], dtype=np.int64), sizes)
  if hasattr(vars_to_starts, "lookup"):
    return vars_to_starts.lookup(hashes, num_states), factor_indices
  starts = np.fromiter(
      (vars_to_starts[(h, n)] for h, n in zip(hashes.tolist(), num_states.tolist())),
      dtype=np.int64, count=hashes.shape[0],
  )
  return starts, factor_indices
