"""Named variables (the interface of pgmax/vgroup/vdict.py:26-189).

A ``VarDict`` numbers its names in the order given: name ``k`` is variable
``(group hash + k, num_states[k])`` and owns flat states
``[first_state[k], first_state[k] + num_states[k])`` of the group's
per-state layout. Data moves between a ``{name: array}`` mapping and that
flat layout; any leading batch axes of a flat array are kept.
"""

import dataclasses
from typing import Any, Dict, Hashable, Mapping, Tuple

import numpy as np

from pgmax_amd import _arrays
from pgmax_amd.utils import cached_property
from pgmax_amd.vgroup import vgroup


def _state_counts(num_states: Any, num_names: int) -> np.ndarray:
  """Per-name state counts from a scalar or an int array of one per name."""
  if np.isscalar(num_states):
    return np.full((num_names,), int(num_states), dtype=np.int64)
  if isinstance(num_states, np.ndarray) and np.issubdtype(num_states.dtype, np.integer):
    if num_states.shape != (num_names,):
      raise ValueError(
          f"Expected num_states shape ({num_names},). Got {num_states.shape}."
      )
    return num_states
  raise ValueError("num_states should be an integer or a NumPy array of dtype int")


@dataclasses.dataclass(frozen=True, eq=False)
class VarDict(vgroup.VarGroup):
  """Variables addressed by name (pgmax/vgroup/vdict.py:26-59)."""

  variable_names: Tuple[Any, ...]
  _hash: int = dataclasses.field(init=False)

  def __post_init__(self):
    super().__post_init__()
    object.__setattr__(
        self, "num_states", _state_counts(self.num_states, len(self.variable_names))
    )
    object.__setattr__(
        self, "_position", {name: k for k, name in enumerate(self.variable_names)}
    )

  @cached_property
  def variable_hashes(self) -> np.ndarray:
    return self._hash + np.arange(len(self.variable_names), dtype=np.int64)

  @cached_property
  def _first_state(self) -> np.ndarray:
    counts = np.asarray(self.num_states, dtype=np.int64)
    return np.concatenate([[0], np.cumsum(counts)])

  def __getitem__(self, var_name: Any) -> Tuple[int, int]:
    """The variable ``(hash, num_states)`` of a name (pgmax/vgroup/vdict.py:71-85)."""
    k = self._position.get(var_name)
    if k is None:
      raise ValueError(f"Variable {var_name} is not in VarDict")
    return (self._hash + k, int(self.num_states[k]))

  def flatten(self, data: Mapping[Any, Any]):
    """``{name: array}`` to the flat layout (pgmax/vgroup/vdict.py:87-126).

    Every name's array holds one value per state (per-state data) or one
    value (per-variable data).
    """
    for var_name, value in data.items():
      k = self._position.get(var_name)
      if k is None:
        raise ValueError(f"data is referring to a non-existent variable {var_name}.")
      shape = tuple(value.shape)
      if shape not in ((int(self.num_states[k]),), (1,)):
        raise ValueError(
            f"Variable {var_name} expects a data array of shape"
            f" {(int(self.num_states[k]),)} or (1,). Got {shape}."
        )
    pieces = [_arrays.reshape(data[name], (-1,)) for name in self.variable_names]
    return _arrays.concatenate(pieces, axis=0)

  def unflatten(self, flat_data, per_state: bool) -> Dict[Hashable, Any]:
    """The flat layout to ``{name: array}`` (pgmax/vgroup/vdict.py:128-189).

    ``flat_data``'s last axis is the flat layout (per state, or one value per
    variable); leading axes are batch axes and stay on every value.
    """
    num_names = len(self.variable_names)
    size = int(self._first_state[-1]) if per_state else num_names
    ndim = flat_data.ndim
    if ndim == 0 or (ndim > 1 and flat_data.shape[-1] != size):
      raise ValueError(f"Can only unflatten 1D array. Got a {ndim}D array.")
    if flat_data.shape[-1] != size:
      what = (f"num_variable_states(={size})" if per_state
              else f"num_variables(={size})")
      raise ValueError(f"flat_data should be shape ({what},). Got {tuple(flat_data.shape)}")
    if per_state:
      bounds = self._first_state.tolist()
      return {
          name: flat_data[..., bounds[k]:bounds[k + 1]]
          for k, name in enumerate(self.variable_names)
      }
    return {name: flat_data[..., k] for k, name in enumerate(self.variable_names)}
