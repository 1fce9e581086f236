"""Dictionary of named variables (pgmax/vgroup/vdict.py:26-189)."""

import dataclasses
from typing import Any, Dict, Hashable, Mapping, Tuple

import numpy as np

from pgmax_amd import _arrays
from pgmax_amd.utils import cached_property
from pgmax_amd.vgroup import vgroup


@dataclasses.dataclass(frozen=True, eq=False)
class VarDict(vgroup.VarGroup):
  """A set of named variables (pgmax/vgroup/vdict.py:26-59)."""

  variable_names: Tuple[Any, ...]
  _hash: int = dataclasses.field(init=False)

  def __post_init__(self):
    super().__post_init__()
    if np.isscalar(self.num_states):
      num_states = np.full(
          (len(self.variable_names),), fill_value=self.num_states, dtype=np.int64
      )
      object.__setattr__(self, "num_states", num_states)
    elif isinstance(self.num_states, np.ndarray) and np.issubdtype(
        self.num_states.dtype, int
    ):
      if self.num_states.shape != (len(self.variable_names),):
        raise ValueError(
            f"Expected num_states shape ({len(self.variable_names)},). Got"
            f" {self.num_states.shape}."
        )
    else:
      raise ValueError(
          "num_states should be an integer or a NumPy array of dtype int"
      )
    object.__setattr__(
        self,
        "_name_to_index",
        {name: ii for ii, name in enumerate(self.variable_names)},
    )

  @cached_property
  def variable_hashes(self) -> np.ndarray:
    return self.__hash__() + np.arange(len(self.variable_names))

  def __getitem__(self, var_name: Any) -> Tuple[int, int]:
    """Variable by name (pgmax/vgroup/vdict.py:71-85)."""
    if var_name not in self._name_to_index:
      raise ValueError(f"Variable {var_name} is not in VarDict")
    var_idx = self._name_to_index[var_name]
    return (self.variable_hashes[var_idx], self.num_states[var_idx])

  def flatten(self, data: Mapping[Any, Any]):
    """Mapping name -> data to a flat array (pgmax/vgroup/vdict.py:87-126)."""
    for var_name in data:
      if var_name not in self._name_to_index:
        raise ValueError(
            f"data is referring to a non-existent variable {var_name}."
        )
      var_idx = self._name_to_index[var_name]
      shape = tuple(data[var_name].shape)
      if shape != (self.num_states[var_idx],) and shape != (1,):
        raise ValueError(
            f"Variable {var_name} expects a data array of shape"
            f" {(self.num_states[var_idx],)} or (1,). Got {shape}."
        )
    return _arrays.concatenate(
        [_arrays.reshape(data[name], (-1,)) for name in self.variable_names],
        axis=0,
    )

  def unflatten(self, flat_data, per_state: bool) -> Dict[Hashable, Any]:
    """Flat -> mapping name -> data (pgmax/vgroup/vdict.py:128-189).

    Leading batch axes are kept on every value.
    """
    if flat_data.ndim < 1:
      raise ValueError(
          f"Can only unflatten 1D array. Got a {flat_data.ndim}D array."
      )
    num_variables = len(self.variable_names)
    num_variable_states = self.num_states.sum()
    size = flat_data.shape[-1]
    if per_state:
      if size != num_variable_states:
        raise ValueError(
            "flat_data should be shape "
            f"(num_variable_states(={num_variable_states}),). Got "
            f"{tuple(flat_data.shape)}"
        )
      limits = np.insert(np.cumsum(self.num_states), 0, 0)
      return {
          name: flat_data[..., limits[ii]:limits[ii + 1]]
          for ii, name in enumerate(self.variable_names)
      }
    if size != num_variables:
      raise ValueError(
          "flat_data should be shape "
          f"(num_variables(={num_variables}),). Got {tuple(flat_data.shape)}"
      )
    return {
        name: flat_data[..., ii] for ii, name in enumerate(self.variable_names)
    }
