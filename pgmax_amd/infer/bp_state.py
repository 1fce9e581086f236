"""Containers for belief-propagation state (pgmax/infer/bp_state.py).

``BPArrays`` holds the three flat float32 arrays the engine works on. On a
GPU machine ``BP.init`` places them in device memory as torch tensors; host
NumPy arrays are accepted everywhere and moved to the device by ``run``.
A leading batch axis on any of the arrays plays the role of the reference's
``jax.vmap`` over ``BPArrays``.
"""

import dataclasses
from typing import Any, Dict, Optional, Tuple, Union, cast

import numpy as np

from pgmax_amd import _arrays
from pgmax_amd import fgraph
from pgmax_amd import fgroup


@dataclasses.dataclass(frozen=True, eq=False)
class BPArrays:
  """Flat log-potentials, ftov messages and evidence (pgmax/infer/bp_state.py:30-55)."""

  log_potentials: Any
  ftov_msgs: Any
  evidence: Any

  def __post_init__(self):
    for field in self.__dataclass_fields__:
      value = getattr(self, field)
      if isinstance(value, np.ndarray):
        value.flags.writeable = False


def _copy(x):
  if _arrays.is_torch(x):
    return x.clone()
  return np.array(x, copy=True)


def _as_like(data, ref):
  """``data`` as an array of ``ref``'s kind, dtype and device."""
  if _arrays.is_torch(ref):
    import torch  # pylint: disable=g-import-not-at-top
    if _arrays.is_torch(data):
      return data.to(device=ref.device, dtype=ref.dtype)
    return torch.as_tensor(np.asarray(data), dtype=ref.dtype, device=ref.device)
  return np.asarray(_arrays.to_numpy(data), dtype=np.asarray(ref).dtype)


def _set_slice(arr, start: int, flat):
  """Copy of ``arr`` with ``arr[..., start:start+n] = flat``.

  Extra leading axes of ``flat`` broadcast ``arr`` to a batch.
  """
  flat = _as_like(flat, arr)
  lead = tuple(flat.shape[:-1])
  if len(lead) > arr.ndim - 1:
    target = lead + tuple(arr.shape[-1:])
    if _arrays.is_torch(arr):
      out = arr.expand(target).clone()
    else:
      out = np.array(np.broadcast_to(arr, target), copy=True)
  else:
    out = _copy(arr)
  out[..., start:start + flat.shape[-1]] = flat
  return out


def update_log_potentials(log_potentials, updates: Dict[Any, Any], fg_state):
  """New log-potentials with factor-group updates (pgmax/infer/bp_state.py:58-101).

  Raises:
    ValueError: wrong shape, or a factor group not in the graph.
  """
  for factor_group, data in updates.items():
    if factor_group in fg_state.factor_group_to_potentials_starts:
      flat_data = factor_group.flatten(data)
      if tuple(flat_data.shape) != factor_group.factor_group_log_potentials.shape:
        raise ValueError(
            "Expected log potentials shape"
            f" {factor_group.factor_group_log_potentials.shape} for"
            f" factor group. Got incompatible data shape {tuple(data.shape)}."
        )
      start = fg_state.factor_group_to_potentials_starts[factor_group]
      log_potentials = _set_slice(log_potentials, start, flat_data)
    else:
      raise ValueError("Invalid FactorGroup for log potentials updates.")
  return log_potentials


@dataclasses.dataclass(frozen=True, eq=False)
class LogPotentials:
  """pgmax/infer/bp_state.py:104-169."""

  fg_state: fgraph.FactorGraphState
  value: Optional[np.ndarray] = None

  def __post_init__(self):
    if self.value is None:
      object.__setattr__(self, "value", self.fg_state.log_potentials)
    elif tuple(self.value.shape) != self.fg_state.log_potentials.shape:
      raise ValueError(
          "Expected log potentials shape"
          f" {self.fg_state.log_potentials.shape}. Got {tuple(self.value.shape)}."
      )

  def __getitem__(self, factor_group: fgroup.FactorGroup):
    if factor_group not in self.fg_state.factor_group_to_potentials_starts:
      raise ValueError("Invalid FactorGroup queried to access log potentials.")
    start = self.fg_state.factor_group_to_potentials_starts[factor_group]
    n = factor_group.factor_group_log_potentials.shape[0]
    return self.value[start:start + n]

  def __setitem__(self, factor_group: fgroup.FactorGroup, data):
    value = update_log_potentials(
        np.asarray(self.value), {factor_group: _arrays.to_numpy(data)}, self.fg_state
    )
    object.__setattr__(self, "value", np.asarray(value))


def _var_states_for_edge_states(fg_state) -> np.ndarray:
  cached = getattr(fg_state, "_pgx_vsfes", None)
  if cached is None:
    cached = np.concatenate(
        [w.var_states_for_edges[..., 0] for w in fg_state.wiring.values()]
    )
    object.__setattr__(fg_state, "_pgx_vsfes", cached)
  return cached


def update_ftov_msgs(ftov_msgs, updates: Dict[Any, Any], fg_state):
  """New ftov messages with per-type or per-variable updates (pgmax/infer/bp_state.py:172-234).

  A per-variable update spreads ``data / n_edges`` over every edge of the
  variable (pgmax/infer/bp_state.py:208-228).
  """
  for name, data in updates.items():
    if name in fg_state.factor_type_to_msgs_range:
      start, end = fg_state.factor_type_to_msgs_range[name]
      if tuple(data.shape)[-1:] != (end - start,):
        raise ValueError(
            f"Expected ftov_msgs shape {(end - start,)}"
            f" for factor type {name}. Got incompatible shape {tuple(data.shape)}."
        )
      ftov_msgs = _set_slice(ftov_msgs, start, data)
    elif name in fg_state.vars_to_starts:
      num_var_states = name[1]
      if tuple(data.shape)[-1:] != (num_var_states,):
        raise ValueError(
            f"Expected ftov_msgs shape {(num_var_states,)} for variable {name}."
            f" Got incompatible shape {tuple(data.shape)}."
        )
      vsfes = _var_states_for_edge_states(fg_state)
      starts = np.nonzero(vsfes == fg_state.vars_to_starts[name])[0]
      for start in starts:
        ftov_msgs = _set_slice(ftov_msgs, int(start), data / starts.shape[0])
    else:
      raise ValueError("Provided variable or factor type is not in the FactorGraph")
  return ftov_msgs


@dataclasses.dataclass(frozen=True, eq=False)
class FToVMessages:
  """pgmax/infer/bp_state.py:237-289."""

  fg_state: fgraph.FactorGraphState
  value: Optional[np.ndarray] = None

  def __post_init__(self):
    if self.value is None:
      object.__setattr__(
          self, "value", np.zeros(self.fg_state.total_factor_num_states)
      )
    elif tuple(self.value.shape) != (self.fg_state.total_factor_num_states,):
      raise ValueError(
          "Expected messages shape"
          f" {(self.fg_state.total_factor_num_states,)}. Got"
          f" {tuple(self.value.shape)}."
      )

  def __setitem__(self, variable: Tuple[int, int], data) -> None:
    value = update_ftov_msgs(
        np.asarray(self.value), {variable: _arrays.to_numpy(data)}, self.fg_state
    )
    object.__setattr__(self, "value", np.asarray(value))


def update_evidence(evidence, updates: Dict[Any, Any], fg_state):
  """New evidence with per-group or per-variable updates (pgmax/infer/bp_state.py:292-342)."""
  for name, data in updates.items():
    if name in fg_state.variable_groups:
      flat_data = name.flatten(data)
      start = fg_state.vars_to_starts[name.variables[0]]
      evidence = _set_slice(evidence, int(start), flat_data)
    elif name in fg_state.vars_to_starts:
      start = fg_state.vars_to_starts[name]
      if tuple(data.shape)[-1:] != (name[1],):
        raise ValueError(
            f"Expected evidence shape {(name[1],)} for variable {name}."
            f" Got incompatible shape {tuple(data.shape)}."
        )
      evidence = _set_slice(evidence, int(start), data)
    else:
      raise ValueError(
          "Got evidence for a variable or a VarGroup not in the FactorGraph!"
      )
  return evidence


@dataclasses.dataclass(frozen=True, eq=False)
class Evidence:
  """pgmax/infer/bp_state.py:345-405."""

  fg_state: fgraph.FactorGraphState
  value: Optional[np.ndarray] = None

  def __post_init__(self):
    if self.value is None:
      object.__setattr__(self, "value", np.zeros((self.fg_state.num_var_states,)))
    elif tuple(self.value.shape) != (self.fg_state.num_var_states,):
      raise ValueError(
          f"Expected evidence shape {(self.fg_state.num_var_states,)}. "
          f"Got {tuple(self.value.shape)}."
      )

  def __getitem__(self, variable: Tuple[int, int]):
    start = self.fg_state.vars_to_starts[variable]
    return self.value[start:start + variable[1]]

  def __setitem__(self, name: Any, data) -> None:
    if isinstance(data, dict):
      data = {k: _arrays.to_numpy(v) for k, v in data.items()}
    else:
      data = _arrays.to_numpy(data)
    value = update_evidence(np.asarray(self.value), {name: data}, self.fg_state)
    object.__setattr__(self, "value", np.asarray(value))


@dataclasses.dataclass(frozen=True, eq=False)
class BPState:
  """pgmax/infer/bp_state.py:408-436."""

  log_potentials: LogPotentials
  ftov_msgs: FToVMessages
  evidence: Evidence

  def __post_init__(self):
    if (self.log_potentials.fg_state != self.ftov_msgs.fg_state) or (
        self.ftov_msgs.fg_state != self.evidence.fg_state
    ):
      raise ValueError(
          "log_potentials, ftov_msgs and evidence should be derived from the"
          " same fg_state."
      )

  @property
  def fg_state(self) -> fgraph.FactorGraphState:
    return self.log_potentials.fg_state
