"""Shared inference context (pgmax/infer/inferer.py).

``InfererContext`` flattens the compiled wiring exactly as
pgmax/infer/inferer.py:65-118 does, and owns the native engine handle that
``BP.run`` drives.
"""

import dataclasses
from typing import Any, Callable, Dict, Hashable, Optional, Sequence

import numpy as np

from pgmax_amd import _arrays
from pgmax_amd import factor
from pgmax_amd import vgroup
from pgmax_amd.factor import FAC_TO_VAR_UPDATES
from pgmax_amd.infer import bp_state as bpstate
from pgmax_amd.infer.bp_state import BPArrays
from pgmax_amd.infer.bp_state import BPState
from pgmax_amd.infer.bp_state import Evidence
from pgmax_amd.infer.bp_state import FToVMessages
from pgmax_amd.infer.bp_state import LogPotentials


@dataclasses.dataclass(frozen=True, eq=False)
class Inferer:
  """Inferer functions (pgmax/infer/inferer.py:36-52)."""

  init: Callable[..., BPArrays]
  update: Callable[..., BPArrays]
  to_bp_state: Callable[..., BPState]
  get_beliefs: Callable[..., Dict[Hashable, Any]]
  run: Callable[..., BPArrays]


def device_put(x):
  """Float32 copy of ``x`` in GPU memory when a GPU is present, else on the host."""
  from pgmax_amd import _lib  # pylint: disable=g-import-not-at-top
  return _lib.to_device_f32(x)


@dataclasses.dataclass(frozen=True, eq=False)
class InfererContext:
  """pgmax/infer/inferer.py:55-225."""

  bp_state: BPState

  def __post_init__(self):
    object.__setattr__(self, "wiring", self.bp_state.fg_state.wiring)
    object.__setattr__(
        self, "evidence_to_vars", self.bp_state.fg_state.evidence_to_vars
    )
    vsfe = factor.concatenate_var_states_for_edges(
        [self.wiring[t].var_states_for_edges for t in FAC_TO_VAR_UPDATES]
    )
    object.__setattr__(self, "var_states_for_edge_states", vsfe[..., 0])
    object.__setattr__(self, "edge_indices_for_edge_states", vsfe[..., 1])
    object.__setattr__(self, "factor_indices_for_edge_states", vsfe[..., 2])
    num_variables = (
        int(self.bp_state.fg_state.evidence_to_vars[-1]) + 1
        if self.bp_state.fg_state.evidence_to_vars.shape[0] else 0
    )
    num_edges = int(vsfe[-1, 1]) + 1 if vsfe.shape[0] else 0
    num_factors = int(vsfe[-1, 2]) + 1 if vsfe.shape[0] else 0
    object.__setattr__(self, "num_variables", num_variables)
    object.__setattr__(self, "num_edges", num_edges)
    object.__setattr__(self, "num_factors", num_factors)
    object.__setattr__(self, "inference_arguments", {
        t: self.wiring[t].get_inference_arguments() for t in FAC_TO_VAR_UPDATES
    })
    object.__setattr__(
        self, "factor_type_to_msgs_range",
        self.bp_state.fg_state.factor_type_to_msgs_range,
    )
    object.__setattr__(
        self, "factor_type_to_potentials_range",
        self.bp_state.fg_state.factor_type_to_potentials_range,
    )
    object.__setattr__(self, "_engine", None)

  @property
  def engine(self):
    """The native engine for this wiring, created on first use."""
    if self._engine is None:
      from pgmax_amd import engine  # pylint: disable=g-import-not-at-top
      object.__setattr__(self, "_engine", engine.Engine(self))
    return self._engine

  def update(
      self,
      bp_arrays: Optional[BPArrays] = None,
      log_potentials_updates: Optional[Dict[Any, Any]] = None,
      ftov_msgs_updates: Optional[Dict[Any, Any]] = None,
      evidence_updates: Optional[Dict[Any, Any]] = None,
  ) -> BPArrays:
    """pgmax/infer/inferer.py:120-168."""
    if bp_arrays is not None:
      log_potentials = bp_arrays.log_potentials
      evidence = bp_arrays.evidence
      ftov_msgs = bp_arrays.ftov_msgs
    else:
      log_potentials = device_put(self.bp_state.log_potentials.value)
      ftov_msgs = device_put(self.bp_state.ftov_msgs.value)
      evidence = device_put(self.bp_state.evidence.value)
    fg_state = self.bp_state.fg_state
    if log_potentials_updates is not None:
      log_potentials = bpstate.update_log_potentials(
          log_potentials, log_potentials_updates, fg_state
      )
    if ftov_msgs_updates is not None:
      ftov_msgs = bpstate.update_ftov_msgs(ftov_msgs, ftov_msgs_updates, fg_state)
    if evidence_updates is not None:
      evidence = bpstate.update_evidence(evidence, evidence_updates, fg_state)
    return BPArrays(
        log_potentials=log_potentials, ftov_msgs=ftov_msgs, evidence=evidence
    )

  def init(
      self,
      log_potentials_updates: Optional[Dict[Any, Any]] = None,
      ftov_msgs_updates: Optional[Dict[Any, Any]] = None,
      evidence_updates: Optional[Dict[Any, Any]] = None,
  ) -> BPArrays:
    """pgmax/infer/inferer.py:170-189."""
    return self.update(
        bp_arrays=None,
        log_potentials_updates=log_potentials_updates,
        ftov_msgs_updates=ftov_msgs_updates,
        evidence_updates=evidence_updates,
    )

  def to_bp_state(self, bp_arrays: BPArrays) -> BPState:
    """pgmax/infer/inferer.py:191-209."""
    fg_state = self.bp_state.fg_state
    return BPState(
        log_potentials=LogPotentials(
            fg_state=fg_state, value=_arrays.to_numpy(bp_arrays.log_potentials)
        ),
        ftov_msgs=FToVMessages(
            fg_state=fg_state, value=_arrays.to_numpy(bp_arrays.ftov_msgs)
        ),
        evidence=Evidence(
            fg_state=fg_state, value=_arrays.to_numpy(bp_arrays.evidence)
        ),
    )

  def get_flat_beliefs(self, bp_arrays: BPArrays):
    """evidence + sum of incoming ftov per variable state (pgmax/infer/inferer.py:218-222).

    Computed on the GPU by the same CSR segment-sum kernel the BP loop uses.
    """
    return self.engine.beliefs(bp_arrays.evidence, bp_arrays.ftov_msgs)

  def get_beliefs_map_marginals(self, bp_arrays: BPArrays):
    """Beliefs, MAP states and marginals per variable group from one device
    call (pgx_beliefs_decode): the results of get_beliefs, decode_map_states
    and get_marginals (pgmax/infer/inferer.py:211-264, pgmax/infer/bp.py:263-288)
    without separate passes over the beliefs."""
    flat, states, marg = self.engine.beliefs_decode(bp_arrays.evidence, bp_arrays.ftov_msgs)
    groups = self.bp_state.fg_state.variable_groups
    beliefs, map_states, marginals = {}, {}, {}
    vs0 = v0 = 0
    for group in groups:
      n_vs, n_v = int(group.num_states.sum()), int(np.size(group.num_states))
      beliefs[group] = group.unflatten(flat[..., vs0:vs0 + n_vs], True)
      map_states[group] = group.unflatten(states[..., v0:v0 + n_v], False)
      marginals[group] = _zero_padding(group.unflatten(marg[..., vs0:vs0 + n_vs], True))
      vs0, v0 = vs0 + n_vs, v0 + n_v
    return beliefs, map_states, marginals

  def get_beliefs(self, bp_arrays: BPArrays) -> Dict[Hashable, Any]:
    """pgmax/infer/inferer.py:211-225."""
    return unflatten_beliefs(
        self.get_flat_beliefs(bp_arrays), self.bp_state.fg_state.variable_groups
    )


def _zero_padding(x):
  """Per-state arrays come padded with -inf for variables with fewer states
  than their group's maximum; a marginal's padding is 0 (the softmax of
  pgmax/infer/bp.py:263-288 gives exp(-inf) there)."""
  if isinstance(x, dict):
    return {k: _zero_padding(v) for k, v in x.items()}
  if _arrays.is_torch(x):
    return x.masked_fill(x == float("-inf"), 0.0)
  return np.where(x == -np.inf, 0.0, x)


def unflatten_beliefs(
    flat_beliefs, variable_groups: Sequence[vgroup.VarGroup]
) -> Dict[Hashable, Any]:
  """pgmax/infer/inferer.py:228-248."""
  beliefs = {}
  start = 0
  for variable_group in variable_groups:
    length = int(variable_group.num_states.sum())
    beliefs[variable_group] = variable_group.unflatten(
        flat_beliefs[..., start:start + length], True
    )
    start += length
  return beliefs


def _argmax_last(x):
  if _arrays.is_torch(x):
    if x.shape[-1] == 0:
      import torch  # pylint: disable=g-import-not-at-top
      return torch.zeros(x.shape[:-1], device=x.device)
    return x.argmax(dim=-1)
  x = np.asarray(x)
  if x.size == 0:
    return np.zeros(x.shape[:-1])
  return np.argmax(x, axis=-1)


def decode_map_states(beliefs: Dict[Hashable, Any]) -> Dict[Hashable, Any]:
  """MAP states: first maximal state per variable (pgmax/infer/inferer.py:251-264)."""
  out = {}
  for key, value in beliefs.items():
    if isinstance(value, dict):
      out[key] = {k: _argmax_last(v) for k, v in value.items()}
    else:
      out[key] = _argmax_last(value)
  return out
