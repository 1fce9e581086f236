"""Energy of a decoding, evaluated on the host (pgmax/infer/energy.py).

Post-hoc evaluation only, not on the BP hot path; NumPy float64.
"""

from typing import Any, Dict, Hashable, Tuple

import numpy as np

from pgmax_amd import _arrays
from pgmax_amd import factor
from pgmax_amd import vgroup
from pgmax_amd.factor import FAC_TO_VAR_UPDATES
from pgmax_amd.infer import bp_state as bpstate
from pgmax_amd.infer.bp_state import BPArrays
from pgmax_amd.infer.bp_state import BPState
from pgmax_amd.infer.bp_state import Evidence


def get_vars_to_map_states(map_states: Dict[Hashable, Any]) -> Dict[Hashable, Any]:
  """pgmax/infer/energy.py:29-49."""
  out = {}
  for variable_group, vg_map_states in map_states.items():
    if isinstance(variable_group, vgroup.NDVarArray) and np.prod(
        variable_group.shape
    ) == 0:
      continue
    if isinstance(vg_map_states, dict):
      vg_map_states = {k: np.asarray(_arrays.to_numpy(v)).reshape(-1)
                       for k, v in vg_map_states.items()}
    else:
      vg_map_states = _arrays.to_numpy(vg_map_states)
    flat = _arrays.to_numpy(variable_group.flatten(vg_map_states))
    out.update(zip(variable_group.variables, list(np.asarray(flat))))
  return out


def compute_energy(
    bp_state: BPState,
    bp_arrays: BPArrays,
    map_states: Dict[Hashable, Any],
    debug_mode=False,
) -> Tuple[float, Any, Any]:
  """Energy of a MAP decoding (pgmax/infer/energy.py:52-148); lower is better."""
  if debug_mode:
    return _compute_energy_debug_mode(bp_state, bp_arrays, map_states)
  fg_state = bp_state.fg_state
  wiring = fg_state.wiring
  vsfe = factor.concatenate_var_states_for_edges(
      [wiring[t].var_states_for_edges for t in FAC_TO_VAR_UPDATES]
  )
  log_potentials = _arrays.to_numpy(bp_arrays.log_potentials).astype(np.float64)
  evidence = _arrays.to_numpy(bp_arrays.evidence).astype(np.float64)
  one_hot = np.zeros_like(evidence)
  starts = fg_state.vars_to_starts
  for variable_group in fg_state.variable_groups:
    decoded = map_states[variable_group]
    if isinstance(decoded, dict):
      for name, state in decoded.items():
        var = variable_group[name]
        one_hot[starts[var] + int(np.asarray(_arrays.to_numpy(state)).reshape(-1)[0])] = 1.0
    else:
      decoded = _arrays.to_numpy(decoded).reshape(-1).astype(np.int64)
      first = starts.lookup(np.asarray(variable_group.variable_hashes).reshape(-1),
                            np.asarray(variable_group.num_states).reshape(-1))
      one_hot[first + decoded] = 1.0
  energy = -np.sum(one_hot * evidence)
  es_one_hot = one_hot[vsfe[..., 0]]
  for factor_type in FAC_TO_VAR_UPDATES:
    m0, m1 = fg_state.factor_type_to_msgs_range[factor_type]
    p0, p1 = fg_state.factor_type_to_potentials_range[factor_type]
    if m0 != m1:
      energy += factor_type.compute_energy(
          edge_states_one_hot_decoding=es_one_hot[m0:m1],
          log_potentials=log_potentials[p0:p1],
          **wiring[factor_type].get_inference_arguments(),
      )
  return energy, None, None


def _compute_energy_debug_mode(bp_state, bp_arrays, map_states):
  """pgmax/infer/energy.py:151-204."""
  energy = 0.0
  vars_energies = {}
  factors_energies = {}
  vars_to_map_states = get_vars_to_map_states(map_states)
  evidence = Evidence(
      bp_state.fg_state, value=_arrays.to_numpy(bp_arrays.evidence).astype(np.float64)
  )
  for variable_group in bp_state.fg_state.variable_groups:
    for var in variable_group.variables:
      var_energy = -float(evidence[var][int(vars_to_map_states[var])])
      vars_energies[var] = var_energy
      energy += var_energy
  for factor_group in bp_state.fg_state.factor_group_to_potentials_starts:
    for this_factor in factor_group.factors:
      factor_energy = factor_group.factor_type.compute_factor_energy(
          variables=this_factor.variables,
          vars_to_map_states=vars_to_map_states,
          factor_configs=factor_group.factor_configs,
          log_potentials=this_factor.log_potentials,
      )
      energy += factor_energy
      factors_energies[frozenset(this_factor.variables)] = factor_energy
  return energy, vars_energies, factors_energies
