"""Inference (mirror of pgmax/infer/__init__.py).

``build_inferer`` supports the "bp" backend; the smooth dual LP inferer of
the reference is not part of this package.
"""

from pgmax_amd.infer.bp import BeliefPropagation
from pgmax_amd.infer.bp import BP
from pgmax_amd.infer.bp import get_marginals
from pgmax_amd.infer.bp_state import BPArrays
from pgmax_amd.infer.bp_state import BPState
from pgmax_amd.infer.bp_state import Evidence
from pgmax_amd.infer.bp_state import FToVMessages
from pgmax_amd.infer.bp_state import LogPotentials
from pgmax_amd.infer.energy import compute_energy
from pgmax_amd.infer.inferer import decode_map_states
from pgmax_amd.infer.inferer import Inferer
from pgmax_amd.infer.inferer import InfererContext


def build_inferer(bp_state: BPState, backend: str) -> Inferer:
  """pgmax/infer/__init__.py:34-41 (the "bp" backend)."""
  if backend == "bp":
    return BP(bp_state)
  raise NotImplementedError(f"Inferer {backend} is not supported.")
