"""Base class for groups of variables (pgmax/vgroup/vgroup.py:29-137).

A variable is the tuple ``(hash, num_states)``. Hashes are ``id(group) % 2**32
* 1e9 + index`` exactly as in the reference, so flat layouts built from the
same groups in the same order are identical.
"""

import dataclasses
import functools
from typing import Any, List, Tuple, Union

import numpy as np

from pgmax_amd.utils import cached_property

MAX_SIZE = 1e9


@functools.total_ordering
@dataclasses.dataclass(frozen=True, eq=False)
class VarGroup:
  """A group of variables (pgmax/vgroup/vgroup.py:29-60)."""

  num_states: Union[int, np.ndarray]
  _hash: int = dataclasses.field(init=False)

  def __post_init__(self):
    this_hash = (id(self) % 2**32) * int(MAX_SIZE)
    assert this_hash < 2**63
    object.__setattr__(self, "_hash", this_hash)

  def __hash__(self):
    return self._hash

  def __eq__(self, other):
    return hash(self) == hash(other)

  def __lt__(self, other):
    return hash(self) < hash(other)

  def __getitem__(self, val: Any) -> Union[Tuple[int, int], List[Tuple[int, int]]]:
    raise NotImplementedError(
        "Please subclass the VarGroup class and override this method"
    )

  @cached_property
  def variable_hashes(self) -> np.ndarray:
    raise NotImplementedError(
        "Please subclass the VarGroup class and override this method"
    )

  @cached_property
  def variables(self) -> List[Tuple[int, int]]:
    """All variables as (hash, num_states) (pgmax/vgroup/vgroup.py:92-106)."""
    assert isinstance(self.variable_hashes, np.ndarray)
    assert isinstance(self.num_states, np.ndarray)
    return list(
        zip(self.variable_hashes.flatten(), self.num_states.flatten())
    )

  def flatten(self, data: Any) -> np.ndarray:
    raise NotImplementedError(
        "Please subclass the VarGroup class and override this method"
    )

  def unflatten(self, flat_data: Any, per_state: bool) -> Any:
    raise NotImplementedError(
        "Please subclass the VarGroup class and override this method"
    )
