"""Groups of factors (pgmax/fgroup/fgroup.py:28-296)."""

import dataclasses
import functools
import inspect
from typing import Any, FrozenSet, List, Mapping, OrderedDict, Sequence, Tuple, Type

import numpy as np

from pgmax_amd.factor import factor
from pgmax_amd.utils import cached_property


@functools.total_ordering
@dataclasses.dataclass(frozen=True, eq=False)
class FactorGroup:
  """A group of factors of one type (pgmax/fgroup/fgroup.py:28-56)."""

  variables_for_factors: Sequence[List[Tuple[int, int]]]
  factor_configs: np.ndarray = dataclasses.field(init=False)
  log_potentials: np.ndarray = dataclasses.field(
      init=False, default_factory=lambda: np.empty((0,))
  )
  factor_type: Type[Any] = dataclasses.field(init=False)

  def __post_init__(self):
    if not self.variables_for_factors:
      raise ValueError("Cannot create a FactorGroup with no Factor.")

  def __hash__(self):
    return id(self)

  def __eq__(self, other):
    return hash(self) == hash(other)

  def __lt__(self, other):
    return hash(self) < hash(other)

  def __getitem__(self, variables: Sequence[Tuple[int, int]]) -> Any:
    """The factor over a set of variables (pgmax/fgroup/fgroup.py:67-86)."""
    variables = frozenset(variables)
    if variables not in self._variables_to_factors:
      raise ValueError(
          f"The queried factor connected to the set of variables {variables} is"
          " not present in the factor group."
      )
    return self._variables_to_factors[variables]

  @cached_property
  def factor_sizes(self) -> np.ndarray:
    return np.array([len(v) for v in self.variables_for_factors])

  @cached_property
  def factor_edges_num_states(self) -> np.ndarray:
    """Number of states of every (factor, variable) edge (pgmax/fgroup/fgroup.py:102-122)."""
    return np.array(
        [variable[1] for vs in self.variables_for_factors for variable in vs],
        dtype=np.int64,
    ).reshape(-1)

  @cached_property
  def _variables_to_factors(self) -> Mapping[FrozenSet[Any], factor.Factor]:
    return self._get_variables_to_factors()

  @cached_property
  def factor_group_log_potentials(self) -> np.ndarray:
    return self.log_potentials.flatten()

  @cached_property
  def factors(self) -> Tuple[factor.Factor, ...]:
    return tuple(self._variables_to_factors.values())

  @cached_property
  def num_factors(self) -> int:
    return len(self.variables_for_factors)

  def _get_variables_to_factors(self) -> OrderedDict[FrozenSet[Any], Any]:
    raise NotImplementedError(
        "Please subclass the FactorGroup class and override this method"
    )

  def flatten(self, data):
    raise NotImplementedError(
        "Please subclass the FactorGroup class and override this method"
    )

  def unflatten(self, flat_data):
    raise NotImplementedError(
        "Please subclass the FactorGroup class and override this method"
    )

  def compile_wiring(self, vars_to_starts: Mapping[Tuple[int, int], int]) -> Any:
    """Compile this group's wiring (pgmax/fgroup/fgroup.py:195-220).

    The factor type's ``compile_wiring`` arguments are looked up on the group
    by name, as in the reference.
    """
    names = inspect.getfullargspec(self.factor_type.compile_wiring).args
    names.remove("vars_to_starts")
    kwargs = {key: getattr(self, key) for key in names}
    return self.factor_type.compile_wiring(vars_to_starts=vars_to_starts, **kwargs)


@dataclasses.dataclass(frozen=True, eq=False)
class SingleFactorGroup(FactorGroup):
  """Group wrapping one factor (pgmax/fgroup/fgroup.py:223-296)."""

  single_factor: factor.Factor

  def __post_init__(self):
    super().__post_init__()
    if len(self.variables_for_factors) != 1:
      raise ValueError(
          "SingleFactorGroup should only contain one factor. Got"
          f" {len(self.variables_for_factors)}"
      )
    object.__setattr__(self, "factor_type", type(self.single_factor))
    names = inspect.getfullargspec(self.factor_type.compile_wiring).args
    names.remove("vars_to_starts")
    for key in names:
      if not hasattr(self, key):
        object.__setattr__(self, key, getattr(self.single_factor, key))
    lp = getattr(self.single_factor, "log_potentials")
    if lp.shape[0] > 0:
      lp = lp[None]
    object.__setattr__(self, "log_potentials", lp)

  def _get_variables_to_factors(self):
    return collections_ordered(
        [(frozenset(self.variables_for_factors[0]), self.single_factor)]
    )

  def flatten(self, data):
    raise NotImplementedError(
        "SingleFactorGroup does not support vectorized factor operations."
    )

  def unflatten(self, flat_data):
    raise NotImplementedError(
        "SingleFactorGroup does not support vectorized factor operations."
    )


def collections_ordered(items):
  import collections  # pylint: disable=g-import-not-at-top
  return collections.OrderedDict(items)
