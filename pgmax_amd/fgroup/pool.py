"""Groups of pool factors (pgmax/fgroup/pool.py:27-61)."""

import dataclasses
from typing import Optional, Type

import numpy as np

from pgmax_amd.factor import factor
from pgmax_amd.factor import pool
from pgmax_amd.fgroup import fgroup


@dataclasses.dataclass(frozen=True, eq=False)
class PoolFactorGroup(fgroup.FactorGroup):
  """pgmax/fgroup/pool.py:27-61."""

  factor_configs: Optional[np.ndarray] = dataclasses.field(init=False, default=None)
  factor_type: Type[factor.Factor] = dataclasses.field(
      init=False, default=pool.PoolFactor
  )

  def _make_factor(self, f: int) -> pool.PoolFactor:
    return pool.PoolFactor(variables=list(self.given_factor(f)))
