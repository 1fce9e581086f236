"""Numerical constants shared by the host code, the kernels and the oracle.

Values follow pgmax/utils/__init__.py:26-37 of the reference. The HIP kernels
hard-code the same numbers (csrc/pgx_common.h); tests check they agree.
"""

import functools
from typing import Any, Callable

# Lower clip applied to every normalised message (pgmax/utils/__init__.py:26).
MSG_NEG_INF = -1e32

# Clip applied to log-potentials at the start of every run
# (pgmax/utils/__init__.py:32, used at pgmax/infer/bp.py:85-87).
LOG_POTENTIAL_MAX_ABS = 1e6

# "Impossible" log value (pgmax/utils/__init__.py:37).
NEG_INF = -float("inf")


def cached_property(func: Callable[..., Any]) -> property:
  """Property computed once per instance (pgmax/utils/__init__.py:40-49)."""
  return property(functools.lru_cache(None)(func))
