"""Numerical constants shared by the host code, the kernels and the oracle.

Values follow pgmax/utils/__init__.py:26-37 of the reference. The HIP kernels
hard-code the same numbers (csrc/pgx_common.h); tests check they agree.
"""

import functools

# Lower clip applied to every normalised message (pgmax/utils/__init__.py:26).
MSG_NEG_INF = -1e32

# Clip applied to log-potentials at the start of every run
# (pgmax/utils/__init__.py:32, used at pgmax/infer/bp.py:85-87).
LOG_POTENTIAL_MAX_ABS = 1e6

# "Impossible" log value (pgmax/utils/__init__.py:37).
NEG_INF = -float("inf")


# Property computed once per instance (the role of pgmax/utils/__init__.py:40-49).
# The value lives in the instance's own __dict__, so it is freed with the
# instance (a process-wide cache would keep every factor group ever built
# alive, arrays included).
cached_property = functools.cached_property
