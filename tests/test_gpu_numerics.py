"""Device checks of two numerical building blocks that the kernels use in place
of their textbook forms (pgx_check_numerics), each over every input it can
see: the hardware FP32 -> BF16 conversion of the warm-start kernels, and the
division-free quotient f / (2 + f) of the logarithm that every log-sum-exp
and lse2 in the BP kernels evaluates."""

import struct

import pytest


def _bits(x):
  return struct.unpack("<I", struct.pack("<f", x))[0]


def _mismatches(which, ranges):
  import torch
  from pgmax_amd import _lib
  lib = _lib.load()
  bad = torch.zeros(1, dtype=torch.int64, device="cuda")
  for first, count in ranges:
    _lib.check(lib.pgx_check_numerics(which, first, count, _lib.ptr(bad), _lib.current_stream()),
               "pgx_check_numerics")
  torch.cuda.synchronize()
  return int(bad.item())


@pytest.mark.gpu
def test_gpu_bf16_conversion_is_round_to_nearest_even(gpu):
  """All 2^32 bit patterns (each non-negative one and its negation), NaNs
  aside: the conversion equals integer round-to-nearest-even."""
  chunk = 1 << 27
  assert _mismatches(0, [(c * chunk, chunk) for c in range(16)]) == 0


@pytest.mark.gpu
def test_gpu_log_quotient_equals_the_division(gpu):
  """Every mantissa m the logarithm reduces to, [sqrt(1/2), sqrt(2)) with a
  margin of 16 patterns on each side: f / (2 + f) for f = m - 1 has the bits
  of the IEEE division."""
  lo, hi = _bits(0.70710678118654752) - 16, _bits(1.41421356237309505) + 16
  assert hi - lo > 8_000_000
  assert _mismatches(1, [(lo, hi - lo)]) == 0


@pytest.mark.gpu
def test_gpu_check_numerics_rejects_bad_arguments(gpu):
  import ctypes
  from pgmax_amd import _lib
  lib = _lib.load()
  assert lib.pgx_check_numerics(2, 0, 1, ctypes.c_void_p(8), None) != 0
  assert lib.pgx_check_numerics(0, 0, 1, None, None) != 0
