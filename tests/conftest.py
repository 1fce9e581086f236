"""Pytest configuration: registers the ``gpu`` marker and puts the repo on sys.path."""

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
  sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
  sys.path.insert(0, TESTS)


def pytest_configure(config):
  config.addinivalue_line(
      "markers", "gpu: needs an AMD GPU and the built HIP library"
  )


@pytest.fixture(scope="session")
def gpu():
  """Skips unless a GPU is visible; returns the torch device."""
  import torch  # pylint: disable=g-import-not-at-top
  if not torch.cuda.is_available():
    pytest.skip("no GPU visible")
  return torch.device("cuda:0")


_DEFAULT_OPTIONS = {"mode": 0, "block": 0, "tile": 0, "fuse": 1, "vec": 3, "profile": 0, "graph": 0,
                    "pack": 1, "lanes": 1, "xcd": 0, "share": 1}


@pytest.fixture(autouse=True)
def _engine_options():
  """Restores the engine options a test may change (shape, tiles, profiling)."""
  yield
  try:
    from pgmax_amd import _lib  # pylint: disable=g-import-not-at-top
    if os.path.exists(_lib.LIB_PATH):
      for key, value in _DEFAULT_OPTIONS.items():
        _lib.set_option(key, value)
  except Exception:  # pylint: disable=broad-except
    pass
