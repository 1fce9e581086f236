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
