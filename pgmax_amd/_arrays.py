"""Small array helpers that work on NumPy arrays and torch tensors alike.

Host-side bookkeeping (flattening evidence, spreading messages, unflattening
beliefs) is written once against these helpers so the same code handles host
NumPy inputs and device-resident torch tensors without copies.
"""

from typing import Sequence

import numpy as np

try:  # torch is the device-memory container; it is optional for host logic.
  import torch
except ImportError:  # pragma: no cover
  torch = None


def is_torch(x) -> bool:
  return torch is not None and isinstance(x, torch.Tensor)


def reshape(x, shape):
  return x.reshape(shape)


def take_last(x, idx: np.ndarray):
  """x[..., idx] for an integer index array."""
  if is_torch(x):
    return x[..., torch.as_tensor(idx, device=x.device, dtype=torch.long)]
  return np.asarray(x)[..., idx]


def put_last(out, idx: np.ndarray, values):
  """out[..., idx] = values; returns out."""
  if is_torch(out):
    out[..., torch.as_tensor(idx, device=out.device, dtype=torch.long)] = values
    return out
  out[..., idx] = np.asarray(values)
  return out


def full_like_shape(ref, shape, value):
  if is_torch(ref):
    return torch.full(shape, value, dtype=ref.dtype, device=ref.device)
  return np.full(shape, value, dtype=np.asarray(ref).dtype)


def concatenate(xs: Sequence, axis: int = -1):
  if any(is_torch(x) for x in xs):
    dev = next(x.device for x in xs if is_torch(x))
    ts = [
        x if is_torch(x) else torch.as_tensor(np.asarray(x), device=dev)
        for x in xs
    ]
    dt = torch.result_type(*ts) if len(ts) > 1 else ts[0].dtype
    return torch.cat([t.to(dt) for t in ts], dim=axis)
  return np.concatenate([np.asarray(x) for x in xs], axis=axis)


def to_numpy(x) -> np.ndarray:
  if is_torch(x):
    return x.detach().cpu().numpy()
  return np.asarray(x)


def clip_inf(x):
  """Clip to [-inf, inf]: a no-op kept to mirror bp_state.py:80-82."""
  return x
