"""ctypes binding of libpgx.so (the C ABI declared in include/pgx.h).

There is no fallback: if the library is missing or no GPU is visible, the
calls that need it raise.
"""

import atexit
import ctypes
import os
import sys
import threading
import weakref

import numpy as np

from pgmax_amd import _arrays

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpgx.so")

PGX_OK, PGX_EINVAL, PGX_EUNSUP, PGX_EHIP, PGX_ENOMEM = 0, -1, -2, -3, -4
PGX_ENUM, PGX_OR, PGX_AND, PGX_POOL = 0, 1, 2, 3
PGX_MPNN_NUM_PARAMS = 37697

_i64 = ctypes.c_int64
_i32 = ctypes.c_int32
_p = ctypes.c_void_p
_pi32 = ctypes.POINTER(ctypes.c_int32)


class LogicalWiring(ctypes.Structure):
  _fields_ = [
      ("num_parents", _i64),
      ("num_factors", _i64),
      ("parents_factor_indices", _pi32),
      ("parents_msg_indices", _pi32),
      ("children_edge_states", _pi32),
      ("edge_states_offset", _i32),
  ]


class Wiring(ctypes.Structure):
  _fields_ = [
      ("num_edge_states", _i64),
      ("num_var_states", _i64),
      ("num_log_potentials", _i64),
      ("var_states_for_edge_states", _pi32),
      ("edge_indices_for_edge_states", _pi32),
      ("factor_indices_for_edge_states", _pi32),
      ("evidence_to_vars", _pi32),
      ("msgs_range", (_i64 * 2) * 4),
      ("potentials_range", (_i64 * 2) * 4),
      ("enum_num_rows", _i64),
      ("enum_factor_configs_indices", _pi32),
      ("enum_factor_configs_edge_states", _pi32),
      ("logical", LogicalWiring * 4),
  ]


SYMBOLS = {
    "pgx_graph_create": (ctypes.c_int, [ctypes.POINTER(Wiring), ctypes.POINTER(_p)]),
    "pgx_graph_destroy": (ctypes.c_int, [_p]),
    "pgx_wiring_check": (ctypes.c_int, [ctypes.POINTER(Wiring), ctypes.POINTER(_i64)]),
    "pgx_graph_info": (ctypes.c_int, [_p, ctypes.POINTER(_i64)]),
    "pgx_run": (ctypes.c_int, [_p, _p, _i64, _p, _i64, _p, _i64, _p, _p, _i32, _i32,
                               ctypes.c_float, ctypes.c_float, _p]),
    "pgx_run_noalloc": (ctypes.c_int, [_p, _p, _i64, _p, _i64, _p, _i64, _p, _p, _i32, _i32,
                                       ctypes.c_float, ctypes.c_float, _p]),
    "pgx_run_converge": (ctypes.c_int, [_p, _p, _i64, _p, _i64, _p, _i64, _p, _p, _p, _i32,
                                        _i32, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                        _p]),
    "pgx_graph_tasks": (ctypes.c_int, [_p, _p]),
    "pgx_autotune": (ctypes.c_int, [_p, _p, _i64, _p, _i64, _p, _i64, _i32, ctypes.c_float,
                                    ctypes.c_float, _i32, _i32, _p, _p]),
    "pgx_get_beliefs": (ctypes.c_int, [_p, _p, _i64, _p, _i64, _p, _i32, _p]),
    "pgx_decode_map": (ctypes.c_int, [_p, _p, _p, _i32, _p]),
    "pgx_beliefs_decode": (ctypes.c_int, [_p, _p, _i64, _p, _i64, _p, _p, _p, _i32, _p]),
    "pgx_fac_to_var": (ctypes.c_int, [_p, _i32, _p, _p, _p, ctypes.c_float, _i32, _p]),
    "pgx_enum_rows_fac_to_var": (ctypes.c_int, [_p, _p, _p, _p, _i64, _i64, _p, _p, _p, _p,
                                                 ctypes.c_float, _p]),
    "pgx_set_option": (_i64, [ctypes.c_char_p, _i64]),
    "pgx_profile_read": (ctypes.c_int, [ctypes.POINTER(ctypes.c_double),
                                        ctypes.POINTER(_i64)]),
    "pgx_mpnn_create": (ctypes.c_int, [_p, ctypes.POINTER(_p)]),
    "pgx_mpnn_destroy": (ctypes.c_int, [_p]),
    "pgx_mpnn_predict": (ctypes.c_int, [_p, _p, _p, _i64, _p, _i64, _p, _i32, _p]),
    "pgx_mpnn_grad": (ctypes.c_int, [_p, _p, _p, _i64, _p, _i64, _p, _i64, _p, _p, _i32, _p]),
    "pgx_mpnn_adam": (ctypes.c_int, [_p, _p, _p, _p, _i64, _i32, ctypes.c_float, ctypes.c_float,
                                     ctypes.c_float, ctypes.c_float, _p]),
    "pgx_check_numerics": (ctypes.c_int, [_i32, ctypes.c_uint32, ctypes.c_uint32, _p, _p]),
    "pgx_vjp_create": (ctypes.c_int, [_p, _p, _p, _p, _p, _i64, ctypes.POINTER(_p)]),
    "pgx_vjp_destroy": (ctypes.c_int, [_p]),
    "pgx_vjp_set_logical": (ctypes.c_int, [_p, _i32, _p, _p, _i64, _p, _i64, _i32]),
    "pgx_vjp_set_pool": (ctypes.c_int, [_p, _p, _p, _i64, _p, _i64]),
    "pgx_run_vjp": (ctypes.c_int, [_p, _p, _i64, _p, _i64, _p, _i64, _p, _p, _p, _p, _i32, _i32,
                                   ctypes.c_float, ctypes.c_float, _p]),
    "pgx_profile_kernels": (ctypes.c_int, [_p, _p]),
    "pgx_last_error": (ctypes.c_char_p, []),
}

_lib = None
_lock = threading.Lock()


def load():
  """Loads libpgx.so (raises if it has not been built)."""
  global _lib
  with _lock:
    if _lib is None:
      if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m pgmax_amd.csrc.build`"
        )
      lib = ctypes.CDLL(LIB_PATH)
      for name, (res, args) in SYMBOLS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
      _lib = lib
  return _lib


class PgxError(RuntimeError):
  pass


class NativeHandle:
  """An owned native object (pgx_graph, pgx_mpnn) released exactly once.

  ``close()`` releases it. Handles still open at interpreter exit are
  released by an atexit hook, which runs while the HIP runtime is still up;
  ``__del__`` during interpreter finalisation releases nothing, since by then
  the runtime (or a profiler's tool library) may already be torn down and a
  hipFree there can crash the process.
  """

  _destroy_symbol = ""

  def _adopt(self, handle) -> None:
    self.handle = handle
    _LIVE.add(self)

  def close(self) -> None:
    h = getattr(self, "handle", None)
    self.handle = None
    if h:
      getattr(load(), self._destroy_symbol)(h)
    _LIVE.discard(self)

  def __enter__(self):
    return self

  def __exit__(self, *exc):
    self.close()

  def __del__(self, _finalizing=sys.is_finalizing):
    # _finalizing is bound at class creation: module globals (sys included)
    # may already be gone when this runs at interpreter exit
    if _finalizing():
      return
    try:
      self.close()
    except Exception:  # pylint: disable=broad-except
      pass


_LIVE: "weakref.WeakSet[NativeHandle]" = weakref.WeakSet()


@atexit.register
def _close_live_handles() -> None:
  for h in list(_LIVE):
    try:
      h.close()
    except Exception:  # pylint: disable=broad-except
      pass


def check(rc: int, what: str) -> None:
  if rc == PGX_OK:
    return
  msg = load().pgx_last_error().decode(errors="replace")
  if rc in (PGX_EINVAL, PGX_EUNSUP):
    raise ValueError(f"{what}: {msg}")
  raise PgxError(f"{what} failed ({rc}): {msg}")


def set_option(key: str, value: int) -> int:
  return int(load().pgx_set_option(key.encode(), int(value)))


def profile_read():
  """(summed milliseconds, number of runs) recorded since the last read."""
  ms = ctypes.c_double()
  n = _i64()
  check(load().pgx_profile_read(ctypes.byref(ms), ctypes.byref(n)), "pgx_profile_read")
  return float(ms.value), int(n.value)


KERNEL_NAMES = ["task_loop", "var_phase", "var_combine", "fac_fused", "fac_phase_enum2x2",
                "fac_phase_enum_pair", "fac_phase_enum", "fac_phase_or", "fac_phase_and",
                "fac_phase_pool", "bipartite", "vjp_var_sums", "vjp_pair_fwd", "vjp_pair_bwd",
                "vjp_var_bwd"]


def profile_kernels():
  """{kernel name: (summed ms, launches)} recorded with option "profile" = 2."""
  ms = (ctypes.c_double * len(KERNEL_NAMES))()
  n = (_i64 * len(KERNEL_NAMES))()
  check(load().pgx_profile_kernels(ms, n), "pgx_profile_kernels")
  return {k: (float(ms[i]), int(n[i])) for i, k in enumerate(KERNEL_NAMES) if n[i]}


def require_gpu():
  import torch  # pylint: disable=g-import-not-at-top
  if not torch.cuda.is_available():
    raise RuntimeError(
        "The pgmax_amd BP engine runs on an AMD GPU through HIP; no GPU is visible."
    )
  return torch


def to_device_f32(x):
  """Float32 array in GPU memory (torch tensor) when a GPU is visible, else NumPy."""
  try:
    import torch  # pylint: disable=g-import-not-at-top
  except ImportError:  # pragma: no cover
    return np.asarray(x, dtype=np.float32)
  if torch.cuda.is_available():
    if _arrays.is_torch(x):
      return x.to(device="cuda", dtype=torch.float32)
    return torch.as_tensor(np.asarray(x, dtype=np.float32), device="cuda")
  return np.asarray(_arrays.to_numpy(x), dtype=np.float32)


def current_stream() -> int:
  import torch  # pylint: disable=g-import-not-at-top
  return int(torch.cuda.current_stream().cuda_stream)


def ptr(t) -> int:
  return int(t.data_ptr()) if t is not None and t.numel() > 0 else 0


def enum_fac_to_var(*args, **kwargs):
  from pgmax_amd import engine  # pylint: disable=g-import-not-at-top
  return engine.op_enum_fac_to_var(*args, **kwargs)


def logical_fac_to_var(*args, **kwargs):
  from pgmax_amd import engine  # pylint: disable=g-import-not-at-top
  return engine.op_logical_fac_to_var(*args, **kwargs)


def pool_fac_to_var(*args, **kwargs):
  from pgmax_amd import engine  # pylint: disable=g-import-not-at-top
  return engine.op_pool_fac_to_var(*args, **kwargs)
