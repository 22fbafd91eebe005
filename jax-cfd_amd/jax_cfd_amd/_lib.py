"""Loader for libjcfd_amd.so and the DLPack hand-off to its C ABI.

The library is the only compute path: if it is missing, or a call fails, an
exception is raised -- there is no CPU fallback.  torch is imported first so
the library binds to the HIP runtime torch already loaded (same soname).
"""
from __future__ import annotations

import contextlib
import ctypes
import os
import threading

import torch  # noqa: F401  (must load before the HIP library; see docstring)
import torch.utils.dlpack

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                        "libjcfd_amd.so")

JCFD_OK = 0
JCFD_ERR_INVALID_ARGUMENT = -1
JCFD_ERR_UNSUPPORTED = -2
JCFD_ERR_DEVICE = -3
JCFD_ERR_OUT_OF_MEMORY = -4
JCFD_ERR_BUDGET = -5

CONVECT_VAN_LEER = 0
CONVECT_VAN_LEER_LIMITERS = 1
CONVECT_LINEAR = 2
CONVECT_UPWIND = 3
PRESSURE_FAST_DIAG = 0
PRESSURE_CG = 1
DIFFUSION_EXPLICIT = 0
DIFFUSION_IMPLICIT_FAST_DIAG = 1
DIFFUSION_IMPLICIT_CG = 2
MAX_RK_STAGES = 8  # JCFD_MAX_RK_STAGES


class BudgetExceededError(RuntimeError):
  """A request would exceed the configured device-memory budget."""


class JcfdDeviceError(RuntimeError):
  """The HIP runtime reported an error."""


class FvmConfig(ctypes.Structure):
  _fields_ = [
      ("n", ctypes.c_int32),
      ("dx", ctypes.c_double), ("dy", ctypes.c_double),
      ("dt", ctypes.c_double),
      ("density", ctypes.c_double),
      ("viscosity", ctypes.c_double),
      ("has_viscosity", ctypes.c_int32),
      ("convect", ctypes.c_int32),
      ("pressure", ctypes.c_int32),
      ("cg_rtol", ctypes.c_double), ("cg_atol", ctypes.c_double),
      ("cg_maxiter", ctypes.c_int32),
      ("has_forcing", ctypes.c_int32),
      ("linear_coefficient", ctypes.c_double),
      ("forcing_x", ctypes.POINTER(ctypes.c_float)),
      ("batch", ctypes.c_int32),
      ("diffusion", ctypes.c_int32),
      ("rk_stages", ctypes.c_int32),
      ("rk_a", ctypes.c_double * 6),
      ("rk_b", ctypes.c_double * 4),
      ("diff_cg_rtol", ctypes.c_double), ("diff_cg_atol", ctypes.c_double),
      ("diff_cg_maxiter", ctypes.c_int32),
  ]


class Ns2dConfig(ctypes.Structure):
  _fields_ = [
      ("n", ctypes.c_int32),
      ("dx", ctypes.c_double), ("dy", ctypes.c_double),
      ("viscosity", ctypes.c_double), ("drag", ctypes.c_double),
      ("smooth", ctypes.c_int32),
      ("dt", ctypes.c_double),
      ("batch", ctypes.c_int32),
      ("n_forcing", ctypes.c_int32),
      ("forcing_index", ctypes.POINTER(ctypes.c_int32)),
      ("forcing_value", ctypes.POINTER(ctypes.c_float)),
      ("rk_stages", ctypes.c_int32),
      ("rk_alpha", ctypes.c_double * (MAX_RK_STAGES + 1)),
      ("rk_beta", ctypes.c_double * MAX_RK_STAGES),
      ("rk_gamma", ctypes.c_double * MAX_RK_STAGES),
  ]


LI_MAX_LAYERS = 16  # JCFD_LI_MAX_LAYERS
LI_COEF_CHANNELS = 128


class LiConfig(ctypes.Structure):
  """jcfd_li_config."""
  _fields_ = [
      ("fvm", FvmConfig),
      ("n_layers", ctypes.c_int32),
      ("hidden_channels", ctypes.c_int32),
      ("n_logits", ctypes.c_int32),
      ("interp_key", ctypes.c_int32 * 8),
  ]


_P = ctypes.c_void_p
_FP = ctypes.POINTER(ctypes.c_float)
_SIGNATURES = {
    "jcfd_li_plan_create": (ctypes.c_int, [ctypes.POINTER(LiConfig), ctypes.POINTER(_P)]),
    "jcfd_li_plan_destroy": (ctypes.c_int, [_P]),
    "jcfd_li_plan_device_bytes": (ctypes.c_int64, [_P]),
    "jcfd_li_set_layer": (ctypes.c_int, [_P, ctypes.c_int32, _FP, _FP]),
    "jcfd_li_set_logits_layer": (ctypes.c_int, [_P, _FP, _FP]),
    "jcfd_li_step": (ctypes.c_int, [_P, _P, _P, _P, _P, ctypes.c_int64, _P]),
    "jcfd_li_tower": (ctypes.c_int, [_P, _P, _P, _P, ctypes.c_int32, _P]),
    "jcfd_li_explicit_terms": (ctypes.c_int, [_P, _P, _P, _P, _P, _P]),
    "jcfd_li_interpolate": (ctypes.c_int, [_P, _P, _P, ctypes.c_int32, _P, _P]),
    "jcfd_li_time_kernels": (ctypes.c_int, [_P, _P, _P, ctypes.c_int32,
                                            ctypes.POINTER(ctypes.c_double), _P]),
    "jcfd_li_kernel_flops": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_double)]),
    "jcfd_ns2d_plan_create": (ctypes.c_int, [ctypes.POINTER(Ns2dConfig), ctypes.POINTER(_P)]),
    "jcfd_ns2d_plan_destroy": (ctypes.c_int, [_P]),
    "jcfd_ns2d_plan_device_bytes": (ctypes.c_int64, [_P]),
    "jcfd_ns2d_plan_set_graphs": (ctypes.c_int, [_P, ctypes.c_int32]),
    "jcfd_fvm_plan_set_graphs": (ctypes.c_int, [_P, ctypes.c_int32]),
    "jcfd_fvm_plan_set_cg_persistent": (ctypes.c_int, [_P, ctypes.c_int32]),
    "jcfd_ns2d_step": (ctypes.c_int, [_P, _P, _P, ctypes.c_int64, _P]),
    "jcfd_ns2d_explicit_terms": (ctypes.c_int, [_P, _P, _P, _P]),
    "jcfd_ns2d_time_kernels": (ctypes.c_int, [_P, _P, ctypes.c_int32,
                                              ctypes.POINTER(ctypes.c_double), _P]),
    "jcfd_ns2d_kernel_bytes": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int64)]),
    "jcfd_ns2d_rfft2": (ctypes.c_int, [_P, _P, _P, _P]),
    "jcfd_ns2d_irfft2": (ctypes.c_int, [_P, _P, _P, _P]),
    "jcfd_last_error": (ctypes.c_char_p, []),
    "jcfd_version": (ctypes.c_char_p, []),
    "jcfd_fvm_plan_create": (ctypes.c_int, [ctypes.POINTER(FvmConfig), ctypes.POINTER(_P)]),
    "jcfd_fvm_plan_destroy": (ctypes.c_int, [_P]),
    "jcfd_fvm_plan_device_bytes": (ctypes.c_int64, [_P]),
    "jcfd_fvm_step": (ctypes.c_int, [_P, _P, _P, _P, _P, ctypes.c_int64, _P]),
    "jcfd_fvm_step_slots": (ctypes.c_int, [_P, _P, _P, _P, ctypes.c_int64, _P]),
    "jcfd_fvm_step_noalloc": (ctypes.c_int, [_P, _P, _P, _P, _P, ctypes.c_int64, _P]),
    "jcfd_fvm_cg_iterations": (ctypes.c_int32, [_P]),
    "jcfd_fvm_diffusion_cg_iterations": (ctypes.c_int32, [_P, ctypes.c_int32]),
    "jcfd_fvm_advect_semilagrangian": (ctypes.c_int, [_P, _P, ctypes.c_double, ctypes.c_double,
                                                      _P, _P, ctypes.c_double, _P, _P]),
    "jcfd_ns2d_step_slots": (ctypes.c_int, [_P, _P, _P, ctypes.c_int64, _P]),
    "jcfd_fvm_explicit_terms": (ctypes.c_int, [_P, _P, _P, _P, _P, _P]),
    "jcfd_fvm_projection": (ctypes.c_int, [_P, _P, _P, _P, _P, _P]),
    "jcfd_fvm_divergence": (ctypes.c_int, [_P, _P, _P, _P, _P]),
    "jcfd_fvm_diffuse": (ctypes.c_int, [_P, _P, _P, _P, _P, _P]),
    "jcfd_fvm_curl": (ctypes.c_int, [_P, _P, _P, _P, _P]),
    "jcfd_fvm_step_vjp": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _P, _P]),
    "jcfd_fvm_curl_vjp": (ctypes.c_int, [_P, _P, _P, _P, _P]),
    "jcfd_fvm_time_kernels": (ctypes.c_int, [_P, _P, _P, ctypes.c_int32,
                                             ctypes.POINTER(ctypes.c_double), _P]),
    "jcfd_fvm_kernel_bytes": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int64)]),
    "jcfd_fvm_spectral_filter": (ctypes.c_int, [_P, _P, _P, _P, _P]),
    "jcfd_ns2d_plan_get_tiles": (ctypes.c_int, [_P, _P]),
    "jcfd_ns2d_plan_set_tiles": (ctypes.c_int, [_P, _P]),
    "jcfd_fvm_plan_get_tiles": (ctypes.c_int, [_P, _P]),
    "jcfd_fvm_plan_set_tiles": (ctypes.c_int, [_P, _P]),
    "jcfd_fvm_group_create": (ctypes.c_int, [ctypes.POINTER(_P)]),
    "jcfd_fvm_group_destroy": (ctypes.c_int, [_P]),
    "jcfd_fvm_group_set_graphs": (ctypes.c_int, [_P, ctypes.c_int32]),
    "jcfd_fvm_group_step": (ctypes.c_int, [_P, _P, ctypes.c_int32, ctypes.c_int64, _P]),
    "jcfd_fvm_group_time_kernels": (ctypes.c_int, [_P, ctypes.c_int32,
                                                   ctypes.POINTER(ctypes.c_double), _P]),
    "jcfd_fvm_group_launch_shape": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int32),
                                                   ctypes.POINTER(ctypes.c_int32),
                                                   ctypes.POINTER(ctypes.c_int64)]),
    "jcfd_ns2d_group_create": (ctypes.c_int, [ctypes.POINTER(_P)]),
    "jcfd_ns2d_group_destroy": (ctypes.c_int, [_P]),
    "jcfd_ns2d_group_set_graphs": (ctypes.c_int, [_P, ctypes.c_int32]),
    "jcfd_ns2d_group_step": (ctypes.c_int, [_P, _P, ctypes.c_int32, ctypes.c_int64, _P]),
    "jcfd_ns2d_group_time_kernels": (ctypes.c_int, [_P, ctypes.c_int32,
                                                    ctypes.POINTER(ctypes.c_double), _P]),
    "jcfd_ns2d_group_launch_shape": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int32),
                                                    ctypes.POINTER(ctypes.c_int32),
                                                    ctypes.POINTER(ctypes.c_int64)]),
}


class FvmTiles(ctypes.Structure):
  """jcfd_fvm_tiles."""
  _fields_ = [("rows_forward", ctypes.c_int32), ("poisson_cols", ctypes.c_int32),
              ("rows_project", ctypes.c_int32), ("radix", ctypes.c_int32)]


class Ns2dTiles(ctypes.Structure):
  """jcfd_ns2d_tiles."""
  _fields_ = [("cols", ctypes.c_int32), ("rows", ctypes.c_int32), ("radix", ctypes.c_int32)]


class FvmGroupMember(ctypes.Structure):
  """jcfd_fvm_group_member."""
  _fields_ = [("plan", _P), ("u", _P), ("v", _P),
              ("slots", ctypes.POINTER(ctypes.c_int32)), ("n_slots", ctypes.c_int32)]


class Ns2dGroupMember(ctypes.Structure):
  """jcfd_ns2d_group_member."""
  _fields_ = [("plan", _P), ("w_hat", _P),
              ("slots", ctypes.POINTER(ctypes.c_int32)), ("n_slots", ctypes.c_int32)]

_lib = None
_lock = threading.Lock()


def exported_symbols():
  return sorted(_SIGNATURES)


def load():
  """Load the library once; raise if it has not been built."""
  global _lib
  if _lib is not None:
    return _lib
  with _lock:
    if _lib is None:
      if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python jax-cfd_amd/build.py`"
            " (there is no non-HIP fallback)")
      lib = ctypes.CDLL(LIB_PATH)
      for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
      _lib = lib
  return _lib


def last_error() -> str:
  return load().jcfd_last_error().decode()


def check(rc: int, what: str) -> None:
  if rc == JCFD_OK:
    return
  msg = f"{what}: {last_error()}"
  if rc == JCFD_ERR_INVALID_ARGUMENT:
    raise ValueError(msg)
  if rc == JCFD_ERR_UNSUPPORTED:
    raise NotImplementedError(msg)
  if rc == JCFD_ERR_OUT_OF_MEMORY:
    raise MemoryError(msg)
  if rc == JCFD_ERR_BUDGET:
    raise BudgetExceededError(msg)
  raise JcfdDeviceError(msg)


# -------------------------------------------------------------- DLPack
_PyCapsule_GetPointer = ctypes.pythonapi.PyCapsule_GetPointer
_PyCapsule_GetPointer.restype = ctypes.c_void_p
_PyCapsule_GetPointer.argtypes = [ctypes.py_object, ctypes.c_char_p]


class _DLHandles:
  """Keeps DLPack capsules alive while the C call uses their DLTensor*."""

  def __init__(self):
    self._caps = []

  def __call__(self, t: torch.Tensor) -> int:
    cap = torch.utils.dlpack.to_dlpack(t)
    self._caps.append(cap)
    # DLManagedTensor starts with its DLTensor, whose layout jcfd_tensor shares.
    return _PyCapsule_GetPointer(cap, b"dltensor")


@contextlib.contextmanager
def dlpack_args():
  h = _DLHandles()
  try:
    yield h
  finally:
    h._caps.clear()


def stream_handle(device: torch.device) -> int:
  return torch.cuda.current_stream(device).cuda_stream


def as_device_tensor(x, dtype: torch.dtype, device=None) -> torch.Tensor:
  """Zero-copy view of a DLPack-capable array, or a copy onto the GPU.

  Device placement and dtype conversion are data movement, not compute.
  """
  if not isinstance(x, torch.Tensor):
    if hasattr(x, "__dlpack__"):
      try:
        x = torch.from_dlpack(x)
      except Exception:  # e.g. numpy arrays of unsupported dtype
        x = torch.as_tensor(x)
    else:
      x = torch.as_tensor(x)
  if device is None:
    device = x.device if x.is_cuda else torch.device("cuda", torch.cuda.current_device())
  if x.device != device or x.dtype != dtype:
    x = x.to(device=device, dtype=dtype)
  return x.contiguous()
