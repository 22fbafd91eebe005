"""Spectral plans: the Python side of the jcfd_ns2d_* C ABI."""
from __future__ import annotations

import ctypes
import dataclasses
from typing import Optional, Tuple

import numpy as np
import torch

from jax_cfd_amd import _lib


@dataclasses.dataclass(frozen=True)
class Ns2dSpec:
  n: int
  dx: float
  dy: float
  viscosity: float
  drag: float
  smooth: bool
  dt: float
  # sparse constant forcing curl: ((p, c, re, im), ...)
  forcing: Tuple[Tuple[int, int, float, float], ...] = ()
  # low-storage CN-RK tableau (alphas, betas, gammas); None: crank_nicolson_rk4
  tableau: Optional[Tuple[Tuple[float, ...], Tuple[float, ...], Tuple[float, ...]]] = None

  @property
  def stages(self) -> int:
    return 5 if self.tableau is None else len(self.tableau[1])


class Ns2dPlan:
  """Owns one jcfd_ns2d_plan for `batch` same-size instances on one GPU."""

  def __init__(self, spec: Ns2dSpec, batch: int = 1,
               device: Optional[torch.device] = None):
    lib = _lib.load()
    self.spec = spec
    self.batch = int(batch)
    self.device = device if device is not None else torch.device(
        'cuda', torch.cuda.current_device())
    cfg = _lib.Ns2dConfig()
    cfg.n = spec.n
    cfg.dx, cfg.dy = spec.dx, spec.dy
    cfg.viscosity, cfg.drag = spec.viscosity, spec.drag
    cfg.smooth = int(bool(spec.smooth))
    cfg.dt = spec.dt
    cfg.batch = self.batch
    cfg.n_forcing = len(spec.forcing)
    self._fi = np.ascontiguousarray(
        np.array([(p, c) for p, c, _, _ in spec.forcing] or [(0, 0)], np.int32))
    self._fv = np.ascontiguousarray(
        np.array([(re, im) for _, _, re, im in spec.forcing] or [(0, 0)], np.float32))
    cfg.forcing_index = self._fi.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
    cfg.forcing_value = self._fv.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    if spec.tableau is not None:
      alphas, betas, gammas = spec.tableau
      cfg.rk_stages = len(betas)
      if not (len(alphas) == len(betas) + 1 == len(gammas) + 1) or \
          len(betas) > _lib.MAX_RK_STAGES:
        raise ValueError('a tableau needs k + 1 alphas, k betas and k gammas, '
                         f'k <= {_lib.MAX_RK_STAGES}')
      for i, a in enumerate(alphas):
        cfg.rk_alpha[i] = float(a)
      for i, (b, g) in enumerate(zip(betas, gammas)):
        cfg.rk_beta[i], cfg.rk_gamma[i] = float(b), float(g)
    handle = ctypes.c_void_p()
    with torch.cuda.device(self.device):
      _lib.check(lib.jcfd_ns2d_plan_create(ctypes.byref(cfg), ctypes.byref(handle)),
                 'jcfd_ns2d_plan_create')
    self._handle = handle
    self._lib = lib

  @property
  def shape(self):
    n = self.spec.n
    return (n, n // 2 + 1) if self.batch == 1 else (self.batch, n, n // 2 + 1)

  @property
  def device_bytes(self) -> int:
    return int(self._lib.jcfd_ns2d_plan_device_bytes(self._handle))

  def set_graphs(self, enabled: bool):
    _lib.check(self._lib.jcfd_ns2d_plan_set_graphs(self._handle, int(bool(enabled))),
               'jcfd_ns2d_plan_set_graphs')

  @property
  def tiles(self) -> dict:
    """Tile shapes and radix schedule of the step kernels."""
    t = _lib.Ns2dTiles()
    _lib.check(self._lib.jcfd_ns2d_plan_get_tiles(self._handle, ctypes.byref(t)),
               'jcfd_ns2d_plan_get_tiles')
    return {k: int(getattr(t, k)) for k, _ in _lib.Ns2dTiles._fields_}

  def set_tiles(self, cols: int, rows: int, radix: int = 8):
    t = _lib.Ns2dTiles(cols, rows, radix)
    _lib.check(self._lib.jcfd_ns2d_plan_set_tiles(self._handle, ctypes.byref(t)),
               'jcfd_ns2d_plan_set_tiles')

  def close(self):
    if getattr(self, '_handle', None) is not None and self._handle.value:
      self._lib.jcfd_ns2d_plan_destroy(self._handle)
      self._handle = None

  def __del__(self):
    try:
      self.close()
    except Exception:
      pass

  def _field(self, x):
    t = _lib.as_device_tensor(x, torch.complex64, self.device)
    if tuple(t.shape) != self.shape:
      t = t.reshape(self.shape)
    return t

  def step(self, w, nsteps: int = 1, out=None):
    w = self._field(w)
    o = out if out is not None else torch.empty(self.shape, dtype=torch.complex64,
                                                device=self.device)
    with _lib.dlpack_args() as dl:
      rc = self._lib.jcfd_ns2d_step(self._handle, dl(w), dl(o), int(nsteps),
                                    _lib.stream_handle(self.device))
    _lib.check(rc, 'jcfd_ns2d_step')
    return o

  KERNELS = ('ns_cols_inverse', 'ns_rows', 'ns_cols_update_inverse', 'ns_cols_update')

  @property
  def LAUNCHES_PER_STEP(self):  # pylint: disable=invalid-name
    k = self.spec.stages
    return {'ns_cols_inverse': 1, 'ns_rows': k, 'ns_cols_update_inverse': k - 1,
            'ns_cols_update': 1}

  def time_kernels(self, w, reps: int = 50):
    w = self._field(w)
    ms = (ctypes.c_double * 4)()
    with _lib.dlpack_args() as dl:
      rc = self._lib.jcfd_ns2d_time_kernels(self._handle, dl(w), int(reps), ms,
                                            _lib.stream_handle(self.device))
    _lib.check(rc, 'jcfd_ns2d_time_kernels')
    return dict(zip(self.KERNELS, list(ms)))

  def kernel_bytes(self):
    b = (ctypes.c_int64 * 4)()
    _lib.check(self._lib.jcfd_ns2d_kernel_bytes(self._handle, b), 'jcfd_ns2d_kernel_bytes')
    return dict(zip(self.KERNELS, [int(x) for x in b]))

  @property
  def real_shape(self):
    n = self.spec.n
    return (n, n) if self.batch == 1 else (self.batch, n, n)

  def rfft2(self, x):
    x = _lib.as_device_tensor(x, torch.float32, self.device).reshape(self.real_shape)
    o = torch.empty(self.shape, dtype=torch.complex64, device=self.device)
    with _lib.dlpack_args() as dl:
      rc = self._lib.jcfd_ns2d_rfft2(self._handle, dl(x), dl(o),
                                     _lib.stream_handle(self.device))
    _lib.check(rc, 'jcfd_ns2d_rfft2')
    return o

  def irfft2(self, x_hat):
    x_hat = self._field(x_hat)
    o = torch.empty(self.real_shape, dtype=torch.float32, device=self.device)
    with _lib.dlpack_args() as dl:
      rc = self._lib.jcfd_ns2d_irfft2(self._handle, dl(x_hat), dl(o),
                                      _lib.stream_handle(self.device))
    _lib.check(rc, 'jcfd_ns2d_irfft2')
    return o

  def step_slots(self, w: torch.Tensor, slots: torch.Tensor, nsteps: int):
    """Advance only `slots` (int32 device tensor) of the (batch, n, n/2+1) w in place."""
    with _lib.dlpack_args() as dl:
      rc = self._lib.jcfd_ns2d_step_slots(self._handle, dl(w), dl(slots), int(nsteps),
                                          _lib.stream_handle(self.device))
    _lib.check(rc, 'jcfd_ns2d_step_slots')

  def explicit_terms(self, w):
    w = self._field(w)
    o = torch.empty(self.shape, dtype=torch.complex64, device=self.device)
    with _lib.dlpack_args() as dl:
      rc = self._lib.jcfd_ns2d_explicit_terms(self._handle, dl(w), dl(o),
                                              _lib.stream_handle(self.device))
    _lib.check(rc, 'jcfd_ns2d_explicit_terms')
    return o


class Ns2dGroup:
  """Mixed-size group stepping (jcfd_ns2d_group_*): one launch per kernel
  class and tableau stage advances every listed slot of every member
  plan, whatever its grid size, each with exactly the arithmetic of a lone
  step."""

  def __init__(self, device: Optional[torch.device] = None):
    lib = _lib.load()
    self.device = device if device is not None else torch.device(
        'cuda', torch.cuda.current_device())
    handle = ctypes.c_void_p()
    with torch.cuda.device(self.device):
      _lib.check(lib.jcfd_ns2d_group_create(ctypes.byref(handle)), 'jcfd_ns2d_group_create')
    self._handle = handle
    self._lib = lib

  def set_graphs(self, enabled: bool):
    _lib.check(self._lib.jcfd_ns2d_group_set_graphs(self._handle, int(bool(enabled))),
               'jcfd_ns2d_group_set_graphs')

  def close(self):
    if getattr(self, '_handle', None) is not None and self._handle.value:
      self._lib.jcfd_ns2d_group_destroy(self._handle)
      self._handle = None

  def __del__(self):
    try:
      self.close()
    except Exception:  # interpreter shutdown
      pass

  def step(self, members, nsteps: int):
    """members: iterable of (plan, w_hat, slots) with w_hat the plan's
    complex64 device state of shape plan.shape and slots a sequence of ints;
    every listed slot advances nsteps steps in place."""
    members = list(members)
    arr = (_lib.Ns2dGroupMember * max(1, len(members)))()
    keep = []
    with _lib.dlpack_args() as dl:
      for k, (plan, w, slots) in enumerate(members):
        idx = np.ascontiguousarray(np.asarray(slots, dtype=np.int32).reshape(-1))
        keep.append(idx)
        arr[k].plan = plan._handle
        arr[k].w_hat = dl(w)
        arr[k].slots = idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        arr[k].n_slots = int(idx.size)
      rc = self._lib.jcfd_ns2d_group_step(self._handle, ctypes.cast(arr, ctypes.c_void_p),
                                          len(members), int(nsteps),
                                          _lib.stream_handle(self.device))
    _lib.check(rc, 'jcfd_ns2d_group_step')

  def time_kernels(self, reps: int = 20):
    """Mean ms per launch of each kernel class over `reps` further steps of
    the last member set (on the plans' internal states)."""
    ms = (ctypes.c_double * 4)()
    rc = self._lib.jcfd_ns2d_group_time_kernels(self._handle, int(reps), ms,
                                                _lib.stream_handle(self.device))
    _lib.check(rc, 'jcfd_ns2d_group_time_kernels')
    return dict(zip(Ns2dPlan.KERNELS, list(ms)))

  def launch_shape(self):
    """(grid, items, algorithmic bytes) per kernel class of the last step."""
    grid, items = (ctypes.c_int32 * 4)(), (ctypes.c_int32 * 4)()
    nbytes = (ctypes.c_int64 * 4)()
    _lib.check(self._lib.jcfd_ns2d_group_launch_shape(self._handle, grid, items, nbytes),
               'jcfd_ns2d_group_launch_shape')
    return {k: {'grid': int(g), 'items': int(i), 'bytes': int(b)}
            for k, g, i, b in zip(Ns2dPlan.KERNELS, grid, items, nbytes)}
