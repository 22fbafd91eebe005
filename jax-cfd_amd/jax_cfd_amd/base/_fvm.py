"""FVM plans: the Python side of the jcfd_fvm_* C ABI."""
from __future__ import annotations

import ctypes
import dataclasses
import functools
from typing import Optional, Tuple

import numpy as np
import torch

from jax_cfd_amd import _lib


@dataclasses.dataclass(frozen=True)
class FvmSpec:
  """Everything a plan depends on (hashable, so plans can be shared)."""
  n: int
  dx: float
  dy: float
  dt: float = 0.0
  density: float = 1.0
  viscosity: Optional[float] = None
  convect: int = _lib.CONVECT_VAN_LEER
  pressure: int = _lib.PRESSURE_FAST_DIAG
  cg_rtol: float = 1e-6
  cg_atol: float = 1e-6
  cg_maxiter: int = 0
  linear_coefficient: Optional[float] = None  # None: no forcing at all
  forcing_x: Optional[Tuple[float, ...]] = None  # f_x(y_j), n values
  diffusion: int = _lib.DIFFUSION_EXPLICIT
  rk_stages: int = 0  # 0: forward Euler (fused); else the Butcher tableau below
  rk_a: Tuple[float, ...] = ()  # a[i-1][j] packed row by row
  rk_b: Tuple[float, ...] = ()
  diff_cg_rtol: float = 1e-6  # diffusion.solve_cg (diffusion == IMPLICIT_CG)
  diff_cg_atol: float = 1e-6
  diff_cg_maxiter: int = 0


def spec_from_grid(grid, **kw) -> FvmSpec:
  if grid.ndim != 2 or grid.shape[0] != grid.shape[1]:
    raise NotImplementedError(
        f'the FVM kernels are built for square 2-D grids, got shape {grid.shape}')
  return FvmSpec(n=grid.shape[0], dx=grid.step[0], dy=grid.step[1], **kw)


class FvmPlan:
  """Owns one jcfd_fvm_plan (device constants + scratch) on one GPU."""

  def __init__(self, spec: FvmSpec, device: Optional[torch.device] = None,
               batch: int = 1):
    lib = _lib.load()
    self.spec = spec
    self.batch = int(batch)
    self.device = device if device is not None else torch.device(
        'cuda', torch.cuda.current_device())
    cfg, self._fx = FvmPlan.config(spec, self.batch)
    handle = ctypes.c_void_p()
    with torch.cuda.device(self.device):
      _lib.check(lib.jcfd_fvm_plan_create(ctypes.byref(cfg), ctypes.byref(handle)),
                 'jcfd_fvm_plan_create')
    self._handle = handle
    self._lib = lib

  @staticmethod
  def config(spec: FvmSpec, batch: int = 1):
    """(jcfd_fvm_config of `spec`, the host forcing column it points at)."""
    cfg = _lib.FvmConfig()
    cfg.n = spec.n
    cfg.dx, cfg.dy = spec.dx, spec.dy
    cfg.dt = spec.dt
    cfg.density = spec.density
    cfg.has_viscosity = int(spec.viscosity is not None)
    cfg.viscosity = spec.viscosity if spec.viscosity is not None else 0.0
    cfg.convect = spec.convect
    cfg.pressure = spec.pressure
    cfg.cg_rtol, cfg.cg_atol = spec.cg_rtol, spec.cg_atol
    cfg.cg_maxiter = spec.cg_maxiter
    cfg.has_forcing = int(spec.linear_coefficient is not None)
    cfg.linear_coefficient = spec.linear_coefficient or 0.0
    fx = None
    if spec.forcing_x is not None:
      fx = np.ascontiguousarray(np.asarray(spec.forcing_x, np.float32))
      if fx.shape != (spec.n,):
        raise ValueError('forcing_x must hold n values')
      cfg.forcing_x = fx.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    cfg.batch = int(batch)
    cfg.diffusion = spec.diffusion
    cfg.rk_stages = spec.rk_stages
    for k, x in enumerate(spec.rk_a):
      cfg.rk_a[k] = x
    for k, x in enumerate(spec.rk_b):
      cfg.rk_b[k] = x
    cfg.diff_cg_rtol, cfg.diff_cg_atol = spec.diff_cg_rtol, spec.diff_cg_atol
    cfg.diff_cg_maxiter = spec.diff_cg_maxiter
    return cfg, fx

  @property
  def device_bytes(self) -> int:
    return int(self._lib.jcfd_fvm_plan_device_bytes(self._handle))

  @property
  def cg_iterations(self) -> int:
    """Iterations of the last CG pressure solve (CG plans only)."""
    return int(self._lib.jcfd_fvm_cg_iterations(self._handle))

  def diffusion_cg_iterations(self, component: int) -> int:
    """Iterations of the last CG diffusion solve of component 0 (u) or 1 (v)."""
    return int(self._lib.jcfd_fvm_diffusion_cg_iterations(self._handle, int(component)))

  def set_graphs(self, enabled: bool):
    _lib.check(self._lib.jcfd_fvm_plan_set_graphs(self._handle, int(bool(enabled))),
               'jcfd_fvm_plan_set_graphs')

  def set_cg_persistent(self, enabled: bool):
    """CG solves as one cooperative launch, or two launches per iteration
    (default, faster on MI355X); same results."""
    _lib.check(self._lib.jcfd_fvm_plan_set_cg_persistent(self._handle, int(bool(enabled))),
               'jcfd_fvm_plan_set_cg_persistent')

  @property
  def tiles(self) -> dict:
    """Tile shapes and radix schedule of the step kernels."""
    t = _lib.FvmTiles()
    _lib.check(self._lib.jcfd_fvm_plan_get_tiles(self._handle, ctypes.byref(t)),
               'jcfd_fvm_plan_get_tiles')
    return {k: int(getattr(t, k)) for k, _ in _lib.FvmTiles._fields_}

  def set_tiles(self, rows_forward: int, poisson_cols: int, rows_project: int,
                radix: int = 8):
    t = _lib.FvmTiles(rows_forward, poisson_cols, rows_project, radix)
    _lib.check(self._lib.jcfd_fvm_plan_set_tiles(self._handle, ctypes.byref(t)),
               'jcfd_fvm_plan_set_tiles')

  def close(self):
    if getattr(self, '_handle', None) is not None and self._handle.value:
      self._lib.jcfd_fvm_plan_destroy(self._handle)
      self._handle = None

  def __del__(self):
    try:
      self.close()
    except Exception:  # interpreter shutdown
      pass

  def _field(self, x) -> torch.Tensor:
    return _lib.as_device_tensor(x, torch.float32, self.device)

  @property
  def shape(self):
    n = self.spec.n
    return (n, n) if self.batch == 1 else (self.batch, n, n)

  def _empty(self, batched: bool = False):
    n = self.spec.n
    return torch.empty(self.shape if batched else (n, n), dtype=torch.float32,
                       device=self.device)

  def step_slots(self, u: torch.Tensor, v: torch.Tensor, slots: torch.Tensor,
                 nsteps: int):
    """Advance only `slots` (int32 device tensor) of (batch, n, n) u, v in place."""
    with _lib.dlpack_args() as dl:
      rc = self._lib.jcfd_fvm_step_slots(self._handle, dl(u), dl(v), dl(slots),
                                         int(nsteps), _lib.stream_handle(self.device))
    _lib.check(rc, 'jcfd_fvm_step_slots')

  def step(self, u, v, nsteps: int = 1, out=None):
    """nsteps forward-Euler + projection steps; returns (u, v) tensors."""
    u, v = self._field(u), self._field(v)
    uo, vo = out if out is not None else (self._empty(True), self._empty(True))
    with _lib.dlpack_args() as dl:
      rc = self._lib.jcfd_fvm_step(self._handle, dl(u), dl(v), dl(uo), dl(vo),
                                   int(nsteps), _lib.stream_handle(self.device))
    _lib.check(rc, 'jcfd_fvm_step')
    return uo, vo

  def step_noalloc(self, u, v, nsteps: int = 1, out=None):
    """As `step`, through jcfd_fvm_step_noalloc: the library allocates
    nothing (the step's scratch is a static per-GPU array); `out` should be
    passed by callers that cannot allocate either."""
    u, v = self._field(u), self._field(v)
    uo, vo = out if out is not None else (self._empty(True), self._empty(True))
    with _lib.dlpack_args() as dl:
      rc = self._lib.jcfd_fvm_step_noalloc(self._handle, dl(u), dl(v), dl(uo), dl(vo),
                                           int(nsteps), _lib.stream_handle(self.device))
    _lib.check(rc, 'jcfd_fvm_step_noalloc')
    return uo, vo

  def explicit_terms(self, u, v):
    u, v = self._field(u), self._field(v)
    du, dv = self._empty(), self._empty()
    with _lib.dlpack_args() as dl:
      rc = self._lib.jcfd_fvm_explicit_terms(self._handle, dl(u), dl(v), dl(du),
                                             dl(dv), _lib.stream_handle(self.device))
    _lib.check(rc, 'jcfd_fvm_explicit_terms')
    return du, dv

  def projection(self, u, v):
    u, v = self._field(u), self._field(v)
    uo, vo = self._empty(), self._empty()
    with _lib.dlpack_args() as dl:
      rc = self._lib.jcfd_fvm_projection(self._handle, dl(u), dl(v), dl(uo), dl(vo),
                                         _lib.stream_handle(self.device))
    _lib.check(rc, 'jcfd_fvm_projection')
    return uo, vo

  def diffuse(self, u, v):
    """Implicit diffusion solve (implicit-diffusion plans): (1 - nu dt L)^-1."""
    u, v = self._field(u), self._field(v)
    uo, vo = self._empty(), self._empty()
    with _lib.dlpack_args() as dl:
      rc = self._lib.jcfd_fvm_diffuse(self._handle, dl(u), dl(v), dl(uo), dl(vo),
                                      _lib.stream_handle(self.device))
    _lib.check(rc, 'jcfd_fvm_diffuse')
    return uo, vo

  def divergence(self, u, v):
    u, v = self._field(u), self._field(v)
    out = self._empty()
    with _lib.dlpack_args() as dl:
      rc = self._lib.jcfd_fvm_divergence(self._handle, dl(u), dl(v), dl(out),
                                         _lib.stream_handle(self.device))
    _lib.check(rc, 'jcfd_fvm_divergence')
    return out

  KERNELS = ('fvm_rows_forward', 'fvm_poisson_cols', 'fvm_rows_project')

  def time_kernels(self, u, v, reps: int = 200):
    """Mean ms per launch of each step kernel (HIP events on this stream)."""
    u, v = self._field(u), self._field(v)
    ms = (ctypes.c_double * 3)()
    with _lib.dlpack_args() as dl:
      rc = self._lib.jcfd_fvm_time_kernels(self._handle, dl(u), dl(v), int(reps), ms,
                                           _lib.stream_handle(self.device))
    _lib.check(rc, 'jcfd_fvm_time_kernels')
    return dict(zip(self.KERNELS, list(ms)))

  def kernel_bytes(self):
    b = (ctypes.c_int64 * 3)()
    _lib.check(self._lib.jcfd_fvm_kernel_bytes(self._handle, b), 'jcfd_fvm_kernel_bytes')
    return dict(zip(self.KERNELS, [int(x) for x in b]))

  def step_vjp(self, u, v, lam_u, lam_v):
    """Cotangent of (u, v) through one step, given the cotangent of its output."""
    u, v = self._field(u), self._field(v)
    lu, lv = self._field(lam_u), self._field(lam_v)
    ou, ov = self._empty(), self._empty()
    with _lib.dlpack_args() as dl:
      rc = self._lib.jcfd_fvm_step_vjp(self._handle, dl(u), dl(v), dl(lu), dl(lv), dl(ou),
                                       dl(ov), _lib.stream_handle(self.device))
    _lib.check(rc, 'jcfd_fvm_step_vjp')
    return ou, ov

  def curl_vjp(self, w_bar):
    wb = self._field(w_bar)
    ub, vb = self._empty(), self._empty()
    with _lib.dlpack_args() as dl:
      rc = self._lib.jcfd_fvm_curl_vjp(self._handle, dl(wb), dl(ub), dl(vb),
                                       _lib.stream_handle(self.device))
    _lib.check(rc, 'jcfd_fvm_curl_vjp')
    return ub, vb

  def advect_semilagrangian(self, c, offset, u, v, dt: float):
    """Semi-Lagrangian step of c (at `offset`) by the staggered (u, v)."""
    c, u, v = self._field(c), self._field(u), self._field(v)
    out = self._empty()
    with _lib.dlpack_args() as dl:
      rc = self._lib.jcfd_fvm_advect_semilagrangian(
          self._handle, dl(c), float(offset[0]), float(offset[1]), dl(u), dl(v),
          float(dt), dl(out), _lib.stream_handle(self.device))
    _lib.check(rc, 'jcfd_fvm_advect_semilagrangian')
    return out

  def curl(self, u, v):
    u, v = self._field(u), self._field(v)
    out = self._empty()
    with _lib.dlpack_args() as dl:
      rc = self._lib.jcfd_fvm_curl(self._handle, dl(u), dl(v), dl(out),
                                   _lib.stream_handle(self.device))
    _lib.check(rc, 'jcfd_fvm_curl')
    return out

  def spectral_filter(self, x, multiplier):
    x = self._field(x)
    m = _lib.as_device_tensor(multiplier, torch.float32, self.device)
    out = self._empty()
    with _lib.dlpack_args() as dl:
      rc = self._lib.jcfd_fvm_spectral_filter(self._handle, dl(x), dl(m), dl(out),
                                              _lib.stream_handle(self.device))
    _lib.check(rc, 'jcfd_fvm_spectral_filter')
    return out


class FvmGroup:
  """Mixed-size group stepping (jcfd_fvm_group_*): one launch per kernel
  class advances every listed slot of every member plan, whatever its
  grid size, each with exactly the arithmetic of a lone step."""

  def __init__(self, device: Optional[torch.device] = None):
    lib = _lib.load()
    self.device = device if device is not None else torch.device(
        'cuda', torch.cuda.current_device())
    handle = ctypes.c_void_p()
    with torch.cuda.device(self.device):
      _lib.check(lib.jcfd_fvm_group_create(ctypes.byref(handle)), 'jcfd_fvm_group_create')
    self._handle = handle
    self._lib = lib

  def set_graphs(self, enabled: bool):
    _lib.check(self._lib.jcfd_fvm_group_set_graphs(self._handle, int(bool(enabled))),
               'jcfd_fvm_group_set_graphs')

  def close(self):
    if getattr(self, '_handle', None) is not None and self._handle.value:
      self._lib.jcfd_fvm_group_destroy(self._handle)
      self._handle = None

  def __del__(self):
    try:
      self.close()
    except Exception:  # interpreter shutdown
      pass

  def step(self, members, nsteps: int):
    """members: iterable of (plan, u, v, slots) with u, v the plan's
    (batch, n, n) float32 device fields and slots a sequence of ints; every
    listed slot advances nsteps steps in place."""
    members = list(members)
    arr = (_lib.FvmGroupMember * max(1, len(members)))()
    keep = []
    with _lib.dlpack_args() as dl:
      for k, (plan, u, v, slots) in enumerate(members):
        idx = np.ascontiguousarray(np.asarray(slots, dtype=np.int32).reshape(-1))
        keep.append(idx)
        arr[k].plan = plan._handle
        arr[k].u = dl(u)
        arr[k].v = dl(v)
        arr[k].slots = idx.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))
        arr[k].n_slots = int(idx.size)
      rc = self._lib.jcfd_fvm_group_step(self._handle, ctypes.cast(arr, ctypes.c_void_p),
                                         len(members), int(nsteps),
                                         _lib.stream_handle(self.device))
    _lib.check(rc, 'jcfd_fvm_group_step')

  def time_kernels(self, reps: int = 20):
    """Mean ms per launch of each phase over `reps` further steps of the last
    member set (their states advance)."""
    ms = (ctypes.c_double * 3)()
    rc = self._lib.jcfd_fvm_group_time_kernels(self._handle, int(reps), ms,
                                               _lib.stream_handle(self.device))
    _lib.check(rc, 'jcfd_fvm_group_time_kernels')
    return dict(zip(FvmPlan.KERNELS, list(ms)))

  def launch_shape(self):
    """(grid, items, algorithmic bytes) per phase of the last step."""
    grid, items = (ctypes.c_int32 * 3)(), (ctypes.c_int32 * 3)()
    nbytes = (ctypes.c_int64 * 3)()
    _lib.check(self._lib.jcfd_fvm_group_launch_shape(self._handle, grid, items, nbytes),
               'jcfd_fvm_group_launch_shape')
    return {k: {'grid': int(g), 'items': int(i), 'bytes': int(b)}
            for k, g, i, b in zip(FvmPlan.KERNELS, grid, items, nbytes)}


@functools.lru_cache(maxsize=64)
def _cached_plan(spec: FvmSpec, device_index: int) -> FvmPlan:
  return FvmPlan(spec, torch.device('cuda', device_index))


def plan_for(spec: FvmSpec, device: Optional[torch.device] = None) -> FvmPlan:
  if device is None:
    device = torch.device('cuda', torch.cuda.current_device())
  return _cached_plan(spec, device.index if device.index is not None else 0)
