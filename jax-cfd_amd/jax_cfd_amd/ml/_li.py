"""Learned-interpolation plans: the Python side of the jcfd_li_* C ABI.

The tower's output layer is handed to the library twice: as the logits
layer (the network's own output, for the logits seam) and folded with the
interpolations' constraints into the 128-channel coefficient layer the step
uses.  Folding is done here in fp64 and rounded once to fp32:
W' = W . N_k and b' = b . N_k + bias_k per interpolation k (N_k its
nullspace, bias_k its low-order solution), the same map the reference
applies to the logits after the network (layers.py:531-544).
"""
from __future__ import annotations

import ctypes
import dataclasses
from typing import List, Optional, Sequence, Tuple

import numpy as np
import torch

from jax_cfd_amd import _lib
from jax_cfd_amd.base import _fvm

COEF = _lib.LI_COEF_CHANNELS


@dataclasses.dataclass(frozen=True)
class LiSpec:
  fvm: _fvm.FvmSpec
  n_layers: int
  hidden_channels: int
  n_logits: int
  interp_key: Tuple[int, ...]  # 8 first-coefficient channels, see jcfd_li_config


def interp_keys(tags: Sequence) -> Tuple[int, ...]:
  """interp_key of jcfd_li_config for interpolations split in the order
  (component m, face axis t, tag) with `tags` = ('u', 'c') in any order:
  interpolation (m, t, g) is number (2 m + t) len(tags) + g."""
  tags = tuple(tags)
  if len(tags) != 2 or set(tags) != {'u', 'c'}:
    raise NotImplementedError("the learned self-advection kernels need tags ('u', 'c')")
  tu, tc = tags.index('u'), tags.index('c')
  idx = lambda m, t, g: 16 * ((2 * m + t) * 2 + g)  # noqa: E731
  # u: x face (advected u, velocity u), y face (advected u, velocity v);
  # v: x face (advected v, velocity u), y face (advected v, velocity v)
  return (idx(0, 0, tc), idx(0, 0, tu), idx(0, 1, tc), idx(1, 0, tu),
          idx(1, 0, tc), idx(0, 1, tu), idx(1, 1, tc), idx(1, 1, tu))


def fold_output_layer(w: np.ndarray, b: np.ndarray, constraints) -> Tuple[np.ndarray, np.ndarray]:
  """(3, 3, cin, n_logits) logits layer -> (3, 3, cin, 128) coefficient
  layer; constraints: [(nullspace (s_k, 16), bias (16,))] in split order."""
  w64, b64 = np.asarray(w, np.float64), np.asarray(b, np.float64)
  ws, bs, off = [], [], 0
  for null, bias in constraints:
    null = np.asarray(null, np.float64)
    s = null.shape[0]
    if null.shape[1] != 16:
      raise NotImplementedError('the kernels are built for 4 x 4 stencils')
    ws.append(np.tensordot(w64[..., off:off + s], null, axes=[-1, 0]))
    bs.append(b64[off:off + s] @ null + np.asarray(bias, np.float64))
    off += s
  if off != w64.shape[-1] or len(constraints) != 8:
    raise ValueError('output channels do not match the 8 interpolations')
  return (np.ascontiguousarray(np.concatenate(ws, -1), np.float32),
          np.ascontiguousarray(np.concatenate(bs), np.float32))


def _fp(a: np.ndarray):
  return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class LiPlan:
  """Owns one jcfd_li_plan: the tower weights on the device and the
  learned-interpolation step."""

  def __init__(self, spec: LiSpec, params, constraints, device: Optional[torch.device] = None,
               batch: int = 1):
    lib = _lib.load()
    self.spec = spec
    self.batch = int(batch)
    self.device = device if device is not None else torch.device(
        'cuda', torch.cuda.current_device())
    fplan_cfg = _fvm.FvmPlan.config(spec.fvm, self.batch)
    cfg = _lib.LiConfig()
    cfg.fvm = fplan_cfg[0]
    self._keep = fplan_cfg[1]
    cfg.n_layers = spec.n_layers
    cfg.hidden_channels = spec.hidden_channels
    cfg.n_logits = spec.n_logits
    for k, x in enumerate(spec.interp_key):
      cfg.interp_key[k] = x
    handle = ctypes.c_void_p()
    with torch.cuda.device(self.device):
      _lib.check(lib.jcfd_li_plan_create(ctypes.byref(cfg), ctypes.byref(handle)),
                 'jcfd_li_plan_create')
    self._handle = handle
    self._lib = lib
    self.set_params(params, constraints)

  def set_params(self, params, constraints):
    params = list(params)
    if len(params) != self.spec.n_layers:
      raise ValueError(f'expected {self.spec.n_layers} layers')
    for k, (w, b) in enumerate(params[:-1]):
      w = np.ascontiguousarray(w, np.float32)
      b = np.ascontiguousarray(b, np.float32)
      _lib.check(self._lib.jcfd_li_set_layer(self._handle, k, _fp(w), _fp(b)), 'jcfd_li_set_layer')
    w, b = params[-1]
    wc, bc = fold_output_layer(w, b, constraints)
    _lib.check(self._lib.jcfd_li_set_layer(self._handle, self.spec.n_layers - 1, _fp(wc), _fp(bc)),
               'jcfd_li_set_layer')
    wl = np.ascontiguousarray(w, np.float32)
    bl = np.ascontiguousarray(b, np.float32)
    _lib.check(self._lib.jcfd_li_set_logits_layer(self._handle, _fp(wl), _fp(bl)),
               'jcfd_li_set_logits_layer')

  @property
  def shape(self):
    n = self.spec.fvm.n
    return (n, n) if self.batch == 1 else (self.batch, n, n)

  @property
  def device_bytes(self) -> int:
    return int(self._lib.jcfd_li_plan_device_bytes(self._handle))

  def close(self):
    if getattr(self, '_handle', None) is not None and self._handle.value:
      self._lib.jcfd_li_plan_destroy(self._handle)
      self._handle = None

  def __del__(self):
    try:
      self.close()
    except Exception:
      pass

  def _field(self, x):
    t = _lib.as_device_tensor(x, torch.float32, self.device)
    if tuple(t.shape) != self.shape:
      t = t.reshape(self.shape)
    return t

  def _empty(self, *lead):
    return torch.empty(lead + self.shape if lead else self.shape, dtype=torch.float32,
                       device=self.device)

  def step(self, u, v, nsteps: int = 1, out=None):
    u, v = self._field(u), self._field(v)
    uo, vo = out if out is not None else (self._empty(), self._empty())
    with _lib.dlpack_args() as dl:
      rc = self._lib.jcfd_li_step(self._handle, dl(u), dl(v), dl(uo), dl(vo), int(nsteps),
                                  _lib.stream_handle(self.device))
    _lib.check(rc, 'jcfd_li_step')
    return uo, vo

  def tower(self, u, v, coefficients: bool = True):
    """(batch, 128, n, n): stencil coefficients, or the logits (first
    n_logits channels) when coefficients is False."""
    u, v = self._field(u), self._field(v)
    n = self.spec.fvm.n
    out = torch.empty((self.batch * COEF, n, n), dtype=torch.float32, device=self.device)
    with _lib.dlpack_args() as dl:
      rc = self._lib.jcfd_li_tower(self._handle, dl(u), dl(v), dl(out), int(bool(coefficients)),
                                   _lib.stream_handle(self.device))
    _lib.check(rc, 'jcfd_li_tower')
    return out.reshape(self.batch, COEF, n, n)

  def explicit_terms(self, u, v):
    u, v = self._field(u), self._field(v)
    du, dv = self._empty(), self._empty()
    with _lib.dlpack_args() as dl:
      rc = self._lib.jcfd_li_explicit_terms(self._handle, dl(u), dl(v), dl(du), dl(dv),
                                            _lib.stream_handle(self.device))
    _lib.check(rc, 'jcfd_li_explicit_terms')
    return du, dv

  def interpolate(self, coefficients: torch.Tensor, x, index: int):
    n = self.spec.fvm.n
    coef = coefficients.reshape(self.batch * COEF, n, n)
    x = self._field(x)
    out = self._empty()
    with _lib.dlpack_args() as dl:
      rc = self._lib.jcfd_li_interpolate(self._handle, dl(coef), dl(x), int(index), dl(out),
                                         _lib.stream_handle(self.device))
    _lib.check(rc, 'jcfd_li_interpolate')
    return out

  def time_kernels(self, u, v, reps: int = 20) -> List[float]:
    """ms per launch: each convolution, the explicit-terms kernel, then the
    combination with its projection (three launches together)."""
    u, v = self._field(u), self._field(v)
    ms = (ctypes.c_double * (self.spec.n_layers + 2))()
    with _lib.dlpack_args() as dl:
      rc = self._lib.jcfd_li_time_kernels(self._handle, dl(u), dl(v), int(reps), ms,
                                          _lib.stream_handle(self.device))
    _lib.check(rc, 'jcfd_li_time_kernels')
    return list(ms)

  def kernel_flops(self) -> List[float]:
    f = (ctypes.c_double * self.spec.n_layers)()
    _lib.check(self._lib.jcfd_li_kernel_flops(self._handle, f), 'jcfd_li_kernel_flops')
    return list(f)
