"""Grid geometry and offset-carrying arrays (jax_cfd/base/grids.py).

Same constructor semantics, offsets and ``'ij'`` layout as the reference
(grids.py:40-108 GridArray, 265-395 GridVariable, 499-658 Grid).  Array data
are torch tensors; the hot-path kernels receive them through DLPack.  The
arithmetic on GridArray is element-wise torch arithmetic with the reference's
offset-consistency rules -- user-side plumbing, never the step itself.
"""
from __future__ import annotations

import dataclasses
import numbers
import operator
from typing import Optional, Sequence, Tuple, Union

import numpy as np
import torch


class InconsistentOffsetError(Exception):
  """Raised for cases of inconsistent offset in GridArrays (grids.py:457)."""


class InconsistentGridError(Exception):
  """Raised for cases of inconsistent grids between GridArrays."""


@dataclasses.dataclass(init=False, frozen=True)
class Grid:
  """Arakawa C-grid geometry (grids.py:499-555)."""
  shape: Tuple[int, ...]
  step: Tuple[float, ...]
  domain: Tuple[Tuple[float, float], ...]

  def __init__(self, shape: Sequence[int],
               step: Optional[Union[float, Sequence[float]]] = None,
               domain: Optional[Union[float, Sequence[Tuple[float, float]]]] = None):
    shape = tuple(operator.index(s) for s in shape)
    object.__setattr__(self, 'shape', shape)
    if step is not None and domain is not None:
      raise TypeError('cannot provide both step and domain')
    elif domain is not None:
      if isinstance(domain, (int, float)):
        domain = ((0, domain),) * len(shape)
      else:
        if len(domain) != self.ndim:
          raise ValueError('length of domain does not match ndim: '
                           f'{len(domain)} != {self.ndim}')
        for bounds in domain:
          if len(bounds) != 2:
            raise ValueError(
                f'domain is not sequence of pairs of numbers: {domain}')
      domain = tuple((float(lo), float(hi)) for lo, hi in domain)
    else:
      if step is None:
        step = 1
      if isinstance(step, numbers.Number):
        step = (step,) * self.ndim
      elif len(step) != self.ndim:
        raise ValueError('length of step does not match ndim: '
                         f'{len(step)} != {self.ndim}')
      domain = tuple((0.0, float(s * n)) for s, n in zip(step, shape))
    object.__setattr__(self, 'domain', domain)
    step = tuple((hi - lo) / n for (lo, hi), n in zip(domain, shape))
    object.__setattr__(self, 'step', step)

  @property
  def ndim(self) -> int:
    return len(self.shape)

  @property
  def cell_center(self) -> Tuple[float, ...]:
    return self.ndim * (0.5,)

  @property
  def cell_faces(self) -> Tuple[Tuple[float, ...], ...]:
    d = self.ndim
    offsets = (np.eye(d) + np.ones([d, d])) / 2.
    return tuple(tuple(float(o) for o in offset) for offset in offsets)

  def stagger(self, v):
    return tuple(GridArray(u, o, self) for u, o in zip(v, self.cell_faces))

  def center(self, v):
    return tuple(GridArray(u, self.cell_center, self) for u in v)

  def axes(self, offset=None) -> Tuple[np.ndarray, ...]:
    """Grid points per axis, fp32 like the reference (grids.py:584-602)."""
    if offset is None:
      offset = self.cell_center
    if len(offset) != self.ndim:
      raise ValueError(f'unexpected offset length: {len(offset)} vs {self.ndim}')
    out = []
    for (lo, _), o, n, h in zip(self.domain, offset, self.shape, self.step):
      x = (np.arange(n, dtype=np.int32) + np.float32(o)).astype(np.float32)
      out.append((np.float32(lo) + x * np.float32(h)).astype(np.float32))
    return tuple(out)

  def fft_axes(self) -> Tuple[np.ndarray, ...]:
    return tuple(_fftfreq32(n, s) for n, s in zip(self.shape, self.step))

  def rfft_axes(self) -> Tuple[np.ndarray, ...]:
    fft = tuple(_fftfreq32(n, s) for n, s in zip(self.shape[:-1], self.step[:-1]))
    return fft + (_rfftfreq32(self.shape[-1], self.step[-1]),)

  def mesh(self, offset=None) -> Tuple[np.ndarray, ...]:
    return tuple(np.meshgrid(*self.axes(offset), indexing='ij'))

  def rfft_mesh(self) -> Tuple[np.ndarray, ...]:
    return tuple(np.meshgrid(*self.rfft_axes(), indexing='ij'))

  def eval_on_mesh(self, fn, offset=None) -> 'GridArray':
    if offset is None:
      offset = self.cell_center
    return GridArray(torch.as_tensor(fn(*self.mesh(offset))), offset, self)


def _fftfreq32(n, d):
  k = np.concatenate([np.arange((n + 1) // 2), -np.arange(n // 2, 0, -1)])
  return (k.astype(np.float32) / np.float32(d * n)).astype(np.float32)


def _rfftfreq32(n, d):
  return (np.arange(n // 2 + 1).astype(np.float32) / np.float32(d * n)).astype(np.float32)


def _data(x):
  return x.data if isinstance(x, GridArray) else x


@dataclasses.dataclass
class GridArray:
  """Data with an alignment offset and a grid (grids.py:40-108)."""
  data: torch.Tensor
  offset: Tuple[float, ...]
  grid: Grid

  @property
  def dtype(self):
    return self.data.dtype

  @property
  def shape(self) -> Tuple[int, ...]:
    return tuple(self.data.shape)

  def _binary(self, other, op, reverse=False):
    if isinstance(other, GridArray):
      offset = consistent_offset(self, other)
      grid = consistent_grid(self, other)
    elif isinstance(other, (numbers.Number, torch.Tensor, np.ndarray)):
      offset, grid = self.offset, self.grid
    else:
      return NotImplemented
    a, b = self.data, _data(other)
    if isinstance(b, np.ndarray) and isinstance(a, torch.Tensor):
      b = torch.as_tensor(b, device=a.device)
    return GridArray(op(b, a) if reverse else op(a, b), offset, grid)

  def __add__(self, o): return self._binary(o, operator.add)
  def __radd__(self, o): return self._binary(o, operator.add, True)
  def __sub__(self, o): return self._binary(o, operator.sub)
  def __rsub__(self, o): return self._binary(o, operator.sub, True)
  def __mul__(self, o): return self._binary(o, operator.mul)
  def __rmul__(self, o): return self._binary(o, operator.mul, True)
  def __truediv__(self, o): return self._binary(o, operator.truediv)
  def __rtruediv__(self, o): return self._binary(o, operator.truediv, True)
  def __neg__(self): return GridArray(-self.data, self.offset, self.grid)
  def __abs__(self): return GridArray(self.data.abs(), self.offset, self.grid)


GridArrayVector = Tuple[GridArray, ...]


@dataclasses.dataclass
class GridVariable:
  """A GridArray with boundary conditions (grids.py:265-395)."""
  array: GridArray
  bc: 'BoundaryConditions'

  def __post_init__(self):
    if not isinstance(self.array, GridArray):
      raise ValueError(
          f'Expected array type to be GridArray, got {type(self.array)}')
    if len(self.bc.types) != self.grid.ndim:
      raise ValueError(
          'Incompatible dimension between grid and bc, grid dimension = '
          f'{self.grid.ndim}, bc dimension = {len(self.bc.types)}')

  @property
  def dtype(self):
    return self.array.dtype

  @property
  def shape(self):
    return self.array.shape

  @property
  def data(self):
    return self.array.data

  @property
  def offset(self):
    return self.array.offset

  @property
  def grid(self):
    return self.array.grid

  def shift(self, offset: int, axis: int, mode: Optional[str] = 'extend') -> GridArray:
    return self.bc.shift(self.array, offset, axis, mode)

  def impose_bc(self) -> 'GridVariable':
    return self.bc.impose_bc(self.array)

  def trim_boundary(self) -> GridArray:
    return self.bc.trim_boundary(self.array)


GridVariableVector = Tuple[GridVariable, ...]


def consistent_offset(*arrays) -> Tuple[float, ...]:
  offsets = {a.offset for a in arrays}
  if len(offsets) != 1:
    raise InconsistentOffsetError(f'arrays do not have a unique offset: {offsets}')
  offset, = offsets
  return offset


def consistent_grid(*arrays) -> Grid:
  grids = {a.grid for a in arrays}
  if len(grids) != 1:
    raise InconsistentGridError(f'arrays do not have a unique grid: {grids}')
  grid, = grids
  return grid


def averaged_offset(*arrays) -> Tuple[float, ...]:
  offset = np.mean([a.offset for a in arrays], axis=0)
  return tuple(offset.tolist())


def control_volume_offsets(c) -> Tuple[Tuple[float, ...], ...]:
  """grids.py:448-454."""
  return tuple(tuple(o + .5 if i == j else o for i, o in enumerate(c.offset))
               for j in range(len(c.offset)))

