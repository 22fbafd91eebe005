"""Rollout drivers (jax_cfd/base/funcutils.py:40-126).

``scan`` / ``init_context`` keep jax.lax.scan's contract for host-side loops.

``repeated`` hands a GPU step function the whole step count in one library
call.  ``trajectory`` stacks post-processed frames along a new leading axis,
as ``jax.lax.scan`` stacks pytree leaves, into one tensor per leaf that is
allocated once for the whole rollout.  For a GPU step function with the
identity ``post_process`` the step kernels write each frame straight into its
slot of that tensor (``rollout_frames``), so a frame costs no copy.
"""
from __future__ import annotations

import contextlib
from typing import Any, Callable, Optional

import numpy as np
import torch

from jax_cfd_amd.base import grids

_INITIALIZING = 0


@contextlib.contextmanager
def init_context():
  """Within this context scan() evaluates f() once (funcutils.py:40-56)."""
  global _INITIALIZING
  _INITIALIZING += 1
  try:
    yield
  finally:
    _INITIALIZING -= 1


def _index(xs, i: int):
  """xs[i] of every leaf."""
  if xs is None:
    return None
  if isinstance(xs, grids.GridVariable):
    return grids.GridVariable(_index(xs.array, i), xs.bc)
  if isinstance(xs, grids.GridArray):
    return grids.GridArray(xs.data[i], xs.offset, xs.grid)
  if isinstance(xs, (tuple, list)):
    return type(xs)(_index(x, i) for x in xs)
  if isinstance(xs, dict):
    return {k: _index(v, i) for k, v in xs.items()}
  return xs[i]


def _leading_length(xs) -> int:
  if isinstance(xs, (grids.GridVariable, grids.GridArray)):
    return int(xs.data.shape[0])
  if isinstance(xs, (tuple, list)):
    return _leading_length(xs[0])
  if isinstance(xs, dict):
    return _leading_length(next(iter(xs.values())))
  return int(xs.shape[0])


def scan(f: Callable, init: Any, xs: Any, length: Optional[int] = None):
  """jax.lax.scan semantics (funcutils.py:65-79): carry, y = f(carry, x)
  over the leading axis of xs (or `length` times with xs None); returns the
  final carry and the stacked ys (None if f returns no y, or for length 0).
  Within init_context() f is evaluated once and its y repeated."""
  if length is None:
    if xs is None:
      raise ValueError('scan needs xs or length')
    length = _leading_length(xs)
  length = int(length)
  if _INITIALIZING:
    carry, y0 = f(init, _index(xs, 0))
    if y0 is None:
      return carry, None
    ys = _allocate(y0, length)
    for t in range(length):
      _write(ys, t, y0)
    return carry, ys
  carry, ys = init, None
  for t in range(length):
    carry, y = f(carry, _index(xs, t))
    if y is not None:
      if ys is None:
        ys = _allocate(y, length)
      _write(ys, t, y)
  return carry, ys


def _identity(x):
  return x


def repeated(f: Callable, steps: int) -> Callable:
  """Returns a repeatedly applied version of f() (funcutils.py:82-88)."""
  multistep = getattr(f, 'multistep', None)
  rollout = getattr(f, 'rollout_frames', None)

  def f_repeated(x_initial):
    if multistep is not None:
      return multistep(x_initial, steps)
    x = x_initial
    for _ in range(steps):
      x = f(x)
    return x
  f_repeated.multistep = (
      (lambda x, k: multistep(x, k * steps)) if multistep is not None else None)
  f_repeated.rollout_frames = (
      (lambda x, inner, count, start_with_input:
       rollout(x, inner * steps, count, start_with_input))
      if rollout is not None else None)
  return f_repeated


def _allocate(frame, count: int):
  """Empty stacked storage for `count` frames shaped like `frame`."""
  if isinstance(frame, grids.GridVariable):
    return grids.GridVariable(_allocate(frame.array, count), frame.bc)
  if isinstance(frame, grids.GridArray):
    return grids.GridArray(_allocate(frame.data, count), frame.offset, frame.grid)
  if isinstance(frame, (tuple, list)):
    return type(frame)(_allocate(x, count) for x in frame)
  if isinstance(frame, dict):
    return {k: _allocate(v, count) for k, v in frame.items()}
  if isinstance(frame, torch.Tensor):
    return torch.empty((count,) + tuple(frame.shape), dtype=frame.dtype,
                       device=frame.device)
  if isinstance(frame, (np.ndarray, np.generic, float, int)):
    a = np.asarray(frame)
    return np.empty((count,) + a.shape, dtype=a.dtype)
  raise TypeError(f'cannot stack frames of type {type(frame)}')


def _write(frames, index: int, frame):
  """frames[index] = frame, leaf by leaf."""
  if isinstance(frames, grids.GridVariable):
    _write(frames.array, index, frame.array)
  elif isinstance(frames, grids.GridArray):
    if tuple(frame.offset) != tuple(frames.offset):
      raise grids.InconsistentOffsetError(
          f'frame offset {frame.offset} != {frames.offset}')
    _write(frames.data, index, frame.data)
  elif isinstance(frames, (tuple, list)):
    if len(frames) != len(frame):
      raise ValueError('frames change structure along the trajectory')
    for f, x in zip(frames, frame):
      _write(f, index, x)
  elif isinstance(frames, dict):
    for k, f in frames.items():
      _write(f, index, frame[k])
  elif isinstance(frames, torch.Tensor):
    frames[index].copy_(frame)
  else:
    frames[index] = np.asarray(frame)


def trajectory(step_fn: Callable, steps: int, post_process: Callable = _identity,
               *, start_with_input: bool = False) -> Callable:
  """funcutils.py:95-126: returns (final state, stacked frames).

  Frame t is post_process(state after t+1 steps), or after t steps when
  `start_with_input`; the final state is the one after `steps` steps, before
  post_process.  `steps == 0` gives empty frames and the input state.
  """
  steps = int(steps)
  if steps < 0:
    raise ValueError(f'steps must be >= 0, got {steps}')
  rollout = getattr(step_fn, 'rollout_frames', None)

  def multistep(values):
    if rollout is not None and post_process is _identity:
      return rollout(values, 1, steps, start_with_input)
    if steps == 0:
      return values, _allocate(post_process(values), 0)
    frames = None
    state = values
    for t in range(steps):
      new_state = step_fn(state)
      frame = post_process(state if start_with_input else new_state)
      if frames is None:
        frames = _allocate(frame, steps)
      _write(frames, t, frame)
      state = new_state
    return state, frames
  return multistep
