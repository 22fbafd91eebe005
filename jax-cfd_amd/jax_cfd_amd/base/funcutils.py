"""Rollout drivers (jax_cfd/base/funcutils.py:82-126).

``repeated`` hands a GPU step function the whole step count in one library
call; ``trajectory`` stacks post-processed frames along a new leading axis,
as ``jax.lax.scan`` stacks pytree leaves.
"""
from __future__ import annotations

from typing import Callable

import torch

from jax_cfd_amd.base import grids


def _identity(x):
  return x


def repeated(f: Callable, steps: int) -> Callable:
  """Returns a repeatedly applied version of f() (funcutils.py:82-88)."""
  multistep = getattr(f, 'multistep', None)

  def f_repeated(x_initial):
    if multistep is not None:
      return multistep(x_initial, steps)
    x = x_initial
    for _ in range(steps):
      x = f(x)
    return x
  f_repeated.multistep = (
      (lambda x, k: multistep(x, k * steps)) if multistep is not None else None)
  return f_repeated


def _stack(frames):
  first = frames[0]
  if isinstance(first, grids.GridVariable):
    arr = _stack([fr.array for fr in frames])
    return grids.GridVariable(arr, first.bc)
  if isinstance(first, grids.GridArray):
    return grids.GridArray(torch.stack([fr.data for fr in frames]), first.offset,
                           first.grid)
  if isinstance(first, (tuple, list)):
    return type(first)(_stack([fr[i] for fr in frames]) for i in range(len(first)))
  if isinstance(first, torch.Tensor):
    return torch.stack(frames)
  raise TypeError(f'cannot stack frames of type {type(first)}')


def trajectory(step_fn: Callable, steps: int, post_process: Callable = _identity,
               *, start_with_input: bool = False) -> Callable:
  """funcutils.py:95-126: returns (final state, stacked frames)."""

  def multistep(values):
    frames = []
    state = values
    for _ in range(steps):
      new_state = step_fn(state)
      frames.append(post_process(state if start_with_input else new_state))
      state = new_state
    return state, _stack(frames)
  return multistep
