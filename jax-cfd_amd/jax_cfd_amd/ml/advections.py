"""Advection and convection modules (jax_cfd/ml/advections.py).

modular_self_advection over FusedLearnedInterpolation is the learned
self-advection: on probe velocities it returns a description the step
factory maps onto the learned-interpolation plan; on real velocities the
GPU computes it (the explicit-terms kernel with diffusion and forcing off).
modular_advection with the fixed interpolations maps onto the base
schemes (linear / upwind / Van Leer with limiters).
"""
from __future__ import annotations

from jax_cfd_amd import _symbolic
from jax_cfd_amd.base import _fvm, advection, grids
from jax_cfd_amd.ml import interpolations


def _component(c, v):
  for i, u in enumerate(v):
    if c is u or (c.data is u.data and tuple(c.offset) == tuple(u.offset)):
      return i
  return None


def _learned(interp, c, v):
  if _symbolic.is_probe(c):
    if not isinstance(c.data, _symbolic.Affine) or c.data.coef != 1.0 or c.data.const is not None:
      raise NotImplementedError('learned self-advection advects the velocity components')
    return grids.GridArray(_symbolic.Term('learned_self_advection', comp=c.data.comp,
                                          interpolation=interp), c.offset, c.grid)
  idx = _component(c, v)
  if idx is None or len(v) != 2:
    raise NotImplementedError('learned self-advection advects the components of the velocity '
                              'its interpolation module was built on')
  if interp._plan is None:
    raise NotImplementedError('interpolation module built on probe velocities')
  key = (id(v[0].data), id(v[1].data))
  cache = getattr(interp, '_advection', None)
  if cache is None or cache[0] != key:
    conv = interp.make_plan(_fvm.spec_from_grid(interp.grid), v[0].data.device)
    interp._advection = cache = (key, conv.explicit_terms(v[0].data, v[1].data))
  return grids.GridArray(cache[1][idx], c.offset, c.grid)


def modular_self_advection(grid, dt, physics_specs, interpolation_module, **kwargs):
  """advections.py:48-71: one interpolation module, tag 'u' for the
  velocity and 'c' for the advected quantity."""
  interpolate_fn = interpolation_module(grid, dt, physics_specs, **kwargs)
  if not isinstance(interpolate_fn, interpolations.FusedLearnedInterpolation):
    raise NotImplementedError('modular_self_advection is built for FusedLearnedInterpolation')

  def advect(c, v, dt=None):
    del dt
    return _learned(interpolate_fn, c, v)

  advect.interpolation = interpolate_fn
  return advect


def modular_advection(grid, dt, physics_specs, c_interpolation_module=interpolations.upwind,
                      u_interpolation_module=interpolations.linear, **kwargs):
  """advections.py:23-45 with the fixed interpolations the kernels build."""
  c_fn = c_interpolation_module(grid, dt, physics_specs, **kwargs)
  u_fn = u_interpolation_module(grid, dt, physics_specs, **kwargs)
  schemes = {('linear', 'linear'): lambda c, v, dt: advection.advect_linear(c, v, dt),
             ('linear', 'upwind'): lambda c, v, dt: advection.advect_upwind(c, v, dt),
             ('linear', 'tvd(lax_wendroff)'): advection.advect_van_leer_using_limiters}
  if (u_fn, c_fn) not in schemes:
    raise NotImplementedError(f'modular_advection with {u_fn} / {c_fn} interpolation is not built')
  scheme = schemes[(u_fn, c_fn)]

  def advect(c, v, dt_=None):
    return scheme(c, v, dt if dt_ is None else dt_)

  return advect


def self_advection(grid, dt, physics_specs, advection_module=modular_advection, **kwargs):
  """advections.py:74-85: each velocity component advected by the velocity."""
  advect_fn = advection_module(grid, dt, physics_specs, **kwargs)

  def convect(v):
    return tuple(advect_fn(u, v, dt) for u in v)

  convect.advect = advect_fn
  return convect
