"""Trajectory datasets with jax-cfd's names, coordinates and attributes
(jax_cfd/data/xarray_utils.py:42-153).

`velocity_trajectory_to_xarray` lays a rollout's frames out as jax-cfd does:
variables ``u``, ``v`` (and a leading scalar ``c`` when the trajectory has
one more component than the grid has dimensions) over dimensions
``([sample,] time, x, y)``, each velocity carrying its staggered ``offset``
attribute, with cell-centre ``x``/``y`` coordinates, a ``time`` coordinate and
integer ``sample`` ids.  xarray is not a dependency: the result is a small
`Dataset` holding NumPy arrays that writes itself as a netCDF-3 file
(64-bit offset format, the format of xarray's scipy engine) that
``xarray.open_dataset`` reads back as the reference's Dataset; `to_xarray`
converts in memory where xarray is installed.

Frames reach the host here: this module is the file writer at the end of a
rollout, not part of the time step.
"""
from __future__ import annotations

import dataclasses
import numbers
import os
from typing import Any, Dict, Mapping, Optional, Sequence, Tuple, Union

import numpy as np
import torch

from jax_cfd_amd.base import grids

XR_VELOCITY_NAMES = ('u', 'v', 'w')
XR_SCALAR_NAMES = ('c')  # a one-letter string, indexed like the reference's
XR_SPATIAL_DIMS = ('x', 'y', 'z')
XR_WAVENUMBER_DIMS = ('kx', 'ky', 'kz')
XR_SAMPLE_NAME = 'sample'
XR_TIME_NAME = 'time'
XR_OFFSET_NAME = 'offset'

XR_SAVE_GRID_SIZE_ATTR_NAME = 'save_grid_size'
XR_SAVE_GRID_SIZE_ATTR_NAME_RECTANGLE = ('save_grid_size_x', 'save_grid_size_y')

XR_DOMAIN_SIZE_NAME = 'domain_size'
XR_NDIM_ATTR_NAME = 'ndim'
XR_STABLE_TIME_STEP_ATTR_NAME = 'stable_time_step'

_NETCDF3_VERSION = 2  # 64-bit offset format


@dataclasses.dataclass
class Variable:
  """A named-dimension array with attributes (xarray.Variable's data model)."""
  dims: Tuple[str, ...]
  data: np.ndarray
  attrs: Dict[str, Any] = dataclasses.field(default_factory=dict)

  def __post_init__(self):
    self.dims = tuple(self.dims)
    self.data = np.asarray(self.data)
    if self.data.ndim != len(self.dims):
      raise ValueError(f'{len(self.dims)} dims {self.dims} for an array of '
                       f'shape {self.data.shape}')

  @property
  def shape(self):
    return self.data.shape

  @property
  def values(self) -> np.ndarray:
    return self.data

  @property
  def sizes(self) -> Dict[str, int]:
    return dict(zip(self.dims, self.data.shape))


class Dataset:
  """Data variables, dimension coordinates and global attributes.

  The subset of ``xarray.Dataset`` a trajectory file needs: ``ds['u']``,
  ``ds.data_vars``, ``ds.coords``, ``ds.attrs``, ``ds.sizes``,
  `to_netcdf` and `to_xarray`.
  """

  def __init__(self, data_vars: Mapping[str, Variable],
               coords: Optional[Mapping[str, np.ndarray]] = None,
               attrs: Optional[Mapping[str, Any]] = None):
    self.data_vars: Dict[str, Variable] = dict(data_vars)
    self.coords: Dict[str, Variable] = {}
    for name, values in (coords or {}).items():
      values = values if isinstance(values, Variable) else Variable((name,), values)
      if values.dims != (name,):
        raise ValueError(f'coordinate {name!r} must be 1-D along {name!r}')
      self.coords[name] = values
    self.attrs: Dict[str, Any] = dict(attrs or {})
    sizes: Dict[str, int] = {}
    for name, var in list(self.coords.items()) + list(self.data_vars.items()):
      for dim, size in var.sizes.items():
        if sizes.setdefault(dim, size) != size:
          raise ValueError(f'dimension {dim!r} has sizes {sizes[dim]} and {size} '
                           f'(variable {name!r})')
    self.sizes = sizes

  @property
  def dims(self) -> Dict[str, int]:
    return dict(self.sizes)

  @property
  def variables(self) -> Dict[str, Variable]:
    return {**self.coords, **self.data_vars}

  def __getitem__(self, name: str) -> Variable:
    return self.variables[name]

  def __contains__(self, name: str) -> bool:
    return name in self.variables

  def to_netcdf(self, path: Union[str, os.PathLike]) -> None:
    """Writes a netCDF-3 (64-bit offset) file: dimensions, then coordinate
    variables, then data variables, each with its attributes, and the
    global attributes."""
    from scipy.io import netcdf_file
    with netcdf_file(os.fspath(path), 'w', version=_NETCDF3_VERSION) as f:
      for dim, size in self.sizes.items():
        f.createDimension(dim, size)
      for name, var in self.variables.items():
        data = _nc3_array(var.data, f'variable {name!r}')
        out = f.createVariable(name, data.dtype, var.dims)
        for key, value in var.attrs.items():
          setattr(out, key, _nc3_attr(value, f'{name}.{key}'))
        out[...] = data
      for key, value in self.attrs.items():
        setattr(f, key, _nc3_attr(value, key))

  def to_xarray(self):
    """The same dataset as an ``xarray.Dataset`` (needs xarray)."""
    import xarray  # pylint: disable=import-outside-toplevel
    data_vars = {k: xarray.Variable(v.dims, v.data, v.attrs)
                 for k, v in self.data_vars.items()}
    coords = {k: v.data for k, v in self.coords.items()}
    return xarray.Dataset(data_vars, coords, self.attrs)

  def __repr__(self):
    dims = ', '.join(f'{k}: {v}' for k, v in self.sizes.items())
    names = ', '.join(self.data_vars)
    return f'<Dataset ({dims}) data_vars: {names}>'


def _nc3_array(values, what: str) -> np.ndarray:
  """Casts to a netCDF-3 type: int64 -> int32 (when every value fits),
  bool -> int8; other types netCDF-3 lacks raise."""
  a = np.asarray(values)
  if a.dtype == np.bool_:
    return a.astype(np.int8)
  if a.dtype.kind in 'iu' and a.dtype not in (np.int8, np.int16, np.int32):
    i32 = np.iinfo(np.int32)
    if a.size and (a.min() < i32.min or a.max() > i32.max):
      raise ValueError(f'{what}: integers outside int32 cannot be stored in netCDF-3')
    return a.astype(np.int32)
  if a.dtype in (np.float32, np.float64, np.int8, np.int16, np.int32):
    return a
  if a.dtype == np.float16:
    return a.astype(np.float32)
  raise TypeError(f'{what}: dtype {a.dtype} has no netCDF-3 type')


def _nc3_attr(value, what: str):
  """Attribute values as netCDF-3 stores them: text as text, Python floats
  as float64, Python ints as int32, sequences as 1-D arrays."""
  if isinstance(value, str):
    return value
  if isinstance(value, bytes):
    return value.decode('utf-8')
  if isinstance(value, (bool, np.bool_)):
    return np.asarray(value, np.int8)
  if isinstance(value, numbers.Integral):
    return _nc3_array(np.asarray(int(value), np.int64), what)
  if isinstance(value, numbers.Real) and not isinstance(value, np.floating):
    return np.asarray(float(value), np.float64)
  a = np.asarray(value)
  if a.dtype.kind in 'US':
    raise TypeError(f'{what}: sequences of strings have no netCDF-3 attribute type')
  if a.dtype.kind == 'f' and a.dtype != np.float32:
    a = a.astype(np.float64)
  elif a.dtype.kind in 'iub':
    a = _nc3_array(a, what)
  if a.ndim > 1:
    raise TypeError(f'{what}: attributes must be scalars or 1-D')
  return a


def _host(data) -> np.ndarray:
  if isinstance(data, (grids.GridArray, grids.GridVariable)):
    data = data.data
  if isinstance(data, torch.Tensor):
    return data.detach().cpu().numpy()
  return np.asarray(data)


def velocity_trajectory_to_xarray(
    trajectory: Sequence[Any],
    grid: Optional[grids.Grid] = None,
    time: Optional[np.ndarray] = None,
    attrs: Optional[Dict[str, Any]] = None,
    samples: bool = False,
    prefix_name: str = '',
) -> Dataset:
  """A trajectory of velocities (and an optional leading scalar) as a
  `Dataset` (xarray_utils.py:63-106).

  `trajectory` holds one stacked array per component, e.g. the frames of
  ``funcutils.trajectory``; components after the first ``len - ndim`` are
  the velocity components.  Scalars keep their own offset as the ``offset``
  attribute; velocities get the grid's face offsets when `grid` is given.
  """
  dimension = grid.ndim if grid is not None else len(trajectory)
  dims = (((XR_SAMPLE_NAME,) if samples else ()) + (XR_TIME_NAME,)
          + XR_SPATIAL_DIMS[:dimension])
  num_scalars = len(trajectory) - dimension
  data_vars = {}
  for component in range(num_scalars):
    item = trajectory[component]
    var_attrs = {}
    if isinstance(item, (grids.GridArray, grids.GridVariable)):
      var_attrs[XR_OFFSET_NAME] = tuple(item.offset)
    data_vars[prefix_name + XR_SCALAR_NAMES[component]] = Variable(
        dims, _host(item), var_attrs)
  for component in range(dimension):
    item = trajectory[component + num_scalars]
    var_attrs = {}
    if grid is not None:
      var_attrs[XR_OFFSET_NAME] = tuple(grid.cell_faces[component])
    data_vars[prefix_name + XR_VELOCITY_NAMES[component]] = Variable(
        dims, _host(item), var_attrs)
  sample_ids = None
  if samples:
    sample_ids = np.arange(next(iter(data_vars.values())).shape[0])
  coords = construct_coords(grid, time, sample_ids)
  return Dataset(data_vars, coords, attrs)


def construct_coords(
    grid: Optional[grids.Grid] = None,
    times: Optional[np.ndarray] = None,
    sample_ids: Optional[np.ndarray] = None,
) -> Mapping[str, np.ndarray]:
  """Cell-centre axes, times and sample ids by name (xarray_utils.py:109-122)."""
  coords = {}
  if grid is not None:
    for dim, axis in zip(XR_SPATIAL_DIMS, grid.axes(grid.cell_center)):
      coords[dim] = np.asarray(axis)
  if times is not None:
    coords[XR_TIME_NAME] = np.asarray(times)
  if sample_ids is not None:
    coords[XR_SAMPLE_NAME] = np.asarray(sample_ids)
  return coords


def _attr_scalar(value):
  a = np.asarray(value)
  return a.reshape(()).item() if a.size == 1 else a


def grid_from_attrs(dataset_attrs: Mapping[str, Any]) -> grids.Grid:
  """The `Grid` described by a dataset's attributes (xarray_utils.py:125-153).

  Square grids: ``ndim`` and ``save_grid_size`` with ``domain_size`` (or the
  legacy ``domain_size_multiple`` of 2π); otherwise per-axis
  ``save_grid_size_x/_y`` on the domain ``(0, aspect_ratio) × (0, 1)``.
  """
  ndim = int(_attr_scalar(dataset_attrs[XR_NDIM_ATTR_NAME]))
  if XR_SAVE_GRID_SIZE_ATTR_NAME in dataset_attrs:
    size = int(_attr_scalar(dataset_attrs[XR_SAVE_GRID_SIZE_ATTR_NAME]))
    if XR_DOMAIN_SIZE_NAME in dataset_attrs:
      length = _attr_scalar(dataset_attrs[XR_DOMAIN_SIZE_NAME])
    elif 'domain_size_multiple' in dataset_attrs:
      length = 2 * np.pi * _attr_scalar(dataset_attrs['domain_size_multiple'])
    else:
      raise ValueError(
          f'could not figure out domain size from attrs:\n{dataset_attrs}')
    return grids.Grid((size,) * ndim, domain=[(0, length)] * ndim)
  shape = tuple(int(_attr_scalar(dataset_attrs[name]))
                for name in XR_SAVE_GRID_SIZE_ATTR_NAME_RECTANGLE[:ndim])
  aspect_ratio = _attr_scalar(dataset_attrs['aspect_ratio'])
  return grids.Grid(shape, domain=((0, aspect_ratio), (0, 1)))


def _native(values):
  """A native-byte-order copy (netCDF stores big-endian)."""
  a = np.asarray(values)
  return a.astype(a.dtype.newbyteorder('='), copy=True)


def open_dataset(path: Union[str, os.PathLike]) -> Dataset:
  """Reads a netCDF-3 file (as written by `Dataset.to_netcdf`) into memory:
  1-D variables named after their own dimension become coordinates, text
  attributes come back as str, arrays in native byte order."""
  from scipy.io import netcdf_file
  with netcdf_file(os.fspath(path), 'r', mmap=False) as f:
    def attrs_of(obj_attrs):
      return {k: (v.decode('utf-8') if isinstance(v, bytes) else _native(v))
              for k, v in obj_attrs.items()}
    coords, data_vars = {}, {}
    for name, var in f.variables.items():
      v = Variable(var.dimensions, _native(var.data), attrs_of(var._attributes))
      if v.dims == (name,):
        coords[name] = v
      else:
        data_vars[name] = v
    return Dataset(data_vars, coords, attrs_of(f._attributes))
