"""Trajectories to labelled datasets and NetCDF files (mirrors the
primitive-equation parts of dinosaur/xarray_utils.py:46-57, 134-142, 195-218,
221-383, 384-476, 531-636).

`data_to_xarray` builds an ``xarray.Dataset`` when xarray is importable.
Without it, it builds this module's `Dataset`: the same data variables
(name -> (dims, values)), coordinates and attributes, written by `to_netcdf`
as a NetCDF-3 (64-bit offset) file that xarray, netCDF4 and ncdump read with
the same names, dimensions, coordinate variables and global attributes.
Device tensors are copied to the host; the nodal transforms of a trajectory
run on the GPU before that (`Grid.to_nodal`), so only the finished fields
cross PCIe.
"""
from __future__ import annotations

import dataclasses
import functools
import os
import tempfile
from typing import Any, Mapping, MutableMapping

import numpy as np

from . import coordinate_systems
from . import hybrid_coordinates
from . import primitive_equations
from . import sigma_coordinates
from . import spherical_harmonic

XR_SAMPLE_NAME = 'sample'
XR_TIME_NAME = 'time'
XR_LEVEL_NAME = 'level'
XR_SURFACE_NAME = 'surface'
XR_LON_NAME = 'lon'
XR_LAT_NAME = 'lat'
XR_LON_MODE_NAME = 'longitudinal_mode'
XR_LAT_MODE_NAME = 'total_wavenumber'
XR_REALIZATION_NAME = 'realization'

NODAL_AXES_NAMES = (XR_LON_NAME, XR_LAT_NAME)
MODAL_AXES_NAMES = (XR_LON_MODE_NAME, XR_LAT_MODE_NAME)

HORIZONTAL_COORD_TYPE_KEY = 'horizontal_grid_type'
VERTICAL_COORD_TYPE_KEY = 'vertical_grid_type'

GRID_REGISTRY = {
    'SigmaCoordinates': sigma_coordinates.SigmaCoordinates,
    'HybridCoordinates': hybrid_coordinates.HybridCoordinates,
    'Grid': spherical_harmonic.Grid,
}


def _host(x):
  """numpy array of a tensor (any device), array or scalar."""
  if hasattr(x, 'detach'):
    return x.detach().cpu().numpy()
  return np.asarray(x)


class Dataset:
  """Data variables (dims, values), coordinates and attributes: what
  `data_to_xarray` returns when xarray is not importable."""

  def __init__(self, data_vars, coords=None, attrs=None):
    self.data_vars = {k: (tuple(d), np.asarray(v)) for k, (d, v) in data_vars.items()}
    self.coords = {k: np.asarray(v) for k, v in (coords or {}).items()}
    self.attrs = dict(attrs or {})
    self.dims = {}
    for k, (d, v) in self.data_vars.items():
      if len(d) != v.ndim:
        raise ValueError(f'{k}: {len(d)} dimension names for an array of rank {v.ndim}')
      for name, n in zip(d, v.shape):
        if self.dims.setdefault(name, n) != n:
          raise ValueError(f'dimension {name} has sizes {self.dims[name]} and {n}')
    for k, v in self.coords.items():
      if v.ndim == 1:
        if self.dims.setdefault(k, v.shape[0]) != v.shape[0]:
          raise ValueError(f'coordinate {k} has {v.shape[0]} values, dimension {self.dims[k]}')

  @property
  def sizes(self):
    return dict(self.dims)

  def __contains__(self, key):
    return key in self.data_vars or key in self.coords

  def __iter__(self):
    return iter(self.data_vars)

  def __getitem__(self, key):
    if key in self.data_vars:
      return _Variable(*self.data_vars[key])
    return _Variable((key,), self.coords[key])

  def to_netcdf(self, path=None):
    """Writes a NetCDF-3 file; returns its bytes when `path` is None."""
    if path is not None:
      _write_netcdf3(self, os.fspath(path))
      return None
    with tempfile.TemporaryDirectory() as tmp:
      name = os.path.join(tmp, 'dataset.nc')
      _write_netcdf3(self, name)
      with open(name, 'rb') as f:
        return f.read()


@dataclasses.dataclass
class _Variable:
  dims: tuple
  values: np.ndarray

  @property
  def dtype(self):
    return self.values.dtype

  @property
  def shape(self):
    return self.values.shape


def _netcdf_value(v):
  v = np.asarray(v)
  if v.dtype == np.bool_:
    return v.astype(np.int8)
  if v.dtype == np.int64:
    if np.abs(v).max(initial=0) >= 2 ** 31:
      return v.astype(np.float64)
    return v.astype(np.int32)
  if v.dtype.kind == 'M':  # datetimes: seconds since the epoch
    return v.astype('datetime64[s]').astype(np.float64)
  if v.dtype.kind == 'm':
    return v.astype('timedelta64[s]').astype(np.float64)
  return v


def _write_netcdf3(ds: Dataset, target):
  from scipy.io import netcdf_file  # pylint: disable=import-outside-toplevel
  f = netcdf_file(target, 'w', version=2)
  try:
    for k, v in ds.attrs.items():
      if v is None:
        continue
      if isinstance(v, (list, tuple)):
        v = np.asarray(v, np.float64) if v else ''
      elif isinstance(v, bool):
        v = int(v)
      setattr(f, k, v)
    for name, n in ds.dims.items():
      f.createDimension(name, n)
    for name, v in ds.coords.items():
      if v.ndim != 1:
        continue
      v = _netcdf_value(v)
      var = f.createVariable(name, v.dtype, (name,))
      var[:] = v
    for name, (dims, v) in ds.data_vars.items():
      v = _netcdf_value(v)
      var = f.createVariable(name, v.dtype, dims)
      if v.ndim:
        var[...] = v
      else:
        var.assignValue(v)
  finally:
    f.close()


def save_netcdf(dataset, path: str):
  """Writes `dataset` (xarray or this module's Dataset) as NetCDF
  (xarray_utils.py:215-218)."""
  data = dataset.to_netcdf()
  with open(path, 'wb') as f:
    f.write(data)


def open_netcdf(path: str):
  """Reads a NetCDF-3 file into a Dataset (xarray.load_dataset when xarray is
  importable, xarray_utils.py:206-212)."""
  try:
    import xarray  # pylint: disable=import-outside-toplevel
    return xarray.load_dataset(path)
  except ImportError:
    pass
  from scipy.io import netcdf_file  # pylint: disable=import-outside-toplevel
  with netcdf_file(path, 'r', mmap=False) as f:
    dims = dict(f.dimensions)
    coords, data_vars = {}, {}
    for name, var in f.variables.items():
      values = np.array(var.data, copy=True)
      if var.dimensions == (name,) and name in dims:
        coords[name] = values
      else:
        data_vars[name] = (var.dimensions, values)
    attrs = {k: (v.decode() if isinstance(v, bytes) else v) for k, v in f._attributes.items()}
  return Dataset(data_vars, coords, attrs)


open_dataset = open_netcdf


def _with_realization_time_sample(shape, dims, times, sample_ids, include_realization):
  """Prepends (realization,) sample, time to non-scalar shapes
  (xarray_utils.py:221-255)."""
  not_scalar = bool(shape)
  if times is not None:
    shape = times.shape + shape
    dims = (XR_TIME_NAME,) + dims
  if sample_ids is not None:
    shape = sample_ids.shape + shape
    dims = (XR_SAMPLE_NAME,) + dims
  if not_scalar and include_realization:
    shape = (1,) + shape
    dims = (XR_REALIZATION_NAME,) + dims
  return shape, dims


def _infer_dims_shape_and_coords(coords, times, sample_ids, additional_coords):
  """All coordinates and the shape -> dimension-names table
  (xarray_utils.py:258-335): arrays are matched to dimensions by shape."""
  lon_k, lat_k = coords.horizontal.modal_axes
  lon, sin_lat = coords.horizontal.nodal_axes
  all_coords = {
      XR_LON_NAME: np.asarray(lon) * 180 / np.pi,
      XR_LAT_NAME: np.arcsin(np.asarray(sin_lat)) * 180 / np.pi,
      XR_LON_MODE_NAME: np.asarray(lon_k),
      XR_LAT_MODE_NAME: np.asarray(lat_k),
      XR_LEVEL_NAME: np.asarray(coords.vertical.centers),
      **additional_coords,
  }
  if times is not None:
    all_coords[XR_TIME_NAME] = times
  if sample_ids is not None:
    all_coords[XR_SAMPLE_NAME] = sample_ids
  modal_shape = coords.horizontal.modal_shape
  nodal_shape = coords.horizontal.nodal_shape
  layers = coords.vertical.layers
  basic = {(): (),
           (layers,) + modal_shape: (XR_LEVEL_NAME,) + MODAL_AXES_NAMES,
           (layers,) + nodal_shape: (XR_LEVEL_NAME,) + NODAL_AXES_NAMES,
           nodal_shape: NODAL_AXES_NAMES,
           modal_shape: MODAL_AXES_NAMES,
           (1,) + nodal_shape: NODAL_AXES_NAMES}
  for dim, value in additional_coords.items():
    if dim == XR_REALIZATION_NAME:
      continue
    value = np.asarray(value)
    if value.ndim != 1:
      raise ValueError('`additional_coords` must be 1d vectors, but got: '
                       f'{value.shape=} for {dim=}')
    if value.shape == (layers,):
      raise ValueError(f'`additional_coords` {dim=} has shape={value.shape} that collides '
                       f'with {XR_LEVEL_NAME=}. Since matching of axes is done using shape, '
                       'consider renaming after the fact.')
    basic[value.shape + modal_shape] = (dim,) + MODAL_AXES_NAMES
    basic[value.shape + nodal_shape] = (dim,) + NODAL_AXES_NAMES
    basic[value.shape] = (dim,)
  update = functools.partial(_with_realization_time_sample, times=times, sample_ids=sample_ids,
                             include_realization=XR_REALIZATION_NAME in additional_coords)
  shape_to_dims = {}
  for shape, dims in basic.items():
    full_shape, full_dims = update(shape, dims)
    shape_to_dims[tuple(full_shape)] = full_dims
  return all_coords, shape_to_dims


def data_to_xarray(data: dict, *, coords: coordinate_systems.CoordinateSystem, times,
                   sample_ids=None,
                   additional_coords: MutableMapping[str, Any] | None = None,
                   attrs: Mapping[str, Any] | None = None,
                   serialize_coords_to_attrs: bool = True):
  """A sample/time-referenced dataset of primitive-equation data
  (xarray_utils.py:384-476): each array (prognostic, tracer or diagnostic)
  gets the dimensions its shape maps to; coordinates serialise into the
  attributes."""
  data = {k: v for k, v in data.items()}
  prognostic = set(data) - {'tracers', 'diagnostics'}
  tracers = dict(data.get('tracers') or {})
  diagnostics = dict(data.get('diagnostics') or {})
  if not prognostic.isdisjoint(tracers):
    raise ValueError('Tracer names collide with prognostic variables',
                     f'Tracers: {set(tracers)}; prognostics: {prognostic}')
  if not prognostic.isdisjoint(diagnostics):
    raise ValueError('Diagnostic names collide with prognostic variables',
                     f'Diagnostic: {set(diagnostics)}; ', f'prognostics: {prognostic}')
  times = None if times is None else _host(times)
  sample_ids = None if sample_ids is None else _host(sample_ids)
  additional_coords = {k: _host(v) for k, v in (additional_coords or {}).items()}
  if coords.vertical.layers != 1 and XR_SURFACE_NAME not in additional_coords:
    additional_coords[XR_SURFACE_NAME] = np.ones(1)
  all_coords, shape_to_dims = _infer_dims_shape_and_coords(coords, times, sample_ids,
                                                           additional_coords)
  used = set()
  data_vars = {}
  for group, what in ((({k: data[k] for k in prognostic}), 'prognostic'),
                      (tracers, 'tracer'), (diagnostics, 'diagnostic')):
    for key in sorted(group) if what == 'prognostic' else group:
      value = _host(group[key])
      if value.shape not in shape_to_dims:
        raise ValueError(f'Value of shape {value.shape} is not in {shape_to_dims=}')
      dims = shape_to_dims[value.shape]
      data_vars[key] = (dims, value)
      used.update(dims)
  ds_attrs = coords.asdict() if serialize_coords_to_attrs else {}
  if attrs is not None:
    for key in ds_attrs:
      if key in attrs:
        raise ValueError(f'Key {key} is not allowed in `attrs`.')
    ds_attrs.update(attrs)
  coord_vals = {k: v for k, v in all_coords.items() if k in used}
  try:
    import xarray  # pylint: disable=import-outside-toplevel
  except ImportError:
    return Dataset(data_vars, coord_vals, ds_attrs)
  return xarray.Dataset(data_vars, coord_vals, attrs=ds_attrs)


def coordinate_system_from_attrs(attrs) -> coordinate_systems.CoordinateSystem:
  """The CoordinateSystem serialised into a dataset's attributes
  (xarray_utils.py:356-381)."""
  hcls = GRID_REGISTRY[attrs[HORIZONTAL_COORD_TYPE_KEY]]
  h_keys = ('longitude_wavenumbers', 'total_wavenumbers', 'longitude_nodes', 'latitude_nodes',
            'latitude_spacing', 'longitude_offset', 'radius')
  kw = {}
  for k in h_keys:
    if k in attrs:
      v = attrs[k]
      kw[k] = v.decode() if isinstance(v, bytes) else (v.item() if hasattr(v, 'item') else v)
  horizontal = hcls(**kw)
  vertical = None
  if VERTICAL_COORD_TYPE_KEY in attrs:
    vname = attrs[VERTICAL_COORD_TYPE_KEY]
    vname = vname.decode() if isinstance(vname, bytes) else vname
    vcls = GRID_REGISTRY[vname]
    if vcls is sigma_coordinates.SigmaCoordinates:
      vertical = vcls(np.asarray(attrs['boundaries']))
    else:
      vertical = vcls(np.asarray(attrs['a_boundaries']), np.asarray(attrs['b_boundaries']))
  return coordinate_systems.CoordinateSystem(horizontal, vertical)


def xarray_to_primitive_eq_data(dataset, *, values: str = 'values',
                                tracers_to_include=tuple()) -> dict:
  """Dictionary of primitive-equation State data from a dataset
  (xarray_utils.py:580-605)."""
  get = lambda k: getattr(dataset[k], values)
  return primitive_equations.State(
      vorticity=get('vorticity'), divergence=get('divergence'),
      temperature_variation=get('temperature_variation'),
      log_surface_pressure=get('log_surface_pressure'),
      tracers={k: get(k) for k in tracers_to_include}).asdict()


def xarray_to_primitive_equations_with_time_data(dataset, *, values: str = 'values',
                                                 tracers_to_include=tuple()) -> dict:
  """As xarray_to_primitive_eq_data, with sim_time (xarray_utils.py:608-634)."""
  get = lambda k: getattr(dataset[k], values)
  return primitive_equations.State(
      vorticity=get('vorticity'), divergence=get('divergence'),
      temperature_variation=get('temperature_variation'),
      log_surface_pressure=get('log_surface_pressure'), sim_time=get('sim_time'),
      tracers={k: get(k) for k in tracers_to_include}).asdict()


def nodal_trajectory_data(grid: spherical_harmonic.Grid, trajectory: primitive_equations.State,
                          include_velocity: bool = True) -> dict:
  """Modal trajectory (leaves (time, [member,] level, 2M-1, L) on the GPU) ->
  host dictionary of nodal fields, the form the notebooks pass to
  data_to_xarray: every leaf through one batched inverse transform, plus u, v
  from vorticity and divergence (spherical_harmonic.py:1198-1210)."""
  names = ['vorticity', 'divergence', 'temperature_variation', 'log_surface_pressure']
  leaves = [getattr(trajectory, n) for n in names]
  tracers = dict(trajectory.tracers)
  nodal = grid.to_nodal_batched(leaves + list(tracers.values()))
  out = {n: _host(v) for n, v in zip(names, nodal[:4])}
  out['tracers'] = {k: _host(v) for k, v in zip(tracers, nodal[4:])}
  if include_velocity:
    u, v = spherical_harmonic.vor_div_to_uv_nodal(grid, trajectory.vorticity,
                                                   trajectory.divergence)
    out['diagnostics'] = {'u': _host(u), 'v': _host(v)}
  if trajectory.sim_time is not None:
    out['sim_time'] = _host(trajectory.sim_time)
  return out


def save_trajectory_netcdf(path: str | os.PathLike, grid_coords, trajectory, times,
                           **kwargs):
  """Nodal trajectory fields of a modal State trajectory written as NetCDF."""
  data = nodal_trajectory_data(grid_coords.horizontal, trajectory)
  data.pop('sim_time', None)
  ds = data_to_xarray(data, coords=grid_coords, times=times, **kwargs)
  save_netcdf(ds, os.fspath(path))
  return ds
