"""Trajectory -> dataset -> NetCDF (dinosaur_amd.xarray_utils), following the
reference's test_primitive_eq_data_to_xarray (xarray_utils_test.py:85-228):
the dimension sizes of a state, a trajectory, a batch of trajectories (modal)
and a nodal batch; reconstruction shapes; attributes; plus a NetCDF round
trip through the file (values, dimensions, attributes) and the coordinate
system rebuilt from the attributes."""
import numpy as np
import pytest

from dinosaur_amd import coordinate_systems, hybrid_coordinates, sigma_coordinates
from dinosaur_amd import primitive_equations as pe
from dinosaur_amd import spherical_harmonic as sh
from dinosaur_amd import xarray_utils as xu


def _coords(layers=4):
  return coordinate_systems.CoordinateSystem(sh.Grid.T21(),
                                             sigma_coordinates.SigmaCoordinates.equidistant(layers))


def _state(coords, lead=(), nodal=False, seed=0):
  rs = np.random.RandomState(seed)
  shape = coords.horizontal.nodal_shape if nodal else coords.horizontal.modal_shape
  K = coords.vertical.layers
  f = lambda n: rs.normal(size=lead + (n,) + shape).astype(np.float32)
  return pe.State(f(K), f(K), f(K), f(1), tracers={'test_tracer': f(K)},
                  sim_time=(np.arange(int(np.prod(lead))).reshape(lead).astype(np.float64)
                            if lead else 0.5))


def _check(ds, sizes, coords, attrs):
  assert dict(ds.sizes) == sizes
  for k, v in coords.asdict().items():
    assert k in ds.attrs
  for k, v in attrs.items():
    assert ds.attrs[k] == v


@pytest.mark.parametrize('samples,steps', [(2, 3)])
def test_primitive_eq_data_to_xarray(samples, steps):
  coords = _coords()
  g = coords.horizontal
  attrs = {'foo': 'bar'}
  base = {xu.XR_LON_MODE_NAME: 2 * g.longitude_wavenumbers - 1,
          xu.XR_LAT_MODE_NAME: g.total_wavenumbers, xu.XR_LEVEL_NAME: 4, xu.XR_SURFACE_NAME: 1}
  st = _state(coords)
  ds = xu.data_to_xarray(st.asdict(), sample_ids=None, times=None, coords=coords, attrs=attrs)
  _check(ds, base, coords, attrs)
  rec = xu.xarray_to_primitive_eq_data(ds, tracers_to_include=('test_tracer',))
  assert rec['vorticity'].shape == st.vorticity.shape
  assert rec['log_surface_pressure'].shape == st.log_surface_pressure.shape
  assert ds['log_surface_pressure'].dims == ('surface',) + xu.MODAL_AXES_NAMES
  times = 0.1 * np.arange(steps)
  traj = _state(coords, (steps,))
  ds = xu.data_to_xarray(traj.asdict(), sample_ids=None, times=times, coords=coords, attrs=attrs)
  _check(ds, {**base, xu.XR_TIME_NAME: steps}, coords, attrs)
  assert ds['vorticity'].dims == ('time', 'level') + xu.MODAL_AXES_NAMES
  assert ds['sim_time'].dims == ('time',)
  rec = xu.xarray_to_primitive_equations_with_time_data(ds, tracers_to_include=('test_tracer',))
  np.testing.assert_array_equal(rec['sim_time'], traj.sim_time)
  batch = _state(coords, (samples, steps))
  ds = xu.data_to_xarray(batch.asdict(), sample_ids=np.arange(samples), times=times,
                         coords=coords, attrs=attrs)
  _check(ds, {**base, xu.XR_TIME_NAME: steps, xu.XR_SAMPLE_NAME: samples}, coords, attrs)
  assert ds['test_tracer'].dims == ('sample', 'time', 'level') + xu.MODAL_AXES_NAMES
  nod = _state(coords, (samples, steps), nodal=True)
  ds = xu.data_to_xarray(nod.asdict(), sample_ids=np.arange(samples), times=times,
                         coords=coords, attrs=attrs)
  _check(ds, {xu.XR_LON_NAME: g.longitude_nodes, xu.XR_LAT_NAME: g.latitude_nodes,
              xu.XR_LEVEL_NAME: 4, xu.XR_SURFACE_NAME: 1, xu.XR_TIME_NAME: steps,
              xu.XR_SAMPLE_NAME: samples}, coords, attrs)
  np.testing.assert_allclose(ds['lat'].values,
                             np.degrees(np.arcsin(g.nodal_axes[1])), rtol=1e-12)


def test_name_collisions_and_bad_coords():
  coords = _coords()
  d = _state(coords).asdict()
  d['tracers'] = {'vorticity': d['vorticity']}
  with pytest.raises(ValueError, match='collide'):
    xu.data_to_xarray(d, coords=coords, times=None)
  with pytest.raises(ValueError, match='collides'):
    xu.data_to_xarray(_state(coords).asdict(), coords=coords, times=None,
                      additional_coords={'x': np.arange(4)})
  with pytest.raises(ValueError, match='not allowed'):
    xu.data_to_xarray(_state(coords).asdict(), coords=coords, times=None,
                      attrs={'radius': 2.0})
  bad = _state(coords).asdict()
  bad['vorticity'] = np.zeros((3, 5))
  with pytest.raises(ValueError, match='shape'):
    xu.data_to_xarray(bad, coords=coords, times=None)


def test_netcdf_round_trip(tmp_path):
  coords = _coords(5)
  traj = _state(coords, (3,), nodal=True, seed=4)
  d = traj.asdict()
  d['diagnostics'] = {'u': np.ones((3, 5) + coords.horizontal.nodal_shape, np.float32)}
  times = np.array([0.0, 300.0, 600.0])
  ds = xu.data_to_xarray(d, coords=coords, times=times, attrs={'note': 'synthetic'})
  path = tmp_path / 'traj.nc'
  xu.save_netcdf(ds, str(path))
  with open(path, 'rb') as f:
    assert f.read(4) == b'CDF\x02'
  back = xu.open_netcdf(str(path))
  for k in ('vorticity', 'divergence', 'temperature_variation', 'log_surface_pressure',
            'test_tracer', 'u', 'sim_time'):
    assert back[k].dims == ds[k].dims, k
    np.testing.assert_array_equal(back[k].values, ds[k].values)
  np.testing.assert_array_equal(back['time'].values, times)
  assert back.attrs['note'] == 'synthetic'
  rebuilt = xu.coordinate_system_from_attrs(back.attrs)
  assert rebuilt.horizontal.modal_shape == coords.horizontal.modal_shape
  # the attribute stores the float32 boundaries' exact values in float64
  np.testing.assert_array_equal(rebuilt.vertical.boundaries, coords.vertical.boundaries)


def test_coordinate_system_from_attrs_hybrid():
  lev = hybrid_coordinates.HybridCoordinates.analytic_levels(6)
  c = coordinate_systems.CoordinateSystem(sh.Grid.T21(), lev)
  assert xu.coordinate_system_from_attrs(c.asdict()).vertical == lev
