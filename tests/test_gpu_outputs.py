"""Output formats on the GPU against the oracle:

* vor_div_to_uv_nodal / uv_nodal_to_vor_div_modal (spherical_harmonic.py:
  1181-1210), FP64 1e-12 and FP32 2e-6, batched over levels and members;
* a T21 SIL3 rollout's trajectory through nodal_trajectory_data and
  data_to_xarray into a NetCDF file: the file's nodal fields equal the
  oracle's transforms of the same modal trajectory (FP64, 1e-12), with the
  reference's dimension names and sizes.
"""
import numpy as np
import pytest
import torch

from oracle import sht as osht
from dinosaur_amd import primitive_equations as pe
from dinosaur_amd import spherical_harmonic as sh
from dinosaur_amd import time_integration as ti
from dinosaur_amd import units, workloads
from dinosaur_amd import xarray_utils as xu

pytestmark = pytest.mark.gpu


def _rel(a, b):
  return np.abs(np.asarray(a) - b).max() / np.abs(b).max()


@pytest.mark.parametrize('dtype,tol', [(torch.float64, 1e-12), (torch.float32, 2e-6)])
def test_uv_vor_div_match_oracle(dtype, tol, gpu_device):
  print(f'[test] uv <-> vor/div vs oracle {dtype}', flush=True)
  g, og = sh.Grid.T42(), osht.Grid.T42()
  rs = np.random.RandomState(2)
  _, l = og.modal_mesh
  z = rs.normal(size=(2, 3) + og.modal_shape) * og.mask / (l + 1) ** 2
  d = rs.normal(size=(2, 3) + og.modal_shape) * og.mask / (l + 1) ** 2
  T = lambda a: torch.as_tensor(a, device=gpu_device).to(dtype)
  u, v = sh.vor_div_to_uv_nodal(g, T(z), T(d))
  ou, ov = osht.vor_div_to_uv_nodal(og, z, d)
  assert _rel(u.cpu().numpy(), ou) < tol and _rel(v.cpu().numpy(), ov) < tol
  z2, d2 = sh.uv_nodal_to_vor_div_modal(g, T(ou), T(ov))
  oz2, od2 = osht.uv_nodal_to_vor_div_modal(og, ou, ov)
  assert _rel(z2.cpu().numpy(), oz2) < tol and _rel(d2.cpu().numpy(), od2) < tol
  # clip=False keeps the top wavenumber
  z3, _ = sh.uv_nodal_to_vor_div_modal(g, T(ou), T(ov), clip=False)
  oz3, _ = osht.uv_nodal_to_vor_div_modal(og, ou, ov, clip=False)
  assert _rel(z3.cpu().numpy(), oz3) < tol


def test_rollout_trajectory_to_netcdf(tmp_path, gpu_device):
  print('[test] T21 L6 rollout trajectory -> NetCDF', flush=True)
  eq, st, coords, ps = workloads.build_primitive('T21', 6, moist=True, device=gpu_device,
                                                 dtype=torch.float64)
  dt = ps.nondimensionalize(1200 * units.scales.units.s)
  step = ti.step_with_filters(ti.imex_rk_sil3(eq, dt),
                              [ti.exponential_step_filter(coords.horizontal, dt)])
  traj_fn = ti.trajectory_from_step(step, outer_steps=4, inner_steps=2)
  _, traj = traj_fn(st)
  assert traj.vorticity.shape[0] == 4
  times = np.arange(4) * 2 * 1200.0
  path = tmp_path / 'rollout.nc'
  ds = xu.save_trajectory_netcdf(path, coords, traj, times, attrs={'experiment': 'T21 test'})
  back = xu.open_netcdf(str(path))
  og = osht.Grid.T21()
  assert back['temperature_variation'].dims == ('time', 'level', 'lon', 'lat')
  assert back['log_surface_pressure'].dims == ('time', 'surface', 'lon', 'lat')
  assert back.sizes['time'] == 4 and back.sizes['level'] == 6
  for name in ('vorticity', 'temperature_variation', 'log_surface_pressure'):
    ref = og.to_nodal(getattr(traj, name).cpu().numpy())
    assert _rel(back[name].values, ref) < 1e-12, name
  q = traj.tracers['specific_humidity'].cpu().numpy()
  assert _rel(back['specific_humidity'].values, og.to_nodal(q)) < 1e-12
  ou, _ = osht.vor_div_to_uv_nodal(og, traj.vorticity.cpu().numpy(),
                                   traj.divergence.cpu().numpy())
  assert _rel(back['u'].values, ou) < 1e-12
  assert back.attrs['experiment'] == 'T21 test'
  assert back.attrs['horizontal_grid_type'] == 'Grid'
  del ds
