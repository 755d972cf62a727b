"""Initial states (mirrors dinosaur/primitive_equations_states.py:37-160).

Host-side closed forms in FP64 NumPy, moved to the GPU and transformed with
the package's own transforms."""
from __future__ import annotations

import numpy as np
import torch

from . import primitive_equations
from . import scales

units = scales.units


def _relative_pressure(altitude_m):
  """Barometric formula (primitive_equations_states.py:107-116)."""
  g = 9.80665
  cp = 1004.68506
  t0 = 288.16
  mm = 0.02896968
  r0 = 8.314462618
  return (1 - g * altitude_m / (cp * t0)) ** (cp * mm / r0)


def isothermal_rest_atmosphere(coords, physics_specs, tref=288.0 * units.degK,
                               p0=1e5 * units.pascal, p1=0.0 * units.pascal,
                               surface_height=None, device='cuda', dtype=torch.float32):
  """Returns (state_fn, aux_features) for a rest atmosphere at temperature tref.

  surface_height: nondimensional nodal orography (numpy, nodal shape), or a
  Quantity in meters.  p1 > 0 (the random pressure bump) is not supported.
  """
  if (p1.magnitude if isinstance(p1, scales.Quantity) else p1) != 0:
    raise ValueError('the random surface-pressure perturbation (p1 != 0) is not supported')
  grid = coords.horizontal
  lon, sin_lat = grid.nodal_mesh
  tref_nd = physics_specs.nondimensionalize(tref)
  p0_nd = physics_specs.nondimensionalize(p0)
  if surface_height is None:
    orography = np.zeros_like(sin_lat)
  else:
    orography = np.asarray(physics_specs.nondimensionalize(surface_height), np.float64)
    if orography.shape != sin_lat.shape:
      raise ValueError(f'Expected surface_height to have shape {sin_lat.shape}, '
                       f'got {orography.shape}')
  altitude_m = orography * physics_specs.scale.length

  def state_fn(rng_key=None):
    del rng_key
    sp = p0_nd * np.ones(coords.surface_nodal_shape) * _relative_pressure(altitude_m)
    zeros = torch.zeros(coords.nodal_shape, dtype=dtype, device=device)
    modal_vort = grid.to_modal(zeros)
    lnps = grid.to_modal(torch.as_tensor(np.log(sp), device=device).to(dtype))
    return primitive_equations.State(
        vorticity=modal_vort,
        divergence=torch.zeros_like(modal_vort),
        temperature_variation=torch.zeros_like(modal_vort),
        log_surface_pressure=lnps,
    )

  aux = {'orography': orography,
         'reference_temperature': np.full((coords.vertical.layers,), tref_nd)}
  return state_fn, aux
