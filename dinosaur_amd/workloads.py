"""Seeded synthetic inputs for the benchmark configurations.

Every random array is drawn on the host with numpy.random.default_rng and a
recorded seed, so the same inputs can be rebuilt anywhere:

* orography: modal coefficients N(0,1) exp(-l/20) on the triangular mask
  (seed 1), to nodal, rescaled to a 2 km maximum, nondimensionalised;
* perturbations: N(0,1) * 1e-6 on vorticity, divergence and temperature at
  1 <= l <= 10 (seed 0, or seed 100 + member for ensemble members);
* humidity: uniform 0-15 g/kg nodal field decaying with sigma^3 (seed 200 +
  member).
"""
from __future__ import annotations

import numpy as np
import torch

from . import coordinate_systems
from . import primitive_equations
from . import primitive_equations_states
from . import scales
from . import sigma_coordinates
from . import spherical_harmonic
from . import units as units_lib

GRIDS = {
    'T21': spherical_harmonic.Grid.T21,
    'T42': spherical_harmonic.Grid.T42,
    'T85': spherical_harmonic.Grid.T85,
    'TL127': spherical_harmonic.Grid.TL127,
    'TL255': spherical_harmonic.Grid.TL255,
}


def orography_modal_coefficients(grid, seed: int = 1) -> np.ndarray:
  rng = np.random.default_rng(seed)
  _, l = grid.modal_mesh
  return rng.standard_normal(grid.modal_shape) * np.exp(-l / 20) * grid.mask


def scale_orography(nodal: np.ndarray, max_height_m: float = 2000.0) -> np.ndarray:
  """Nodal field rescaled to the given maximum height, nondimensionalised."""
  nodal = np.asarray(nodal, np.float64)
  return nodal * (max_height_m / nodal.max()) / scales.RADIUS.magnitude


def perturbation_arrays(grid, layers: int, seed: int = 0, amplitude: float = 1e-6):
  rng = np.random.default_rng(seed)
  _, l = grid.modal_mesh
  sel = grid.mask & (l >= 1) & (l <= 10)
  return [rng.standard_normal((layers,) + grid.modal_shape) * amplitude * sel
          for _ in range(3)]


def humidity_nodal(grid, sigma_centers: np.ndarray, seed: int) -> np.ndarray:
  """Specific humidity, uniform 0-15 g/kg decaying with sigma^3 (kg/kg)."""
  rng = np.random.default_rng(seed)
  base = rng.uniform(0.0, 0.015, size=grid.nodal_shape)
  return base[None] * (np.asarray(sigma_centers, np.float64)[:, None, None] ** 3)


def build_primitive(name: str, layers: int, *, moist: bool = False, members: int = 1,
                    device='cuda', dtype=torch.float32, transform_precision='float32',
                    impl=None, fused_nodal: bool = True, first_member: int = 0):
  """Returns (equation, state, coords, physics_specs) for the rest-atmosphere
  recipe with seeded orography and perturbations.  With members > 1 the state
  leaves carry a leading member dimension.  `impl` replaces the spherical
  harmonics implementation (the Grid's plug-in seam).  Members are numbered
  from `first_member`, which sets their seeds, so members of several groups
  of one ensemble get distinct perturbations."""
  import functools  # pylint: disable=g-import-not-at-top
  impl = functools.partial(impl or spherical_harmonic.RealSphericalHarmonics,
                           transform_precision=transform_precision)
  grid = GRIDS[name](spherical_harmonics_impl=impl)
  vertical = sigma_coordinates.SigmaCoordinates.equidistant(layers)
  coords = coordinate_systems.CoordinateSystem(grid, vertical)
  ps = units_lib.SimUnits.from_si()
  coef = orography_modal_coefficients(grid)
  nodal_oro = grid.to_nodal(torch.as_tensor(coef, device=device, dtype=torch.float64))
  oro = scale_orography(nodal_oro.cpu().numpy())
  state_fn, aux = primitive_equations_states.isothermal_rest_atmosphere(
      coords, ps, surface_height=oro, device=device, dtype=dtype)
  base = state_fn()
  # The orography parameter is prepared in FP64 and rounded once, so runs in
  # every precision integrate the same seeded model: the rest state's ln ps
  # balances g*h to round-off, and a parameter carrying FP32 transform error
  # instead would be an imbalance (a different model), not a pipeline error.
  modal_oro = primitive_equations.truncated_modal_orography(
      torch.as_tensor(aux['orography'], device=device, dtype=torch.float64), coords).to(dtype)
  states = []
  for mem in range(first_member, first_member + members):
    seed = 0 if (members == 1 and first_member == 0) else 100 + mem
    dz, dd, dt_ = perturbation_arrays(grid, layers, seed=seed)
    tracers = {}
    if moist:
      q = humidity_nodal(grid, vertical.centers, seed=200 + mem)
      qm = grid.to_modal(torch.as_tensor(q, device=device, dtype=dtype))
      tracers = {
          'specific_humidity': qm,
          'specific_cloud_liquid_water_content': torch.zeros_like(qm),
          'specific_cloud_ice_water_content': torch.zeros_like(qm),
      }
    states.append(primitive_equations.State(
        vorticity=base.vorticity + torch.as_tensor(dz, device=device, dtype=dtype),
        divergence=base.divergence + torch.as_tensor(dd, device=device, dtype=dtype),
        temperature_variation=base.temperature_variation + torch.as_tensor(dt_, device=device, dtype=dtype),
        log_surface_pressure=base.log_surface_pressure.clone(),
        tracers=tracers))
  if members == 1:
    state = states[0]
  else:
    state = primitive_equations.State(
        vorticity=torch.stack([s.vorticity for s in states]),
        divergence=torch.stack([s.divergence for s in states]),
        temperature_variation=torch.stack([s.temperature_variation for s in states]),
        log_surface_pressure=torch.stack([s.log_surface_pressure for s in states]),
        tracers={k: torch.stack([s.tracers[k] for s in states]) for k in states[0].tracers})
  kw = {}
  if moist:
    kw = dict(humidity_key='specific_humidity',
              cloud_keys=('specific_cloud_liquid_water_content',
                          'specific_cloud_ice_water_content'))
  eq = primitive_equations.PrimitiveEquationsSigma(
      aux['reference_temperature'], modal_oro, coords, ps, fused_nodal=fused_nodal, **kw)
  return eq, state, coords, ps
