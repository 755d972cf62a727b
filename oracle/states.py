"""Initial states (oracle restatement).

Restates dinosaur/primitive_equations_states.py:
  isothermal_rest_atmosphere  :37-160 (p1 = 0: no random bump)
  steady_state_jw             :196-386
  baroclinic_perturbation_jw  :389-501
  gaussian_scalar             :504-544
All host-side closed forms followed by Grid.to_modal.
"""
from __future__ import annotations

import numpy as np

from .pe import State


def relative_pressure(altitude_m):
  """Barometric formula (primitive_equations_states.py:107-116)."""
  g = 9.80665
  cp = 1004.68506
  t0 = 288.16
  mm = 0.02896968
  r0 = 8.314462618
  return (1 - g * altitude_m / (cp * t0)) ** (cp * mm / r0)


def isothermal_rest_atmosphere(coords, physics_specs, tref=288.0, p0=1e5,
                               surface_height_nondim=None):
  """Returns (state, nodal_orography, reference_temperatures)."""
  grid = coords.horizontal
  lon, sin_lat = grid.nodal_mesh
  lat = np.arcsin(sin_lat)
  tref_nd = physics_specs.nondim(tref, temperature=1)
  p0_nd = physics_specs.nondim(p0, length=-1, time=-2, mass=1)
  if surface_height_nondim is None:
    orography = np.zeros_like(lat)
  else:
    orography = np.asarray(surface_height_nondim, np.float64)
  altitude_m = physics_specs.dim(orography, length=1)
  sp = p0_nd * np.ones(coords.surface_nodal_shape) * relative_pressure(altitude_m)
  nodal_vort = np.zeros(coords.nodal_shape)
  modal_vort = grid.to_modal(nodal_vort)
  state = State(
      vorticity=modal_vort,
      divergence=np.zeros_like(modal_vort),
      temperature_variation=np.zeros_like(modal_vort),
      log_surface_pressure=grid.to_modal(np.log(sp)),
  )
  ref_temps = np.full((coords.vertical.layers,), tref_nd)
  return state, orography, ref_temps


def steady_state_jw(coords, physics_specs, u0=35.0, p0=1e5, t0=288.0,
                    delta_t=4.8e5, gamma=0.005, eta_tropo=0.2, eta0=0.252):
  """Returns (state, nodal_orography, reference_temperatures, geopotential)."""
  nd = physics_specs.nondim
  u0 = nd(u0, length=1, time=-1)
  t0 = nd(t0, temperature=1)
  delta_t = nd(delta_t, temperature=1)
  p0 = nd(p0, length=-1, time=-2, mass=1)
  gamma = nd(gamma, length=-1, temperature=1)
  a = physics_specs.radius
  g = physics_specs.g
  r_gas = physics_specs.R
  omega = physics_specs.angular_velocity

  def ref_t(eta):
    top = t0 * eta ** (r_gas * gamma / g)
    if eta < eta_tropo:
      return top + delta_t * (eta_tropo - eta) ** 5
    return top

  def ref_geo(eta):
    top = (t0 * g / gamma) * (1 - eta ** (r_gas * gamma / g))
    if eta < eta_tropo:
      return top - r_gas * delta_t * (
          (np.log(eta / eta_tropo) + 137 / 60) * eta_tropo ** 5
          - 5 * eta * eta_tropo ** 4 + 5 * eta ** 2 * eta_tropo ** 3
          - (10 / 3) * eta_tropo ** 2 * eta ** 3
          + (5 / 4) * eta_tropo * eta ** 4 - eta ** 5 / 5)
    return top

  def geo(lat, eta):
    en = (eta - eta0) * np.pi / 2
    return ref_geo(eta) + u0 * np.cos(en) ** 1.5 * (
        ((-2 * np.sin(lat) ** 6 * (np.cos(lat) ** 2 + 1 / 3) + 10 / 63)
         * u0 * np.cos(en) ** 1.5)
        + ((1.6 * np.cos(lat) ** 3 * (np.sin(lat) ** 2 + 2 / 3) - np.pi / 4)
           * a * omega))

  def tvar(lat, eta):
    en = (eta - eta0) * np.pi / 2
    c, s = np.cos(en), np.sin(en)
    return (0.75 * (eta * np.pi * u0 / r_gas) * s * np.sqrt(c) * (
        ((-2 * (np.cos(lat) ** 2 + 1 / 3) * np.sin(lat) ** 6 + 10 / 63)
         * 2 * u0 * c ** 1.5)
        + ((1.6 * np.cos(lat) ** 3 * (np.sin(lat) ** 2 + 2 / 3) - np.pi / 4)
           * a * omega)))

  def vort(lat, eta):
    en = (eta - eta0) * np.pi / 2
    return ((-4 * u0 / a) * np.cos(en) ** 1.5 * np.sin(lat) * np.cos(lat)
            * (2 - 5 * np.sin(lat) ** 2))

  grid = coords.horizontal
  lon, sin_lat = grid.nodal_mesh
  lat = np.arcsin(sin_lat)
  etas = coords.vertical.centers.astype(np.float64)
  nv = np.stack([vort(lat, e) for e in etas])
  nt = np.stack([tvar(lat, e) for e in etas])
  lsp = np.log(p0 * np.ones(lat.shape)[None])
  mv = grid.to_modal(nv)
  state = State(mv, np.zeros_like(mv), grid.to_modal(nt), grid.to_modal(lsp))
  orography = geo(lat, 1.0) / g
  geopotential = np.stack([geo(lat, e) for e in etas])
  ref_temps = np.stack([ref_t(e) for e in etas])
  return state, orography, ref_temps, geopotential


def baroclinic_perturbation_jw(coords, physics_specs, u_perturb=1.0,
                               lon_location=np.pi / 9, lat_location=2 * np.pi / 9,
                               perturbation_radius=0.1):
  u_p = physics_specs.nondim(u_perturb, length=1, time=-1)
  a = physics_specs.radius
  grid = coords.horizontal
  lon, sin_lat = grid.nodal_mesh
  lat = np.arcsin(sin_lat)
  x = (np.sin(lat_location) * np.sin(lat)
       + np.cos(lat_location) * np.cos(lat) * np.cos(lon - lon_location))
  r = a * np.arccos(x)
  rr = a * perturbation_radius
  sq = np.sqrt(np.maximum(1 - x ** 2, 1e-12))
  vp = (u_p / a) * np.exp(-(r / rr) ** 2) * (
      np.tan(lat) - (2 * (a / rr) ** 2 * np.arccos(x)) * (
          np.sin(lat_location) * np.cos(lat)
          - np.cos(lat_location) * np.sin(lat) * np.cos(lon - lon_location)) / sq)
  dp = ((-2 * u_p * a / rr ** 2) * np.exp(-(r / rr) ** 2) * np.arccos(x)
        * (np.cos(lat_location) * np.sin(lon - lon_location) / sq))
  k = coords.vertical.layers
  mv = grid.to_modal(np.stack([vp] * k))
  md = grid.to_modal(np.stack([dp] * k))
  return State(mv, md, np.zeros_like(mv), np.zeros_like(mv[:1]))


def gaussian_scalar(coords, physics_specs, lon_location=np.pi / 9,
                    lat_location=2 * np.pi / 9, perturbation_radius=0.2,
                    amplitude=1.0):
  a = physics_specs.radius
  grid = coords.horizontal
  lon, sin_lat = grid.nodal_mesh
  lat = np.arcsin(sin_lat)
  x = (np.sin(lat_location) * np.sin(lat)
       + np.cos(lat_location) * np.cos(lat) * np.cos(lon - lon_location))
  r = a * np.arccos(np.clip(x, -1, 1))
  rr = a * perturbation_radius
  f = amplitude * np.exp(-(r / rr) ** 2)
  return grid.to_modal(np.stack([f] * coords.vertical.layers))
