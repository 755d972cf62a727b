"""Held-Suarez forcing on sigma levels (mirrors dinosaur/held_suarez.py:47-167)."""
from __future__ import annotations

import numpy as np
import torch

from . import primitive_equations
from . import scales
from . import time_integration

units = scales.units


class HeldSuarezForcingSigma(time_integration.ExplicitODE):
  """Rayleigh drag below sigma_b and Newtonian relaxation toward T_eq."""

  def __init__(self, coords, physics_specs, reference_temperature,
               p0=1e5 * units.pascal, sigma_b=0.7, kf=1 / (1 * units.day),
               ka=1 / (40 * units.day), ks=1 / (4 * units.day), minT=200 * units.degK,
               maxT=315 * units.degK, dTy=60 * units.degK, dThz=10 * units.degK):
    self.coords = coords
    self.physics_specs = physics_specs
    self.reference_temperature = np.asarray(reference_temperature)
    nd = physics_specs.nondimensionalize
    self.p0 = nd(p0)
    self.sigma_b = sigma_b
    self.kf, self.ka, self.ks = nd(kf), nd(ka), nd(ks)
    self.minT, self.maxT, self.dTy, self.dThz = nd(minT), nd(maxT), nd(dTy), nd(dThz)
    self.sigma = coords.vertical.centers
    _, sin_lat = coords.horizontal.nodal_mesh
    self.lat = np.arcsin(sin_lat)
    self._consts = {}

  def kv(self):
    c = self.kf * np.maximum(0, (self.sigma - self.sigma_b) / (1 - self.sigma_b))
    return c[:, np.newaxis, np.newaxis]

  def kt(self):
    cutoff = np.maximum(0, (self.sigma - self.sigma_b) / (1 - self.sigma_b))
    return self.ka + (self.ks - self.ka) * (cutoff[:, np.newaxis, np.newaxis]
                                           * np.cos(self.lat) ** 4)

  def _c(self, name, like):
    key = (name, like.device, like.dtype)
    t = self._consts.get(key)
    if t is None:
      grid = self.coords.horizontal
      host = {
          'kv_cos2': self.kv() / grid.cos_lat ** 2,
          'kt': self.kt(),
          'sigma': self.sigma[:, None, None],
          'sin2': np.sin(self.lat) ** 2,
          'cos2': np.cos(self.lat) ** 2,
          'tref': self.reference_temperature[:, None, None],
      }[name]
      t = torch.as_tensor(np.asarray(host, np.float64), device=like.device).to(like.dtype)
      self._consts[key] = t
    return t

  def equilibrium_temperature(self, nodal_surface_pressure):
    x = nodal_surface_pressure
    p_over_p0 = self._c('sigma', x) * x / self.p0
    t = p_over_p0 ** self.physics_specs.kappa * (
        self.maxT - self.dTy * self._c('sin2', x)
        - self.dThz * torch.log(p_over_p0) * self._c('cos2', x))
    return torch.clamp(t, min=self.minT)

  def explicit_terms(self, state):
    grid = self.coords.horizontal
    a = grid.radius
    zeta, delta, lnps = state.vorticity, state.divergence, state.log_surface_pressure
    psi = grid.inverse_laplacian(zeta)
    chi = grid.inverse_laplacian(delta)
    ucos_m = (grid.d_dlon(chi) - grid.cos_lat_d_dlat(psi)) / a
    vcos_m = (grid.cos_lat_d_dlat(chi) + grid.d_dlon(psi)) / a
    un, vn, tn, lsp = grid.to_nodal_batched([ucos_m, vcos_m, state.temperature_variation, lnps])
    kvc = self._c('kv_cos2', un)
    temp = self._c('tref', tn) + tn
    teq = self.equilibrium_temperature(torch.exp(lsp))
    ttend = -self._c('kt', tn) * (temp - teq)
    vu, vv, tm = grid.to_modal_batched([-kvc * un, -kvc * vn, ttend])
    return primitive_equations.State(
        vorticity=grid.curl_cos_lat((vu, vv)),
        divergence=grid.div_cos_lat((vu, vv)),
        temperature_variation=tm,
        log_surface_pressure=torch.zeros_like(lnps),
        tracers={k: torch.zeros_like(v) for k, v in state.tracers.items()},
        sim_time=None,
    )
