"""Hybrid sigma-pressure coordinates and their vertical operators (mirrors
dinosaur/hybrid_coordinates.py:52-596).  Host metadata is NumPy; the
operators act on torch tensors along the level axis (-3 by default).

Level interfaces sit at pressure p = a + b * ps (a in hPa by convention, b
dimensionless), listed from the model top to the surface.
"""
from __future__ import annotations

import numpy as np
import torch

from . import sigma_coordinates


class HybridCoordinates:
  """Interfaces p = a + b ps from the model top to the surface
  (hybrid_coordinates.py:239-268)."""

  def __init__(self, a_boundaries, b_boundaries):
    self.a_boundaries = np.asarray(a_boundaries)
    self.b_boundaries = np.asarray(b_boundaries)
    if len(self.a_boundaries) != len(self.b_boundaries):
      raise ValueError('Expected `a_boundaries` and `b_boundaries` to have the same length, '
                       f'got {len(self.a_boundaries)} and {len(self.b_boundaries)}.')

  # ---- constructors
  @classmethod
  def from_coefficients(cls, a_coeffs, b_coeffs, p0: float = 1000.0) -> 'HybridCoordinates':
    """p = A p0 + B ps with dimensionless A (:269-289)."""
    return cls(np.array(a_coeffs) * p0, np.array(b_coeffs))

  @classmethod
  def analytic_levels(cls, n_levels: int, p_top: float = 0.0, p0: float = 1000.0,
                      sigma_exponent: float = 3.0,
                      stretch_exponent: float = 2.0) -> 'HybridCoordinates':
    """Power-law levels (:291-346): a master coordinate eta = s^stretch on
    s in [0, 1]; the interface pressure at ps = p0 is p_top + eta (p0 - p_top),
    of which b = eta^sigma_exponent is terrain-following and the rest is a."""
    s = np.linspace(0, 1, n_levels + 1)
    eta = s ** stretch_exponent
    b = eta ** sigma_exponent
    a = (p_top + eta * (p0 - p_top)) / p0 - b
    a[0], b[0] = p_top / p0, 0.0
    a[-1], b[-1] = 0.0, 1.0
    return cls.from_coefficients(a, b, p0)

  @classmethod
  def analytic_ecmwf_like(cls, n_levels: int = 137, p0: float = 1013.25, p_top: float = 0.0,
                          p_transition: float = 80.0, alpha_strat: float = 4.2,
                          alpha_surf: float = 0.25,
                          b_exponent: float = 1.3) -> 'HybridCoordinates':
    """Levels dense near the top and the surface (:349-411): target
    pressures p_top + s^e(s) (p0 - p_top) with the exponent e running
    linearly from alpha_strat to alpha_surf; pure pressure up to
    p_transition, below it b = tau^b_exponent with tau the normalised
    distance from the transition, and a the non-negative remainder."""
    s = np.linspace(0, 1, n_levels + 1)
    p = p_top + s ** np.linspace(alpha_strat, alpha_surf, n_levels + 1) * (p0 - p_top)
    hybrid = p > p_transition
    tau = np.where(hybrid, (p - p_transition) / (p0 - p_transition), 0.0)
    b = np.where(hybrid, tau ** b_exponent, 0.0)
    if hybrid[-1]:
      b[-1] = 1.0
    a = np.where(hybrid, np.maximum(0.0, p - b * p0), p)
    a[0], b[0] = 0.0, 0.0
    return cls(a, b)

  @classmethod
  def from_sigma_levels(cls, sigma_levels: sigma_coordinates.SigmaCoordinates):
    """Hybrid levels with a = 0: sigma coordinates (:465-471)."""
    b = np.array(sigma_levels.boundaries)
    return cls(np.zeros_like(b), b)

  # ---- metadata
  def get_eta(self, p_surface: float) -> np.ndarray:
    return self.a_centers / p_surface + self.b_centers

  @property
  def pressure_thickness(self) -> np.ndarray:
    return np.diff(self.a_boundaries)

  @property
  def sigma_thickness(self) -> np.ndarray:
    return np.diff(self.b_boundaries)

  @property
  def a_centers(self) -> np.ndarray:
    return (self.a_boundaries[1:] + self.a_boundaries[:-1]) / 2

  @property
  def b_centers(self) -> np.ndarray:
    return (self.b_boundaries[1:] + self.b_boundaries[:-1]) / 2

  @property
  def center_to_center(self) -> np.ndarray:
    return np.diff(self.b_centers)

  @property
  def layers(self) -> int:
    return len(self.a_boundaries) - 1

  def _columns(self, values, surface_pressure):
    """values broadcast against a scalar or a (1, ...) / (...) field."""
    if isinstance(surface_pressure, torch.Tensor):
      nd = sum(1 for s in surface_pressure.shape if s > 1)
      v = torch.as_tensor(np.asarray(values, np.float64), device=surface_pressure.device)
      return v.to(surface_pressure.dtype).reshape((-1,) + (1,) * nd)
    nd = sum(1 for s in np.shape(surface_pressure) if s > 1)
    return np.expand_dims(values, tuple(range(1, nd + 1)))

  def pressure_boundaries(self, surface_pressure):
    return (self._columns(self.a_boundaries, surface_pressure)
            + self._columns(self.b_boundaries, surface_pressure) * surface_pressure)

  def pressure_centers(self, surface_pressure):
    p = self.pressure_boundaries(surface_pressure)
    return (p[1:] + p[:-1]) / 2

  def layer_thickness(self, surface_pressure):
    return (self._columns(self.pressure_thickness, surface_pressure)
            + self._columns(self.sigma_thickness, surface_pressure) * surface_pressure)

  def get_sigma_boundaries(self, surface_pressure):
    return self.a_boundaries / surface_pressure + self.b_boundaries

  def get_sigma_centers(self, surface_pressure):
    s = self.get_sigma_boundaries(surface_pressure)
    return (s[1:] + s[:-1]) / 2

  def to_approx_sigma_coords(self, layers: int, surface_pressure: float = 1013.25
                             ) -> sigma_coordinates.SigmaCoordinates:
    """Sigma levels interpolated from these levels at one surface pressure,
    the end points pinned to 0 and 1 (:568-596)."""
    bounds = np.interp(np.linspace(0, 1, num=layers + 1),
                       np.linspace(0, 1, num=self.layers + 1),
                       self.get_sigma_boundaries(surface_pressure))
    bounds[0], bounds[-1] = 0.0, 1.0
    return sigma_coordinates.SigmaCoordinates(bounds)

  def asdict(self):
    return {'a_boundaries': self.a_boundaries.tolist(),
            'b_boundaries': self.b_boundaries.tolist()}

  def __hash__(self):
    return hash((tuple(self.a_boundaries.tolist()), tuple(self.b_boundaries.tolist())))

  def __eq__(self, other):
    return (isinstance(other, HybridCoordinates)
            and np.array_equal(self.a_boundaries, other.a_boundaries)
            and np.array_equal(self.b_boundaries, other.b_boundaries))


# ------------------------------------------------------------------ operators


def _check(x, coordinates, axis):
  if coordinates.layers != x.shape[axis]:
    raise ValueError('`x.shape[axis]` must be equal to `coordinates.layers`;'
                     f'got {x.shape[axis]} and {coordinates.layers}.')


def _vec(values, ndim, axis, like: torch.Tensor):
  shape = [1] * ndim
  shape[axis] = -1
  return torch.as_tensor(np.asarray(values, np.float64), device=like.device).to(
      like.dtype).reshape(shape)


def _cumsum(x, axis, downward):
  if downward:
    return torch.cumsum(x, dim=axis)
  return torch.flip(torch.cumsum(torch.flip(x, (axis,)), dim=axis), (axis,))


def cumulative_integral_over_pressure(x, surface_pressure, coordinates: HybridCoordinates,
                                      axis: int = -3, downward: bool = True):
  """Running sum of x dp, dp = da + db ps (:52-81)."""
  _check(x, coordinates, axis)
  dp = (_vec(coordinates.pressure_thickness, x.ndim, axis, x)
        + _vec(coordinates.sigma_thickness, x.ndim, axis, x) * surface_pressure)
  return _cumsum(x * dp, axis, downward)


def integral_over_pressure(x, surface_pressure, coordinates: HybridCoordinates,
                           axis: int = -3, keepdims: bool = True):
  """sum x da + ps sum x db (:84-106)."""
  _check(x, coordinates, axis)
  ia = (x * _vec(coordinates.pressure_thickness, x.ndim, axis, x)).sum(dim=axis, keepdim=keepdims)
  ib = (x * _vec(coordinates.sigma_thickness, x.ndim, axis, x)).sum(dim=axis, keepdim=keepdims)
  return ia + surface_pressure * ib


def cumulative_integral_over_sigma(x, coordinates: HybridCoordinates, axis: int = -3,
                                   downward: bool = True):
  """Running sum of x db: the sigma part only (:109-136)."""
  _check(x, coordinates, axis)
  return _cumsum(x * _vec(coordinates.sigma_thickness, x.ndim, axis, x), axis, downward)


def integral_over_sigma(x, coordinates: HybridCoordinates, axis: int = -3,
                        keepdims: bool = True):
  """sum x db (:139-155)."""
  _check(x, coordinates, axis)
  return (x * _vec(coordinates.sigma_thickness, x.ndim, axis, x)).sum(dim=axis, keepdim=keepdims)


def centered_difference(x, coordinates: HybridCoordinates, axis: int = -3):
  """(x[n+1] - x[n]) / (b_center[n+1] - b_center[n]) at internal interfaces
  (:158-197)."""
  _check(x, coordinates, axis)
  dx = torch.diff(x, dim=axis)
  return dx * _vec(1 / coordinates.center_to_center, dx.ndim, axis, x)


def centered_vertical_advection(w, x, coordinates: HybridCoordinates, axis: int = -3,
                                w_boundary_values=None, dx_deta_boundary_values=None):
  """-(w dx/deta) at layer centres, the mean of the two bounding interfaces;
  boundary values default to zero (:200-236)."""
  xd = centered_difference(x, coordinates, axis)

  def edge(values, like):
    if values is None:
      shape = list(like.shape)
      shape[axis] = 1
      z = torch.zeros(shape, dtype=like.dtype, device=like.device)
      return z, z
    return values

  wt, wb = edge(w_boundary_values, w)
  xt, xb = edge(dx_deta_boundary_values, xd)
  wx = torch.cat([wt, w, wb], dim=axis) * torch.cat([xt, xd, xb], dim=axis)
  n = wx.shape[axis]
  return -0.5 * (wx.narrow(axis, 1, n - 1) + wx.narrow(axis, 0, n - 1))
