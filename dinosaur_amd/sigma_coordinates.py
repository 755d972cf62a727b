"""Sigma coordinates and vertical operators (mirrors
dinosaur/sigma_coordinates.py:59-405).  Host metadata is NumPy; the operators
act on torch tensors along the level axis (-3 by default)."""
from __future__ import annotations

import numpy as np
import torch


def _with_f64_math(fn, x):
  return fn(x.astype(np.float64)).astype(x.dtype)


class SigmaCoordinates:
  """Layers from the model top (sigma = 0) to the surface (sigma = 1)."""

  def __init__(self, boundaries):
    boundaries = np.asarray(boundaries)
    if not (np.isclose(boundaries[0], 0) and np.isclose(boundaries[-1], 1)):
      raise ValueError('Expected boundaries[0] = 0, boundaries[-1] = 1, '
                       f'got boundaries = {boundaries}')
    if not all(np.diff(boundaries) > 0):
      raise ValueError('Expected `boundaries` to be monotonically increasing, '
                       f'got boundaries = {boundaries}')
    self.boundaries = boundaries

  @property
  def internal_boundaries(self):
    return self.boundaries[1:-1]

  @property
  def centers(self):
    return _with_f64_math(lambda x: (x[1:] + x[:-1]) / 2, self.boundaries)

  @property
  def layer_thickness(self):
    return _with_f64_math(np.diff, self.boundaries)

  @property
  def center_to_center(self):
    return _with_f64_math(np.diff, self.centers)

  @property
  def layers(self):
    return len(self.boundaries) - 1

  @classmethod
  def equidistant(cls, layers, dtype=np.float32):
    return cls(np.linspace(0, 1, layers + 1, dtype=dtype))

  @classmethod
  def from_centers(cls, centers):
    centers = np.asarray(centers)

    def c2b(c):
      n = len(c)
      a = 0.5 * (np.eye(n) + np.eye(n, k=-1))
      return np.pad(np.linalg.solve(a, c), [(1, 0)])
    return cls(_with_f64_math(c2b, centers))

  def asdict(self):
    return {'boundaries': self.boundaries.tolist()}

  def __hash__(self):
    return hash(tuple(self.centers.tolist()))

  def __eq__(self, other):
    return isinstance(other, SigmaCoordinates) and np.array_equal(self.centers, other.centers)


def _vec(values, ndim, axis, like: torch.Tensor):
  shape = [1] * ndim
  shape[axis] = -1
  return torch.as_tensor(np.asarray(values, np.float64), device=like.device).to(
      like.dtype).reshape(shape)


def centered_difference(x, coordinates: SigmaCoordinates, axis: int = -3):
  if coordinates.layers != x.shape[axis]:
    raise ValueError('`x.shape[axis]` must be equal to `coordinates.layers`; '
                     f'got {x.shape[axis]} and {coordinates.layers}.')
  dx = torch.diff(x, dim=axis)
  return dx * _vec(1 / coordinates.center_to_center, dx.ndim, axis, x)


def cumulative_sigma_integral(x, coordinates: SigmaCoordinates, axis: int = -3,
                              downward: bool = True):
  if coordinates.layers != x.shape[axis]:
    raise ValueError('`x.shape[axis]` must be equal to `coordinates.layers`;'
                     f'got {x.shape[axis]} and {coordinates.layers}.')
  xds = x * _vec(coordinates.layer_thickness, x.ndim, axis, x)
  if downward:
    return torch.cumsum(xds, dim=axis)
  return torch.flip(torch.cumsum(torch.flip(xds, (axis,)), dim=axis), (axis,))


def sigma_integral(x, coordinates: SigmaCoordinates, axis: int = -3, keepdims: bool = True):
  if coordinates.layers != x.shape[axis]:
    raise ValueError('`x.shape[axis]` must be equal to `coordinates.layers`;'
                     f'got {x.shape[axis]} and {coordinates.layers}.')
  xds = x * _vec(coordinates.layer_thickness, x.ndim, axis, x)
  return xds.sum(dim=axis, keepdim=keepdims)


def centered_vertical_advection(w, x, coordinates: SigmaCoordinates, axis: int = -3):
  """-(w dx/dsigma) at layer centres from the two bounding faces, zero fluxes
  at the top and bottom boundaries."""
  xd = centered_difference(x, coordinates, axis)
  wx = w * xd
  pad = [0, 0] * (-axis - 1) + [1, 1]
  wxp = torch.nn.functional.pad(wx, pad)
  n = wxp.shape[axis]
  return -0.5 * (wxp.narrow(axis, 1, n - 1) + wxp.narrow(axis, 0, n - 1))
