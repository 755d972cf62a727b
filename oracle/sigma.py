"""Sigma-coordinate operators (oracle restatement).

Restates dinosaur/sigma_coordinates.py:59-153 (SigmaCoordinates, equidistant
with float32 boundaries by default), :161-198 (centered_difference), :201-258
(cumulative_sigma_integral, downward cumsum of x*dsigma), :261-278
(sigma_integral), :344-405 (centered_vertical_advection with zero boundary
values), and the cumsum/diff semantics of dinosaur/jax_numpy_utils.py:32-174.

Pinned in tests/test_oracle_sigma.py by closed-form integrals of polynomials
in sigma, exactness of the midpoint rule on linear integrands, and the
second-order centred difference on quadratics.
"""
from __future__ import annotations

import numpy as np


class SigmaCoordinates:
  """Layers ordered from the model top (sigma=0) to the surface (sigma=1)."""

  def __init__(self, boundaries):
    boundaries = np.asarray(boundaries)
    if not (np.isclose(boundaries[0], 0) and np.isclose(boundaries[-1], 1)):
      raise ValueError('Expected boundaries[0] = 0, boundaries[-1] = 1')
    if not all(np.diff(boundaries) > 0):
      raise ValueError('Expected monotonically increasing boundaries')
    self.boundaries = boundaries

  @classmethod
  def equidistant(cls, layers, dtype=np.float32):
    return cls(np.linspace(0, 1, layers + 1, dtype=dtype))

  @property
  def layers(self):
    return len(self.boundaries) - 1

  def _f64(self, fn, x):
    return fn(x.astype(np.float64)).astype(x.dtype)

  @property
  def centers(self):
    return self._f64(lambda x: (x[1:] + x[:-1]) / 2, self.boundaries)

  @property
  def layer_thickness(self):
    return self._f64(np.diff, self.boundaries)

  @property
  def center_to_center(self):
    return self._f64(np.diff, self.centers)

  @property
  def internal_boundaries(self):
    return self.boundaries[1:-1]


def centered_difference(x, coords, axis=-3):
  if coords.layers != x.shape[axis]:
    raise ValueError('x.shape[axis] must equal coordinates.layers')
  dx = np.diff(x, axis=axis)
  inv = (1 / coords.center_to_center).astype(np.float64)
  shape = [1] * dx.ndim
  shape[axis] = -1
  return dx * inv.reshape(shape).astype(x.dtype)


def _bcast(v, ndim, axis, dtype):
  shape = [1] * ndim
  shape[axis] = -1
  return np.asarray(v, np.float64).reshape(shape).astype(dtype)


def cumulative_sigma_integral(x, coords, axis=-3, downward=True):
  if coords.layers != x.shape[axis]:
    raise ValueError('x.shape[axis] must equal coordinates.layers')
  xds = x * _bcast(coords.layer_thickness, x.ndim, axis, x.dtype)
  if downward:
    return np.cumsum(xds, axis=axis)
  return np.flip(np.cumsum(np.flip(xds, axis), axis), axis)


def sigma_integral(x, coords, axis=-3, keepdims=True):
  if coords.layers != x.shape[axis]:
    raise ValueError('x.shape[axis] must equal coordinates.layers')
  xds = x * _bcast(coords.layer_thickness, x.ndim, axis, x.dtype)
  return xds.sum(axis=axis, keepdims=keepdims)


def centered_vertical_advection(w, x, coords, axis=-3):
  """-(w dx/dsigma) at layer centres, averaged from the two bounding faces."""
  zshape_w = list(w.shape)
  zshape_w[axis] = 1
  wpad = np.concatenate([np.zeros(zshape_w, w.dtype), w,
                         np.zeros(zshape_w, w.dtype)], axis=axis)
  xd = centered_difference(x, coords, axis)
  zshape_x = list(xd.shape)
  zshape_x[axis] = 1
  xdpad = np.concatenate([np.zeros(zshape_x, xd.dtype), xd,
                          np.zeros(zshape_x, xd.dtype)], axis=axis)
  wx = wpad * xdpad
  n = wx.shape[axis]
  hi = np.take(wx, np.arange(1, n), axis=axis)
  lo = np.take(wx, np.arange(0, n - 1), axis=axis)
  return -0.5 * (hi + lo)
