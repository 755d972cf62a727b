"""Pins the oracle's sigma-coordinate operators with closed forms
(sigma_coordinates_test.py:34-393 checks the same quantities)."""
import numpy as np
import pytest

from oracle import sigma


def test_equidistant_layout():
  c = sigma.SigmaCoordinates.equidistant(4)
  np.testing.assert_allclose(c.boundaries, [0, .25, .5, .75, 1])
  np.testing.assert_allclose(c.centers, [.125, .375, .625, .875])
  np.testing.assert_allclose(c.layer_thickness, [.25] * 4)
  assert c.layer_thickness.dtype == np.float32
  with pytest.raises(ValueError):
    sigma.SigmaCoordinates([0.1, 1.0])


def test_cumulative_integral_midpoint_exact_on_linear():
  c = sigma.SigmaCoordinates(np.linspace(0, 1, 11))
  x = (3 * c.centers + 1)[:, None, None] * np.ones((1, 2, 3))
  got = sigma.cumulative_sigma_integral(x, c)
  b = c.boundaries[1:]
  np.testing.assert_allclose(got[:, 0, 0], 1.5 * b ** 2 + b, atol=1e-12)
  up = sigma.cumulative_sigma_integral(x, c, downward=False)
  total = 1.5 + 1
  np.testing.assert_allclose(up[:, 0, 0], total - (1.5 * c.boundaries[:-1] ** 2 + c.boundaries[:-1]), atol=1e-12)
  np.testing.assert_allclose(sigma.sigma_integral(x, c)[0, 0, 0], total, atol=1e-12)


def test_centered_difference_exact_on_quadratic_at_midpoints():
  c = sigma.SigmaCoordinates(np.linspace(0, 1, 9))
  s = c.centers
  x = (s ** 2)[:, None, None]
  d = sigma.centered_difference(x, c)
  mid = (s[1:] + s[:-1]) / 2
  np.testing.assert_allclose(d[:, 0, 0], 2 * mid, atol=1e-12)


def test_centered_vertical_advection_formula():
  c = sigma.SigmaCoordinates(np.linspace(0, 1, 6))
  rs = np.random.RandomState(0)
  w = rs.normal(size=(4, 3, 2))
  x = rs.normal(size=(5, 3, 2))
  got = sigma.centered_vertical_advection(w, x, c)
  dx = np.diff(x, axis=0) / np.diff(c.centers)[:, None, None]
  flux = np.concatenate([np.zeros((1, 3, 2)), w * dx, np.zeros((1, 3, 2))])
  np.testing.assert_allclose(got, -0.5 * (flux[1:] + flux[:-1]), atol=1e-14)
  # convergence: -w dx/dsigma for smooth profiles as layers grow
  errs = []
  for n in (16, 32, 64):
    cc = sigma.SigmaCoordinates(np.linspace(0, 1, n + 1))
    s = cc.centers[:, None, None]
    wb = np.sin(np.pi * cc.internal_boundaries)[:, None, None]
    xx = s ** 3
    out = sigma.centered_vertical_advection(wb, xx, cc)
    exact = -np.sin(np.pi * s) * 3 * s ** 2
    errs.append(np.abs(out - exact)[1:-1].max())
  assert errs[1] < errs[0] / 3 and errs[2] < errs[1] / 3
