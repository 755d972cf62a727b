"""Pins the oracle's transform tables and Grid operators.

Each check is independent of the oracle's own formulas: library routines
(scipy.special.lpmv, numpy.polynomial.legendre.leggauss), closed forms of
test functions and their derivatives, discrete orthonormality, and the
reference's known answers (spherical_harmonic.py:1154-1161 sqrt(4 pi);
spherical_harmonic_test.py:117-155 axes and shapes).
"""
from math import factorial

import numpy as np
import pytest
import scipy.special

from oracle import sht


def test_legendre_matches_lpmv_with_condon_shortley():
  # associated_legendre_test.py:43-56 checks proportionality only; here the
  # value AND sign are pinned by scipy's lpmv (which carries (-1)^m) times the
  # unit-L2 normalisation sqrt((2l+1)/2 (l-m)!/(l+m)!).
  x = np.linspace(-0.97, 0.97, 11)
  p = sht.legendre_table(9, 14, x)
  for m in range(9):
    for l in range(14):
      if l < m:
        assert np.all(p[m, :, l] == 0)
        continue
      c = np.sqrt((2 * l + 1) / 2 * factorial(l - m) / factorial(l + m))
      np.testing.assert_allclose(p[m, :, l], c * scipy.special.lpmv(m, l, x),
                                 rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize('n', [16, 32, 64, 128])
def test_gauss_nodes_match_numpy_leggauss(n):
  x, w = sht.gauss_nodes(n)
  xn, wn = np.polynomial.legendre.leggauss(n)
  np.testing.assert_allclose(x, xn, atol=1e-14)
  np.testing.assert_allclose(w, wn, atol=1e-14)
  assert abs(w.sum() - 2) < 1e-13
  np.testing.assert_allclose(x, -x[::-1], atol=1e-15)


@pytest.mark.parametrize('spacing', ['gauss', 'equiangular'])
def test_discrete_orthonormality(spacing):
  # associated_legendre_test.py:28-41.
  n_l = 24
  x, w = sht.LATITUDE_SPACINGS[spacing](n_l if spacing == 'gauss' else 2 * n_l)
  p = sht.legendre_table(8, n_l, x)
  for m in range(8):
    gram = np.einsum('j,jk,jl->kl', w, p[m][:, m:], p[m][:, m:])
    np.testing.assert_allclose(gram, np.eye(n_l - m), atol=1e-10)


def test_fourier_basis_closed_form():
  f = sht.real_fourier_basis(5, 16)
  xs = 2 * np.pi * np.arange(16) / 16
  np.testing.assert_allclose(f[:, 0], 1 / np.sqrt(2 * np.pi))
  for j in range(1, 5):
    np.testing.assert_allclose(f[:, 2 * j - 1], np.cos(j * xs) / np.sqrt(np.pi), atol=1e-14)
    np.testing.assert_allclose(f[:, 2 * j], np.sin(j * xs) / np.sqrt(np.pi), atol=1e-14)
  # orthonormal under the trapezoid weight 2 pi / I
  np.testing.assert_allclose(f.T @ f * (2 * np.pi / 16), np.eye(9), atol=1e-13)


def test_axes_and_shapes():
  g = sht.Grid(4, 4, 8, 4)
  np.testing.assert_array_equal(g.modal_axes[0], [0, 1, -1, 2, -2, 3, -3])
  assert sht.Grid.T21().nodal_shape == (64, 32)
  assert sht.Grid.T21().modal_shape == (43, 23)
  assert sht.Grid.TL31().modal_shape == (63, 33)
  assert sht.Grid.TL255().modal_shape == (511, 257)


def test_constant_field_kat():
  g = sht.Grid.T21()
  c = g.to_modal(np.ones(g.nodal_shape))
  assert abs(c[0, 0] - 3.5449077) < 1e-7
  assert np.abs(c.ravel()[1:]).max() < 1e-13


@pytest.mark.parametrize('ml', [(0, 0), (3, 7), (11, 20)])
def test_single_harmonic_is_unit_norm(ml):
  # spherical_harmonic_test.py:508-529: integral of Y^2 over the sphere = 1.
  g = sht.Grid.T42()
  m, l = ml
  x = np.zeros(g.modal_shape)
  x[max(2 * m - 1, 0), l] = 1
  y = g.to_nodal(x)
  assert abs(g.integrate(y ** 2) - 1) < 1e-10


def test_round_trip():
  g = sht.Grid(32, 32, 128, 64)
  rs = np.random.RandomState(0)
  _, l = g.modal_mesh
  x = rs.normal(size=g.modal_shape) * g.mask / (l + 1) ** 2
  np.testing.assert_allclose(g.to_modal(g.to_nodal(x)), x, atol=1e-13)


def _f0(lat, lon):
  return np.cos(lat) ** 4 * np.sin(3 * lon)


def _f0_dlat(lat, lon):
  return -4 * np.cos(lat) ** 3 * np.sin(lat) * np.sin(3 * lon)


def _f0_dlon(lat, lon):
  return 3 * np.cos(lat) ** 4 * np.cos(3 * lon)


def test_derivatives_against_analytic():
  # spherical_harmonic_test.py:242-287 with hand-derived derivatives; the
  # test function is not band-limited, so the reference tolerance 1e-5 holds.
  g = sht.Grid.with_wavenumbers(64)
  lon, sin_lat = g.nodal_mesh
  lat = np.arcsin(sin_lat)
  f = _f0(lat, lon)
  fm = g.to_modal(f)
  np.testing.assert_allclose(g.to_nodal(g.d_dlon(fm)), _f0_dlon(lat, lon), atol=1e-5)
  np.testing.assert_allclose(g.to_nodal(g.cos_lat_d_dlat(fm)),
                             _f0_dlat(lat, lon) * np.cos(lat), atol=1e-5)
  # sec(lat) d/dlat (cos^2 f) = sec * (-2 cos sin f + cos^2 f_lat)
  expected = (-2 * np.sin(lat) * np.cos(lat) * f + np.cos(lat) ** 2 * _f0_dlat(lat, lon)) / np.cos(lat)
  np.testing.assert_allclose(g.to_nodal(g.sec_lat_d_dlat_cos2(fm)), expected, atol=1e-5)


def test_laplacian_of_harmonic_and_inverse():
  g = sht.Grid.T21(radius=2.0)
  rs = np.random.RandomState(1)
  _, l = g.modal_mesh
  x = rs.normal(size=g.modal_shape) * g.mask
  x[0, 0] = 0
  np.testing.assert_allclose(g.inverse_laplacian(g.laplacian(x)), x, atol=1e-12)
  np.testing.assert_allclose(g.laplacian(x), -l * (l + 1) / 4.0 * x)


def test_div_curl_of_cos_lat_vector():
  # spherical_harmonic_test.py:422-490: vorticity -> stream -> velocity and
  # divergence -> potential -> velocity round trips on a dealiased grid.
  g = sht.Grid.with_wavenumbers(64)
  rs = np.random.RandomState(2)
  _, l = g.modal_mesh
  zeta = rs.normal(size=g.modal_shape) * g.mask / (l + 1) ** 2
  delta = rs.normal(size=g.modal_shape) * g.mask / (l + 1) ** 2
  zeta[:, -1] = delta[:, -1] = 0
  zeta[0, 0] = delta[0, 0] = 0
  u, v = sht.get_cos_lat_vector(zeta, delta, g, clip=False)
  sec2 = g.sec2_lat
  un = g.to_modal(g.to_nodal(u) * sec2)
  vn = g.to_modal(g.to_nodal(v) * sec2)
  div = g.div_cos_lat((un, vn))
  div[:, -1] = 0
  np.testing.assert_allclose(div, delta, atol=1e-10)
  np.testing.assert_allclose(g.curl_cos_lat((un, vn)), zeta, atol=1e-10)


def test_clip_wavenumbers_kat():
  # spherical_harmonic_test.py:531-537
  g = sht.Grid.T21()
  x = np.ones(g.modal_shape)
  y = g.clip_wavenumbers(x, n=2)
  assert np.all(y[:, -2:] == 0) and np.all(y[:, :-2] == 1)
  with pytest.raises(ValueError):
    g.clip_wavenumbers(x, n=0)


def test_vor_div_uv_closed_forms_and_round_trip():
  """Solid-body rotation u = U cos(lat), v = 0 has vorticity 2 U sin(lat) / a
  and zero divergence; a pure divergent flow u = 0, v = V cos(lat) sin(lat)
  has divergence V (1 - 3 sin^2) / a and no vorticity (spherical_harmonic.py:
  1181-1210).  Band-limited u, v round-trip through vorticity and divergence."""
  g = sht.Grid.T42(radius=2.0)
  lon, lat = np.meshgrid(*g.nodal_axes, indexing='ij')
  cos = np.sqrt(1 - lat ** 2)
  U = 3.0
  vort, div = sht.uv_nodal_to_vor_div_modal(g, U * cos, np.zeros_like(cos))
  np.testing.assert_allclose(g.to_nodal(vort), 2 * U * lat / g.radius, atol=1e-12)
  assert np.abs(div).max() < 1e-12
  V = 0.7
  vort, div = sht.uv_nodal_to_vor_div_modal(g, np.zeros_like(cos), V * cos * lat)
  np.testing.assert_allclose(g.to_nodal(div), V * (1 - 3 * lat ** 2) / g.radius, atol=1e-12)
  assert np.abs(vort).max() < 1e-12
  u, v = sht.vor_div_to_uv_nodal(g, *sht.uv_nodal_to_vor_div_modal(g, U * cos,
                                                                   V * cos * lat))
  np.testing.assert_allclose(u, U * cos, atol=1e-12)
  np.testing.assert_allclose(v, V * cos * lat, atol=1e-12)
  # a planted sign error in the curl fails the closed form
  bad = -g.to_nodal(sht.uv_nodal_to_vor_div_modal(g, U * cos, np.zeros_like(cos))[0])
  assert np.abs(bad - 2 * U * lat / g.radius).max() > 1.0


@pytest.mark.parametrize('base', [1, 8])
def test_fast_layout_restatement_matches_real_layout(base):
  """The Fast layout's own einsum restatement (zero-imaginary Fourier basis,
  unrepeated P applied to (sign, m)-unstacked rows) equals the pinned Real
  layout after the row relabelling (Fast row 0 = Real row 0, Fast row r =
  Real row r - 1 for r >= 2, row 1 zero), both directions, with and without
  padding; its longitudinal derivative equals the Real one; the m = 0' row
  stays zero under to_modal.  A swapped sin/cos stacking order fails."""
  g = sht.Grid.T21()
  fg = sht.FastGrid(g, base_shape_multiple=base)
  rs = np.random.RandomState(0)
  _, l = g.modal_mesh
  x = rs.normal(size=(3,) + g.modal_shape) * g.mask / (l + 1)
  xf = fg.from_real(x)
  assert np.abs(xf[..., ~fg.mask]).max() == 0
  z = g.to_nodal(x)
  zf = fg.to_nodal(xf)
  np.testing.assert_allclose(zf[..., :g.nodal_shape[0], :g.nodal_shape[1]], z, atol=1e-13)
  back = fg.to_modal(zf)
  np.testing.assert_allclose(fg.to_real(back), g.to_modal(z), atol=1e-13)
  assert np.abs(back[..., 1, :]).max() < 1e-15
  np.testing.assert_allclose(fg.to_real(fg.d_dlon(xf)), g.d_dlon(x), atol=1e-14)
  bad = sht.FastGrid(g, base)
  bad.f = bad.f.copy()
  bad.f[:, 2::2], bad.f[:, 3::2] = fg.f[:, 3::2], fg.f[:, 2::2]
  assert np.abs(bad.to_nodal(xf) - zf).max() > 1e-3
