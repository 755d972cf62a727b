"""Spherical harmonic transform and Grid operators (oracle restatement).

Restates, in NumPy, the reference's Real-layout transform exactly as the
reference formulates it (dense einsums against FP64 tables):

* ``legendre_table``      dinosaur/associated_legendre.py:23-101
* ``gauss_nodes``         dinosaur/associated_legendre.py:104-114
* ``equiangular_nodes*``  dinosaur/associated_legendre.py:117-175
* ``real_fourier_basis``  dinosaur/fourier.py:24-60
* ``Grid`` (Real layout)  dinosaur/spherical_harmonic.py:161-283 (basis,
  inverse_transform, transform) and :678-1178 (Grid constructors, operators,
  get_cos_lat_vector), plus the vor/div <-> u,v helpers :1181-1210.

The arithmetic dtype follows the input arrays (FP64 in tests; FP32 for the
CPU baseline timing), with tables cast from FP64 like the reference does.

Pinned in tests/test_oracle_sht.py by: scipy.special.lpmv with the closed-form
normalisation (value and Condon-Shortley sign), numpy.polynomial.legendre.
leggauss nodes/weights, discrete orthonormality, closed-form Fourier columns,
the sqrt(4 pi) constant-field KAT, round trips, analytic derivatives of test
functions, and Laplacian eigen-identities.
"""
from __future__ import annotations

import dataclasses
import functools
import math

import numpy as np
import scipy.special


def legendre_table(n_m: int, n_l: int, x: np.ndarray) -> np.ndarray:
  """Normalised P_l^m(x): array (n_m, len(x), n_l), zero for l < m.

  Sectoral start p[m][m] = -sqrt(1 + 1/(2m)) * sqrt(1 - x^2) * p[m-1][m-1]
  with p[0][0] = 1/sqrt(2), then the three-term recurrence in k = l - m
  (associated_legendre.py:51-67, 95-101).
  """
  if n_m > n_l:
    raise ValueError(f'Expected n_m <= n_l; got n_m = {n_m} and n_l = {n_l}.')
  x = np.asarray(x, np.float64)
  y = np.sqrt(1 - x * x)
  out = np.zeros((n_m, len(x), n_l))
  sect = np.full(len(x), 1 / np.sqrt(2))
  for m in range(n_m):
    if m > 0:
      sect = -np.sqrt(1 + 1 / (2 * m)) * y * sect
    prev2 = np.zeros_like(x)
    prev1 = sect
    out[m, :, m] = sect
    for l in range(m + 1, n_l):
      k = l - m
      mk2 = float((m + k) ** 2)
      mkp2 = float((m + k - 1) ** 2)
      m2 = float(m * m)
      a = np.sqrt((4 * mk2 - 1) / (mk2 - m2))
      b = np.sqrt((mkp2 - m2) / (4 * mkp2 - 1))
      cur = a * (x * prev1 - b * prev2)
      out[m, :, l] = cur
      prev2, prev1 = prev1, cur
  return out


def gauss_nodes(n: int):
  """Ascending Gauss-Legendre nodes/weights (associated_legendre.py:104-114)."""
  return scipy.special.roots_legendre(n)


def _weights_for(x: np.ndarray) -> np.ndarray:
  """Quadrature weights exact for P_k, k < n (associated_legendre.py:165-175)."""
  leg = legendre_table(1, x.shape[0], x)[0].T
  z = np.zeros_like(x)
  z[0] = 1
  w = np.linalg.solve(leg, z)
  return w / w.sum() * 2


def equiangular_nodes(n: int):
  """associated_legendre.py:117-141."""
  spacing = np.pi / n
  theta = np.linspace(-np.pi / 2 + spacing / 2, np.pi / 2 - spacing / 2, n)
  x = np.sin(theta)
  return x, _weights_for(x)


def equiangular_nodes_with_poles(n: int):
  """associated_legendre.py:144-162."""
  theta = np.linspace(-np.pi / 2, np.pi / 2, n)
  x = np.sin(theta)
  return x, _weights_for(x)


LATITUDE_SPACINGS = dict(
    gauss=gauss_nodes,
    equiangular=equiangular_nodes,
    equiangular_with_poles=equiangular_nodes_with_poles,
)


def real_fourier_basis(wavenumbers: int, nodes: int) -> np.ndarray:
  """F (nodes, 2*wavenumbers-1): 1/sqrt(2pi), cos(j x)/sqrt(pi), sin(j x)/sqrt(pi).

  fourier.py:24-60 with x_i = 2 pi i / nodes.
  """
  if nodes < wavenumbers:
    raise ValueError('real basis requires nodes >= wavenumbers')
  i = np.arange(nodes)[:, None]
  j = np.arange(1, wavenumbers)[None, :]
  ang = 2 * np.pi * ((i * j) % nodes) / nodes
  f = np.empty((nodes, 2 * wavenumbers - 1))
  f[:, 0] = 1 / np.sqrt(2 * np.pi)
  f[:, 1::2] = np.cos(ang) / np.sqrt(np.pi)
  f[:, 2::2] = np.sin(ang) / np.sqrt(np.pi)
  return f


def shift(x: np.ndarray, offset: int, axis: int) -> np.ndarray:
  """Zero-padded roll (jax_numpy_utils.py:158-167)."""
  out = np.zeros_like(x)
  n = x.shape[axis]
  if abs(offset) >= n:
    return out
  src = [slice(None)] * x.ndim
  dst = [slice(None)] * x.ndim
  if offset > 0:
    src[axis] = slice(0, n - offset)
    dst[axis] = slice(offset, n)
  else:
    src[axis] = slice(-offset, n)
    dst[axis] = slice(0, n + offset)
  out[tuple(dst)] = x[tuple(src)]
  return out


@dataclasses.dataclass(frozen=True)
class Grid:
  """Real-layout Grid restatement (spherical_harmonic.py:678-1151)."""
  longitude_wavenumbers: int
  total_wavenumbers: int
  longitude_nodes: int
  latitude_nodes: int
  latitude_spacing: str = 'gauss'
  longitude_offset: float = 0.0
  radius: float = 1.0

  # ---- constructors (spherical_harmonic.py:772-917)
  @classmethod
  def construct(cls, max_wavenumber, gaussian_nodes, **kw):
    return cls(max_wavenumber + 1, max_wavenumber + 2, 4 * gaussian_nodes,
               2 * gaussian_nodes, **kw)

  @classmethod
  def with_wavenumbers(cls, longitude_wavenumbers, dealiasing='quadratic', **kw):
    order = {'linear': 2, 'quadratic': 3, 'cubic': 4}[dealiasing]
    nodes = order * longitude_wavenumbers + 1
    return cls(longitude_wavenumbers, longitude_wavenumbers + 1, nodes,
               math.ceil(nodes / 2), **kw)

  @classmethod
  def T21(cls, **kw):
    return cls.construct(21, 16, **kw)

  @classmethod
  def T42(cls, **kw):
    return cls.construct(42, 32, **kw)

  @classmethod
  def T85(cls, **kw):
    return cls.construct(85, 64, **kw)

  @classmethod
  def TL31(cls, **kw):
    return cls.construct(31, 16, **kw)

  @classmethod
  def TL63(cls, **kw):
    return cls.construct(63, 32, **kw)

  @classmethod
  def TL127(cls, **kw):
    return cls.construct(127, 64, **kw)

  @classmethod
  def TL255(cls, **kw):
    return cls.construct(255, 128, **kw)

  # ---- metadata (spherical_harmonic.py:185-217, 950-1013)
  @functools.cached_property
  def _lat_nodes(self):
    return LATITUDE_SPACINGS[self.latitude_spacing](self.latitude_nodes)

  @property
  def nodal_axes(self):
    lon = np.linspace(0, 2 * np.pi, self.longitude_nodes, endpoint=False)
    return lon + self.longitude_offset, self._lat_nodes[0]

  @property
  def nodal_shape(self):
    return (self.longitude_nodes, self.latitude_nodes)

  @property
  def nodal_mesh(self):
    return np.meshgrid(*self.nodal_axes, indexing='ij')

  @property
  def modal_axes(self):
    m_pos = np.arange(1, self.longitude_wavenumbers)
    m = np.concatenate([[0], np.stack([m_pos, -m_pos], axis=1).ravel()])
    return m, np.arange(self.total_wavenumbers)

  @property
  def modal_shape(self):
    return (2 * self.longitude_wavenumbers - 1, self.total_wavenumbers)

  @property
  def modal_mesh(self):
    return np.meshgrid(*self.modal_axes, indexing='ij')

  @property
  def mask(self):
    m, l = self.modal_mesh
    return abs(m) <= l

  @property
  def cos_lat(self):
    return np.sqrt(1 - self.nodal_axes[1] ** 2)

  @property
  def sec2_lat(self):
    return 1 / (1 - self.nodal_axes[1] ** 2)

  @property
  def laplacian_eigenvalues(self):
    _, l = self.modal_axes
    return -l * (l + 1) / (self.radius ** 2)

  # ---- basis (spherical_harmonic.py:219-260)
  @functools.cached_property
  def basis(self):
    f = real_fourier_basis(self.longitude_wavenumbers, self.longitude_nodes)
    wf = 2 * np.pi / self.longitude_nodes
    x, wp = self._lat_nodes
    w = wf * wp
    p = legendre_table(self.longitude_wavenumbers, self.total_wavenumbers, x)
    p = np.repeat(p, 2, axis=0)[1:]
    return f, p, w

  def _basis_as(self, dtype):
    """FP64 tables cast to the data dtype, as the reference's einsum does."""
    cache = self.__dict__.setdefault('_cast_cache', {})
    key = np.dtype(dtype)
    if key not in cache:
      cache[key] = tuple(a.astype(key) for a in self.basis)
    return cache[key]

  # ---- transforms (spherical_harmonic.py:262-280, 1017-1031)
  # The reference's two einsums per direction, evaluated as stacked matrix
  # products (one BLAS GEMM per zonal wavenumber for the Legendre sums, one
  # for the Fourier sums over the whole batch), in the same order: Legendre
  # then Fourier for to_nodal, Fourier then Legendre for to_modal.
  def to_nodal(self, x):
    if isinstance(x, dict):
      return {k: self.to_nodal(v) for k, v in x.items()}
    if isinstance(x, (tuple, list)):
      return type(x)(self.to_nodal(v) for v in x)
    f, p, _ = self._basis_as(x.dtype)
    lead = x.shape[:-2]
    (n_r, n_l), (n_i, n_j) = self.modal_shape, self.nodal_shape
    xb = np.ascontiguousarray(np.moveaxis(x.reshape((-1, n_r, n_l)), 0, -1))  # (m, l, b)
    b = xb.shape[-1]
    px = np.matmul(p, xb)                                       # 'mjl,...ml->...mj'
    y = (f @ px.reshape(n_r, n_j * b)).reshape(n_i, n_j, b)     # 'im,...mj->...ij'
    return np.ascontiguousarray(np.moveaxis(y, -1, 0)).reshape(lead + (n_i, n_j))

  def to_modal(self, z):
    if isinstance(z, dict):
      return {k: self.to_modal(v) for k, v in z.items()}
    if isinstance(z, (tuple, list)):
      return type(z)(self.to_modal(v) for v in z)
    f, p, w = self._basis_as(z.dtype)
    lead = z.shape[:-2]
    (n_r, n_l), (n_i, n_j) = self.modal_shape, self.nodal_shape
    wz = (w * z).reshape((-1, n_i, n_j))
    b = wz.shape[0]
    wzb = np.ascontiguousarray(np.moveaxis(wz, 0, -1)).reshape(n_i, n_j * b)
    fwz = (f.T @ wzb).reshape(n_r, n_j, b)                      # 'im,...ij->...mj'
    out = np.matmul(p.transpose(0, 2, 1), fwz)                  # 'mjl,...mj->...ml'
    return np.ascontiguousarray(np.moveaxis(out, -1, 0)).reshape(lead + (n_r, n_l))

  # ---- spectral operators (spherical_harmonic.py:1033-1140)
  def laplacian(self, x):
    return x * self.laplacian_eigenvalues.astype(x.dtype)

  def inverse_laplacian(self, x):
    with np.errstate(divide='ignore', invalid='ignore'):
      inv = 1 / self.laplacian_eigenvalues
    inv[0] = 0
    inv[self.total_wavenumbers:] = 0
    return x * inv.astype(x.dtype)

  def clip_wavenumbers(self, x, n: int = 1):
    if n <= 0:
      raise ValueError(f'`n` must be >= 0; got {n}.')
    if isinstance(x, dict):
      return {k: self.clip_wavenumbers(v, n) for k, v in x.items()}
    if isinstance(x, (tuple, list)):
      return type(x)(self.clip_wavenumbers(v, n) for v in x)
    mask = np.ones(self.modal_shape[-1], x.dtype)
    mask[-n:] = 0
    return x * mask

  @functools.cached_property
  def _derivative_recurrence_weights(self):
    m, l = self.modal_mesh
    with np.errstate(divide='ignore', invalid='ignore'):
      a = np.sqrt(self.mask * (l ** 2 - m ** 2) / (4 * l ** 2 - 1))
    a[:, 0] = 0
    b = np.sqrt(self.mask * ((l + 1) ** 2 - m ** 2) / (4 * (l + 1) ** 2 - 1))
    b[:, -1] = 0
    return a, b

  def d_dlon(self, x):
    """fourier.py:63-88 along the m axis (-2)."""
    n = x.shape[-2]
    i = np.arange(n).reshape((-1, 1))
    j = (i + 1) // 2
    down = shift(x, -1, -2)
    up = shift(x, +1, -2)
    return (j * np.where(i % 2, down, -up)).astype(x.dtype)

  def cos_lat_d_dlat(self, x):
    _, l = self.modal_mesh
    a, b = self._derivative_recurrence_weights
    lm1 = shift(((l + 1) * a).astype(x.dtype) * x, -1, -1)
    lp1 = shift((-l * b).astype(x.dtype) * x, +1, -1)
    return lm1 + lp1

  def sec_lat_d_dlat_cos2(self, x):
    _, l = self.modal_mesh
    a, b = self._derivative_recurrence_weights
    lm1 = shift(((l - 1) * a).astype(x.dtype) * x, -1, -1)
    lp1 = shift((-(l + 2) * b).astype(x.dtype) * x, +1, -1)
    return lm1 + lp1

  def cos_lat_grad(self, x, clip=True):
    raw = (self.d_dlon(x) / self.radius, self.cos_lat_d_dlat(x) / self.radius)
    return self.clip_wavenumbers(raw) if clip else raw

  def k_cross(self, v):
    return (-v[1], v[0])

  def div_cos_lat(self, v, clip=True):
    raw = (self.d_dlon(v[0]) + self.sec_lat_d_dlat_cos2(v[1])) / self.radius
    return self.clip_wavenumbers(raw) if clip else raw

  def curl_cos_lat(self, v, clip=True):
    raw = (self.d_dlon(v[1]) - self.sec_lat_d_dlat_cos2(v[0])) / self.radius
    return self.clip_wavenumbers(raw) if clip else raw

  def integrate(self, z):
    w = self.basis[2] * self.radius ** 2
    return np.einsum('y,...xy->...', w, z)


def get_cos_lat_vector(vorticity, divergence, grid: Grid, clip=True):
  """spherical_harmonic.py:1164-1178."""
  psi = grid.inverse_laplacian(vorticity)
  chi = grid.inverse_laplacian(divergence)
  a = grid.cos_lat_grad(chi, clip=clip)
  b = grid.k_cross(grid.cos_lat_grad(psi, clip=clip))
  return (a[0] + b[0], a[1] + b[1])


def uv_nodal_to_vor_div_modal(grid, u_nodal, v_nodal, clip=True):
  """spherical_harmonic.py:1181-1195."""
  u = grid.to_modal(u_nodal / grid.cos_lat)
  v = grid.to_modal(v_nodal / grid.cos_lat)
  return (grid.curl_cos_lat((u, v), clip=clip),
          grid.div_cos_lat((u, v), clip=clip))


def vor_div_to_uv_nodal(grid, vorticity, divergence, clip=True):
  """spherical_harmonic.py:1198-1210."""
  u, v = get_cos_lat_vector(vorticity, divergence, grid, clip=clip)
  return grid.to_nodal(u) / grid.cos_lat, grid.to_nodal(v) / grid.cos_lat


# ---------------------------------------------------------------- fast layout


def real_fourier_basis_with_zero_imag(wavenumbers: int, nodes: int) -> np.ndarray:
  """(I, 2M) real basis with a zero column for the m = 0 sine
  (fourier.py:91-109): f[:, 0] = 1/sqrt(2 pi), f[:, 1] = 0, then cos(m lon)
  and sin(m lon) / sqrt(pi) alternating."""
  if nodes < wavenumbers:
    raise ValueError('`real_basis` requires nodes >= wavenumbers')
  lon = 2 * np.pi * np.arange(nodes) / nodes
  m = np.arange(1, wavenumbers)
  f = np.empty((nodes, 2 * wavenumbers))
  f[:, 0] = 1 / np.sqrt(2 * np.pi)
  f[:, 1] = 0
  f[:, 2::2] = np.cos(np.outer(lon, m)) / np.sqrt(np.pi)
  f[:, 3::2] = np.sin(np.outer(lon, m)) / np.sqrt(np.pi)
  return f


def _round_up(x, m):
  return -(-x // m) * m


class FastGrid:
  """FastSphericalHarmonics (spherical_harmonic.py:411-618) restated with its
  own einsums: modal (2M, L) rows [m=0, m=0' (always zero), 1 cos, 1 sin,
  ...], the P table unrepeated (M, J, L) applied to the (sign, m) unstacked
  rows, optional padding to `base_shape_multiple`."""

  def __init__(self, grid: Grid, base_shape_multiple: int = 1):
    self.grid = grid
    self.M, self.L = grid.longitude_wavenumbers, grid.total_wavenumbers
    self.I, self.J = grid.longitude_nodes, grid.latitude_nodes
    b = base_shape_multiple
    self.modal_shape = (_round_up(2 * self.M, 2 * b), _round_up(self.L, b))
    self.nodal_shape = (_round_up(self.I, b), _round_up(self.J, b))
    pmx = self.modal_shape[0] - 2 * self.M
    pmy = self.modal_shape[1] - self.L
    pnx = self.nodal_shape[0] - self.I
    pny = self.nodal_shape[1] - self.J
    f = real_fourier_basis_with_zero_imag(self.M, self.I)
    self.f = np.pad(f, [(0, pnx), (0, pmx)])
    x, wp = grid._lat_nodes
    self.w = np.pad((2 * np.pi / self.I) * wp, [(0, pny)])
    p = legendre_table(self.M, self.L, x)
    self.p = np.pad(p, [(0, pmx // 2), (0, pny), (0, pmy)])

  @property
  def modal_axes(self):
    m_pos = np.arange(1, self.M)
    lon_k = np.concatenate([[0, 0], np.stack([m_pos, -m_pos], axis=1).ravel()])
    lon_k = np.pad(lon_k, [(0, self.modal_shape[0] - 2 * self.M)])
    return lon_k, np.pad(np.arange(self.L), [(0, self.modal_shape[1] - self.L)])

  @property
  def mask(self):
    m, l = np.meshgrid(*self.modal_axes, indexing='ij')
    i, j = np.meshgrid(*(np.arange(s) for s in self.modal_shape), indexing='ij')
    return (abs(m) <= l) & (i != 1) & (i < 2 * self.M) & (j < self.L)

  @staticmethod
  def _unstack(x):
    return np.reshape(x, x.shape[:-2] + (2, x.shape[-2] // 2) + x.shape[-1:], order='F')

  @staticmethod
  def _stack(x):
    return np.reshape(x, x.shape[:-3] + (-1,) + x.shape[-1:], order='F')

  def to_nodal(self, x):
    x = self._unstack(x)
    x = np.einsum('mjl,...sml->...smj', self.p, x)
    x = self._stack(x)
    return np.einsum('im,...mj->...ij', self.f, x)

  def to_modal(self, z):
    x = np.einsum('im,...ij->...mj', self.f, self.w * z)
    x = self._unstack(x)
    x = np.einsum('mjl,...smj->...sml', self.p, x)
    return self._stack(x)

  def d_dlon(self, x):
    """j * (u_down for even rows, -u_up for odd rows), j = row // 2
    (fourier.py:112-126)."""
    i = np.arange(x.shape[-2])[:, None]
    j = i // 2
    down = np.concatenate([x[..., 1:, :], np.zeros_like(x[..., :1, :])], axis=-2)
    up = np.concatenate([np.zeros_like(x[..., :1, :]), x[..., :-1, :]], axis=-2)
    return j * np.where((i + 1) % 2, down, -up)

  def from_real(self, x):
    """Real-layout (2M-1, L) coefficients in this layout."""
    out = np.zeros(x.shape[:-2] + self.modal_shape)
    out[..., 0, :self.L] = x[..., 0, :]
    out[..., 2:2 * self.M, :self.L] = x[..., 1:, :]
    return out

  def to_real(self, x):
    out = np.concatenate([x[..., :1, :self.L], x[..., 2:2 * self.M, :self.L]], axis=-2)
    return out
