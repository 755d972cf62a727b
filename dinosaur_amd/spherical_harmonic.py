"""Spherical harmonic transforms and spectral operators on MI355X.

Drop-in for the Real-layout path of dinosaur/spherical_harmonic.py:
``RealSphericalHarmonics`` (:161-283) and ``Grid`` (:678-1151) with the same
constructors, metadata, transforms and operators, plus ``get_cos_lat_vector``
and the u,v <-> vorticity,divergence helpers (:1164-1210).

Arrays are torch tensors on the GPU (FP32 or FP64).  Other array types that
speak DLPack (NumPy, JAX, CuPy ...) are accepted and returned as the same type
where possible.  The transforms run the gfx950 kernels of libdinosaur_amd.so:
a batched real FFT along longitude and the folded triangular Legendre GEMMs,
every leaf of a batch going through one launch per stage.  Gradients flow
through hand-written adjoint kernels (``torch.autograd.Function``).
"""
from __future__ import annotations

import ctypes
import dataclasses
import functools
import math
import threading
from typing import Any, Callable

import numpy as np
import torch

from . import _lib

Array = Any

# ----------------------------------------------------------------- nodes


def gauss_legendre_nodes(n: int):
  """FP64 Gauss-Legendre nodes/weights (associated_legendre.py:104-114)."""
  return _lib.gauss_nodes_host(n)


def _weights_for_nodes(x: np.ndarray) -> np.ndarray:
  """Weights integrating P_k exactly for k < n (associated_legendre.py:165-175).

  The Legendre matrix comes from the same recurrence as the GPU table, on the
  host in FP64, for the one-time equiangular setup.
  """
  n = x.shape[0]
  leg = np.zeros((n, n))
  p_prev = np.zeros_like(x)
  p = np.full_like(x, 1 / np.sqrt(2))
  leg[0] = p
  for l in range(1, n):
    a = np.sqrt((4 * l * l - 1) / (l * l))
    b = np.sqrt(((l - 1) ** 2) / (4 * (l - 1) ** 2 - 1)) if l > 1 else 0.0
    p, p_prev = a * (x * p - b * p_prev), p
    leg[l] = p
  z = np.zeros_like(x)
  z[0] = 1
  w = np.linalg.solve(leg, z)
  return w / w.sum() * 2


def _symmetric(x: np.ndarray) -> np.ndarray:
  return 0.5 * (x - x[::-1])


def equiangular_nodes(n: int):
  spacing = np.pi / n
  theta = np.linspace(-np.pi / 2 + spacing / 2, np.pi / 2 - spacing / 2, n)
  x = _symmetric(np.sin(theta))
  return x, _weights_for_nodes(x)


def equiangular_nodes_with_poles(n: int):
  theta = np.linspace(-np.pi / 2, np.pi / 2, n)
  x = _symmetric(np.sin(theta))
  return x, _weights_for_nodes(x)


LATITUDE_SPACINGS = dict(
    gauss=gauss_legendre_nodes,
    equiangular=equiangular_nodes,
    equiangular_with_poles=equiangular_nodes_with_poles,
)


@functools.lru_cache(maxsize=64)
def get_latitude_nodes(n: int, spacing: str):
  get_nodes = LATITUDE_SPACINGS.get(spacing)
  if get_nodes is None:
    raise ValueError(f'Unknown spacing: {spacing} available spacings are '
                     f'{list(LATITUDE_SPACINGS.keys())}')
  return get_nodes(n)


# ----------------------------------------------------------------- plans


class Plan:
  """Owns a dino_plan (device tables) for one grid on one device."""

  def __init__(self, M: int, L: int, I: int, J: int, spacing: str, device: torch.device):
    self.M, self.L, self.I, self.J = M, L, I, J
    self.R = 2 * M - 1
    self.device = device
    lib = _lib.lib()
    handle = ctypes.c_void_p()
    with torch.cuda.device(device):
      if spacing == 'gauss':
        rc = lib.dino_plan_create(M, L, I, J, None, None, ctypes.byref(handle))
      else:
        x, w = get_latitude_nodes(J, spacing)
        x = np.ascontiguousarray(x, np.float64)
        w = np.ascontiguousarray(w, np.float64)
        rc = lib.dino_plan_create(M, L, I, J, _lib.dptr(x), _lib.dptr(w), ctypes.byref(handle))
    _lib.check(rc, 'dino_plan_create')
    self.handle = handle
    # autotuned schedules: (batch bucket, dtype code, precision) -> {stage: schedule}
    self.tuned = {}

  def __del__(self):
    try:
      h = getattr(self, 'handle', None)
      if h is not None and h.value:
        _lib.release('dino_plan_destroy', h)
    except Exception:  # pylint: disable=broad-except  (interpreter shutdown)
      pass

  def nodes(self):
    x = np.empty(self.J)
    w = np.empty(self.J)
    _lib.check(_lib.lib().dino_plan_nodes(self.handle, _lib.dptr(x), _lib.dptr(w)))
    return x, w

  def set_schedule(self, stage: int, schedule: int):
    """Kernel family (+ flags) for one transform stage (_lib.STAGE_*, SCHED_*)."""
    _lib.check(_lib.lib().dino_plan_set_schedule(self.handle, stage, schedule),
               'dino_plan_set_schedule')

  def schedule(self, stage: int) -> int:
    return _lib.lib().dino_plan_get_schedule(self.handle, stage)

  def apply_tuned(self, batch: int, dtype_code: int, precision: int):
    """Sets the autotuned schedules for this batch size, if any were tuned:
    the entry of the same power-of-two batch bucket, else the nearest one."""
    if not self.tuned:
      return
    from . import autotune  # pylint: disable=g-import-not-at-top
    entry = autotune.lookup(self.tuned, batch, dtype_code, precision)
    if entry is not None:
      for stage, sched in entry.items():
        if _lib.lib().dino_plan_get_schedule(self.handle, stage) != sched:
          self.set_schedule(stage, sched)

  def folded_table(self) -> np.ndarray:
    n = _lib.lib().dino_plan_table_size(self.handle)
    out = np.empty(n)
    _lib.check(_lib.lib().dino_plan_table(self.handle, _lib.dptr(out)))
    return out


_plans: dict = {}
_plans_lock = threading.Lock()


def get_plan(M, L, I, J, spacing, device) -> Plan:
  device = torch.device(device)
  if device.type != 'cuda':
    raise ValueError(f'dinosaur_amd transforms run on the GPU; got a tensor on {device}')
  if device.index is None:
    device = torch.device('cuda', torch.cuda.current_device())
  key = (M, L, I, J, spacing, device.index)
  with _plans_lock:
    plan = _plans.get(key)
    if plan is None:
      plan = Plan(M, L, I, J, spacing, device)
      _plans[key] = plan
  return plan


def _stream(device) -> ctypes.c_void_p:
  return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _dtype_code(t: torch.Tensor) -> int:
  if t.dtype == torch.float32:
    return _lib.DINO_F32
  if t.dtype == torch.float64:
    return _lib.DINO_F64
  raise TypeError(f'transforms take float32 or float64 tensors, got {t.dtype}')


def _ptr(t: torch.Tensor) -> ctypes.c_void_p:
  return ctypes.c_void_p(t.data_ptr())


def _workspace(plan: Plan, batch: int, t: torch.Tensor) -> torch.Tensor:
  return torch.empty((batch, plan.R, plan.J), dtype=t.dtype, device=t.device)


def sht_call(fn_name: str, plan: Plan, x: torch.Tensor, out_tail: tuple, precision: int):
  """Runs one batched transform entry point on a (batch, ...) tensor."""
  x = x.contiguous()
  batch = x.shape[0]
  out = torch.empty((batch,) + out_tail, dtype=x.dtype, device=x.device)
  if batch == 0:
    return out
  work = _workspace(plan, batch, x)
  flags = precision & _lib.SHT_TRIANGLE_ONLY
  precision &= ~_lib.SHT_TRIANGLE_ONLY
  if x.dtype == torch.float64:
    precision = _lib.PREC_NATIVE  # FP64 data always runs FP64 arithmetic
  plan.apply_tuned(batch, _dtype_code(x), precision)
  precision |= flags
  fn = getattr(_lib.lib(), fn_name)
  with torch.cuda.device(x.device):
    rc = fn(plan.handle, _ptr(x), _ptr(out), _ptr(work), batch, _dtype_code(x), precision,
            _stream(x.device))
  _lib.check(rc, fn_name)
  if _lib._active is not None:
    _tally_transform(fn_name, plan, batch, x)
  return out


def _tally_transform(fn_name: str, plan: Plan, batch: int, x: torch.Tensor):
  """Algorithmic work of the two stages of one transform call: the Fourier
  rows (batch, 2M-1, J) are written once and read once; modal arrays count
  their triangular entries; the folded Legendre sums are J flops per modal
  entry (Jh multiply-adds on each of the two fold parities)."""
  es = x.element_size()
  n_real = _lib.triangle_entries(plan.R, plan.L)
  fourier, nodal = plan.R * plan.J, plan.I * plan.J
  roof = 'f64' if x.dtype == torch.float64 else 'f32_mfma'
  flops = batch * plan.J * n_real
  if fn_name in ('dino_sht_inverse', 'dino_sht_forward_adjoint'):
    _lib.tally('legendre_inverse', es * batch * (n_real + fourier), flops, roof)
    _lib.tally('dft_inverse', es * batch * (fourier + nodal))
  else:
    _lib.tally('dft_forward', es * batch * (nodal + fourier))
    _lib.tally('legendre_forward', es * batch * (fourier + n_real), flops, roof)


class _ToNodal(torch.autograd.Function):
  """modal (B, 2M-1, L) -> nodal (B, I, J); backward = P^T F^T.  partial: the
  modal cotangent is consumed only by kernels that read the triangle l >= m,
  so it is written triangle-only (DINO_SHT_TRIANGLE_ONLY)."""

  @staticmethod
  def forward(ctx, x, plan, precision, partial):
    ctx.plan, ctx.precision, ctx.partial = plan, precision, partial
    return sht_call('dino_sht_inverse', plan, x, (plan.I, plan.J), precision)

  @staticmethod
  def backward(ctx, g):
    plan = ctx.plan
    prec = ctx.precision | (_lib.SHT_TRIANGLE_ONLY if ctx.partial else 0)
    return (sht_call('dino_sht_inverse_adjoint', plan, g, (plan.R, plan.L), prec),
            None, None, None)


class _ToModal(torch.autograd.Function):
  """nodal (B, I, J) -> modal (B, 2M-1, L); backward = w F P.  partial: the
  modal output is consumed only by kernels that read the triangle l >= m, so
  it is written triangle-only."""

  @staticmethod
  def forward(ctx, z, plan, precision, partial):
    ctx.plan, ctx.precision = plan, precision
    if _lib.skip_values():  # linear: the backward reads nothing saved
      return torch.empty((z.shape[0], plan.R, plan.L), dtype=z.dtype, device=z.device)
    prec = precision | (_lib.SHT_TRIANGLE_ONLY if partial else 0)
    return sht_call('dino_sht_forward', plan, z, (plan.R, plan.L), prec)

  @staticmethod
  def backward(ctx, g):
    plan = ctx.plan
    return (sht_call('dino_sht_forward_adjoint', plan, g, (plan.I, plan.J), ctx.precision),
            None, None, None)


# ----------------------------------------------------------------- array interop


def _as_tensor(x):
  """Returns (tensor on GPU, restore_fn) for torch/NumPy/DLPack inputs."""
  if isinstance(x, torch.Tensor):
    if x.device.type != 'cuda':
      raise ValueError('dinosaur_amd runs on the GPU: move the tensor to cuda first '
                       f'(got {x.device})')
    return x, lambda t: t
  if isinstance(x, np.ndarray) or np.isscalar(x):
    arr = np.asarray(x)
    if arr.dtype not in (np.float32, np.float64):
      arr = arr.astype(np.float32)
    t = torch.from_numpy(np.ascontiguousarray(arr)).cuda()
    return t, lambda r: r.detach().cpu().numpy()
  if hasattr(x, '__dlpack__'):
    t = torch.from_dlpack(x)
    if t.device.type != 'cuda':
      t = t.cuda()
    mod = type(x).__module__.split('.')[0]

    def restore(r):
      if mod == 'jax':
        import jax.dlpack  # pylint: disable=g-import-not-at-top
        return jax.dlpack.from_dlpack(r)
      return r
    return t, restore
  raise TypeError(f'unsupported array type {type(x)}')


def tree_map_over_nonscalars(fn, x):
  """Maps fn over array leaves, passing scalars through (pytree_utils.py:86-98)."""
  if isinstance(x, dict):
    return {k: tree_map_over_nonscalars(fn, v) for k, v in x.items()}
  if isinstance(x, (tuple, list)):
    return type(x)(tree_map_over_nonscalars(fn, v) for v in x)
  if x is None:
    return None
  if hasattr(x, 'tree_map_arrays'):
    return x.tree_map_arrays(fn)
  if isinstance(x, (int, float)) or (hasattr(x, 'ndim') and x.ndim == 0):
    return x
  return fn(x)


# ----------------------------------------------------------------- impls


@dataclasses.dataclass(frozen=True)
class _SphericalHarmonicBasis:
  f: np.ndarray
  p: np.ndarray
  w: np.ndarray


@dataclasses.dataclass(frozen=True)
class SphericalHarmonics:
  """Plug-in base (spherical_harmonic.py:76-158)."""
  longitude_wavenumbers: int = 0
  total_wavenumbers: int = 0
  longitude_nodes: int = 0
  latitude_nodes: int = 0
  latitude_spacing: str = 'gauss'


@dataclasses.dataclass(frozen=True)
class RealSphericalHarmonics(SphericalHarmonics):
  """Real layout m = [0, +1, -1, ...] x l (spherical_harmonic.py:161-283).

  ``transform_precision`` picks the Legendre arithmetic for FP32 data:
  'float32' (FP32 operands on MFMA) or 'bf16x3' (split-BF16 precision on
  MFMA with FP32 accumulation).  FP64 data always uses FP64 arithmetic.
  """
  transform_precision: str = 'float32'

  def __post_init__(self):
    if self.transform_precision not in _lib.PRECISIONS:
      raise ValueError(f'unknown transform_precision {self.transform_precision!r}')

  @functools.cached_property
  def nodal_axes(self):
    lon = np.linspace(0, 2 * np.pi, self.longitude_nodes, endpoint=False)
    sin_lat, _ = get_latitude_nodes(self.latitude_nodes, self.latitude_spacing)
    return lon, sin_lat

  @functools.cached_property
  def nodal_shape(self):
    return (self.longitude_nodes, self.latitude_nodes)

  @functools.cached_property
  def nodal_padding(self):
    return (0, 0)

  @functools.cached_property
  def modal_axes(self):
    m_pos = np.arange(1, self.longitude_wavenumbers)
    m_pos_neg = np.stack([m_pos, -m_pos], axis=1).ravel()
    return np.concatenate([[0], m_pos_neg]), np.arange(self.total_wavenumbers)

  @functools.cached_property
  def modal_shape(self):
    return (2 * self.longitude_wavenumbers - 1, self.total_wavenumbers)

  @functools.cached_property
  def modal_padding(self):
    return (0, 0)

  @functools.cached_property
  def modal_dtype(self):
    return np.dtype(np.float32)

  @functools.cached_property
  def mask(self):
    m, l = np.meshgrid(*self.modal_axes, indexing='ij')
    return abs(m) <= l

  def plan(self, device) -> Plan:
    return get_plan(self.longitude_wavenumbers, self.total_wavenumbers, self.longitude_nodes,
                    self.latitude_nodes, self.latitude_spacing, device)

  @property
  def _precision(self) -> int:
    return _lib.PRECISIONS[self.transform_precision]

  @functools.cached_property
  def basis(self) -> _SphericalHarmonicBasis:
    """Dense (f, p, w) in the reference's layout, unfolded from the GPU tables."""
    plan = self.plan(torch.device('cuda', torch.cuda.current_device()))
    M, L, I, J = plan.M, plan.L, plan.I, plan.J
    folded = plan.folded_table()
    jh = (J + 1) // 2
    p = np.zeros((M, J, L))
    pos = 0
    for m in range(M):
      rows = folded[pos:pos + (L - m) * jh].reshape(L - m, jh)
      pos += (L - m) * jh
      up = rows.T  # (jh, L - m) at j = J - jh + jj
      p[m, J - jh:, m:] = up
      sign = np.where((np.arange(m, L) - m) % 2 == 0, 1.0, -1.0)
      p[m, :jh, m:] = (up * sign)[::-1]
    p = np.repeat(p, 2, axis=0)[1:]
    i = np.arange(I)[:, None]
    j = np.arange(1, M)[None, :]
    ang = 2 * np.pi * ((i * j) % I) / I
    f = np.empty((I, 2 * M - 1))
    f[:, 0] = 1 / np.sqrt(2 * np.pi)
    f[:, 1::2] = np.cos(ang) / np.sqrt(np.pi)
    f[:, 2::2] = np.sin(ang) / np.sqrt(np.pi)
    x, wp = plan.nodes()
    return _SphericalHarmonicBasis(f=f, p=p, w=(2 * np.pi / I) * wp)

  def inverse_transform(self, x, _triangle_only_adjoint=False):
    """modal (..., 2M-1, L) -> nodal (..., I, J)."""
    t = x
    lead = t.shape[:-2]
    if t.shape[-2:] != self.modal_shape:
      raise ValueError(f'expected modal shape (..., {self.modal_shape}), got {tuple(t.shape)}')
    plan = self.plan(t.device)
    y = _ToNodal.apply(t.reshape((-1,) + self.modal_shape), plan, self._precision,
                       _triangle_only_adjoint)
    return y.reshape(lead + self.nodal_shape)

  def transform(self, x, _triangle_only=False):
    """nodal (..., I, J) -> modal (..., 2M-1, L)."""
    t = x
    lead = t.shape[:-2]
    if t.shape[-2:] != self.nodal_shape:
      raise ValueError(f'expected nodal shape (..., {self.nodal_shape}), got {tuple(t.shape)}')
    plan = self.plan(t.device)
    y = _ToModal.apply(t.reshape((-1,) + self.nodal_shape), plan, self._precision,
                       _triangle_only)
    return y.reshape(lead + self.modal_shape)

  def longitudinal_derivative(self, x):
    return _real_basis_derivative(x)


def _real_basis_derivative(u):
  """u_x[2j-1] = j u[2j], u_x[2j] = -j u[2j-1], u_x[0] = 0 along axis -2
  (fourier.py:63-88)."""
  n = u.shape[-2]
  if n % 2 != 1:
    raise ValueError(f'{u.shape=} along axis=-2 is not odd')
  idx, coef = _ddlon_tables(n, u.device, u.dtype)
  return torch.index_select(u, -2, idx) * coef[:, None]


@functools.lru_cache(maxsize=256)
def _ddlon_tables(n, device, dtype):
  i = np.arange(n)
  j = (i + 1) // 2
  src = np.where(i % 2 == 1, i + 1, i - 1)
  src[0] = 0
  src = np.clip(src, 0, n - 1)
  coef = np.where(i % 2 == 1, j, -j).astype(np.float64)
  coef[0] = 0
  return (torch.as_tensor(src, device=device),
          torch.as_tensor(coef, dtype=dtype, device=device))


def _shift_l(x, offset: int):
  """Zero-padded shift along the last axis (jax_numpy_utils.py:158-167)."""
  n = x.shape[-1]
  if abs(offset) >= n:
    return torch.zeros_like(x)
  if offset > 0:
    return torch.nn.functional.pad(x[..., :n - offset], (offset, 0))
  return torch.nn.functional.pad(x[..., -offset:], (0, -offset))


def _round_to_multiple(x: int, multiple: int) -> int:
  return multiple * math.ceil(x / multiple)


@dataclasses.dataclass(frozen=True)
class FastSphericalHarmonics(SphericalHarmonics):
  """The reference's "fast" modal layout (spherical_harmonic.py:411-621):
  rows m = [0, 0', +1, -1, +2, -2, ...] with an always-zero imaginary row 0'
  for m = 0, so modal arrays are (2M, L), optionally padded to multiples of
  `base_shape_multiple`.

  On TPUs that layout halves the Legendre operand and shards evenly.  Here
  the transforms run in the Real layout's kernels (the per-m GEMMs never
  materialise a repeated P, so the layout buys nothing on this GPU): the
  modal rows are regathered around them, which costs one extra pass over the
  modal array per transform.  `transform_precision` as for
  RealSphericalHarmonics ('tensorfloat32', the reference default, means FP32
  here); `spmd_mesh` must be None (one GPU per grid).
  """
  spmd_mesh: Any = None
  base_shape_multiple: int | None = None
  reverse_einsum_arg_order: bool | None = None
  stacked_fourier_transforms: bool | None = None
  transform_precision: str = 'tensorfloat32'

  def __post_init__(self):
    if self.spmd_mesh is not None:
      raise ValueError('dinosaur_amd runs one grid per GPU; spmd_mesh is not supported')
    if self.base_shape_multiple is None:
      object.__setattr__(self, 'base_shape_multiple', 1)
    if self.reverse_einsum_arg_order is None:
      object.__setattr__(self, 'reverse_einsum_arg_order', False)
    if self.stacked_fourier_transforms is None:
      unstacked = math.ceil(self.longitude_wavenumbers / 128)
      stacked = 2 * math.ceil(self.longitude_wavenumbers / 256)
      object.__setattr__(self, 'stacked_fourier_transforms', stacked <= unstacked)
    if self._precision_name not in _lib.PRECISIONS:
      raise ValueError(f'unknown transform_precision {self.transform_precision!r}')

  @property
  def _precision_name(self):
    return {'tensorfloat32': 'float32', 'highest': 'float32', 'default': 'float32'}.get(
        self.transform_precision, self.transform_precision)

  @functools.cached_property
  def _real(self) -> RealSphericalHarmonics:
    return RealSphericalHarmonics(self.longitude_wavenumbers, self.total_wavenumbers,
                                  self.longitude_nodes, self.latitude_nodes,
                                  self.latitude_spacing, transform_precision=self._precision_name)

  @functools.cached_property
  def nodal_limits(self):
    return (self.longitude_nodes, self.latitude_nodes)

  @functools.cached_property
  def modal_limits(self):
    return (2 * self.longitude_wavenumbers, self.total_wavenumbers)

  @functools.cached_property
  def nodal_shape(self):
    b = self.base_shape_multiple
    return tuple(_round_to_multiple(x, k) for x, k in zip(self.nodal_limits, (b, b)))

  @functools.cached_property
  def modal_shape(self):
    b = self.base_shape_multiple
    return tuple(_round_to_multiple(x, k) for x, k in zip(self.modal_limits, (2 * b, b)))

  @functools.cached_property
  def nodal_padding(self):
    return tuple(x - y for x, y in zip(self.nodal_shape, self.nodal_limits))

  @functools.cached_property
  def modal_padding(self):
    return tuple(x - y for x, y in zip(self.modal_shape, self.modal_limits))

  @functools.cached_property
  def nodal_axes(self):
    px, py = self.nodal_padding
    lon, sin_lat = self._real.nodal_axes
    return np.pad(lon, [(0, px)]), np.pad(sin_lat, [(0, py)])

  @functools.cached_property
  def modal_axes(self):
    px, py = self.modal_padding
    m_pos = np.arange(1, self.longitude_wavenumbers)
    m_pos_neg = np.stack([m_pos, -m_pos], axis=1).ravel()
    return (np.pad(np.concatenate([[0, 0], m_pos_neg]), [(0, px)]),
            np.pad(np.arange(self.total_wavenumbers), [(0, py)]))

  @functools.cached_property
  def modal_dtype(self):
    return np.dtype(np.float32)

  @functools.cached_property
  def mask(self):
    m, l = np.meshgrid(*self.modal_axes, indexing='ij')
    i, j = np.meshgrid(*(np.arange(s) for s in self.modal_shape), indexing='ij')
    i_lim, j_lim = self.modal_limits
    return (abs(m) <= l) & (i != 1) & (i < i_lim) & (j < j_lim)

  def plan(self, device) -> Plan:
    return self._real.plan(device)

  @functools.cached_property
  def basis(self) -> _SphericalHarmonicBasis:
    """Dense (f, p, w) in this layout (P not repeated over +-m)."""
    b = self._real.basis
    px, py = self.nodal_padding
    mx, my = self.modal_padding
    M = self.longitude_wavenumbers
    f = np.zeros((self.longitude_nodes, 2 * M))
    f[:, 0] = b.f[:, 0]
    f[:, 2:] = b.f[:, 1:]
    f = np.pad(f, [(0, px), (0, mx)])
    if self.stacked_fourier_transforms:
      f = np.reshape(f, (-1, 2, f.shape[-1] // 2), order='F')
    p = b.p[0::2]  # rows 0, +1, +2, ... of the repeated Real table: one per m
    p = np.pad(p, [(0, mx // 2), (0, py), (0, my)])
    w = np.pad(b.w, [(0, py)])
    return _SphericalHarmonicBasis(f=f, p=p, w=w)

  @functools.lru_cache(maxsize=8)
  def _rows(self, device):
    """Real row r of the fast layout's row 0, 2, 3, ..., 2M-1 (row 1 is 0')."""
    R = 2 * self.longitude_wavenumbers - 1
    real_rows = torch.arange(R, device=device)
    fast_of_real = torch.where(real_rows == 0, real_rows, real_rows + 1)
    return fast_of_real

  def inverse_transform(self, x):
    """modal (..., 2M(+pad), L(+pad)) -> nodal (..., I(+pad), J(+pad))."""
    if tuple(x.shape[-2:]) != self.modal_shape:
      raise ValueError(f'expected modal shape (..., {self.modal_shape}), got {tuple(x.shape)}')
    L = self.total_wavenumbers
    xr = torch.index_select(x[..., :, :L], -2, self._rows(x.device))
    y = self._real.inverse_transform(xr)
    px, py = self.nodal_padding
    return torch.nn.functional.pad(y, (0, py, 0, px)) if px or py else y

  def transform(self, x):
    """nodal (..., I(+pad), J(+pad)) -> modal (..., 2M(+pad), L(+pad))."""
    if tuple(x.shape[-2:]) != self.nodal_shape:
      raise ValueError(f'expected nodal shape (..., {self.nodal_shape}), got {tuple(x.shape)}')
    I, J = self.nodal_limits
    y = self._real.transform(x[..., :I, :J])
    out = y.new_zeros(y.shape[:-2] + self.modal_shape)
    out = out.index_copy(-2, self._rows(y.device),
                         torch.nn.functional.pad(y, (0, self.modal_padding[1])))
    return out

  def longitudinal_derivative(self, x):
    """u_x[2m] = m u[2m+1], u_x[2m+1] = -m u[2m] (fourier.py:112-126)."""
    n = x.shape[-2]
    if n % 2:
      raise ValueError(f'{x.shape=} along axis=-2 is not even')
    i = torch.arange(n, device=x.device)
    m = (i // 2).to(x.dtype)
    partner = torch.where(i % 2 == 0, i + 1, i - 1)
    coef = torch.where(i % 2 == 0, m, -m)
    return torch.index_select(x, -2, partner) * coef[:, None]


RealSphericalHarmonicsWithZeroImag = FastSphericalHarmonics


# ----------------------------------------------------------------- Grid

SphericalHarmonicsImpl = Callable[..., SphericalHarmonics]


@dataclasses.dataclass(frozen=True)
class Grid:
  """Real-space and spectral grids over the sphere (spherical_harmonic.py:678-1151)."""
  longitude_wavenumbers: int = 0
  total_wavenumbers: int = 0
  longitude_nodes: int = 0
  latitude_nodes: int = 0
  latitude_spacing: str = 'gauss'
  longitude_offset: float = 0.0
  radius: float | None = None
  spherical_harmonics_impl: SphericalHarmonicsImpl = RealSphericalHarmonics
  spmd_mesh: Any = None

  def __post_init__(self):
    if self.radius is None:
      object.__setattr__(self, 'radius', 1.0)
    if self.latitude_spacing not in LATITUDE_SPACINGS:
      raise ValueError(f'Unsupported `latitude_spacing` "{self.latitude_spacing}". '
                       f'Supported values are: {list(LATITUDE_SPACINGS)}.')
    if self.spmd_mesh is not None:
      raise ValueError('dinosaur_amd runs one grid per GPU; spmd_mesh is not supported')

  @classmethod
  def with_wavenumbers(cls, longitude_wavenumbers, dealiasing='quadratic',
                       latitude_spacing='gauss', longitude_offset=0.0,
                       spherical_harmonics_impl=RealSphericalHarmonics, radius=None,
                       spmd_mesh=None):
    order = {'linear': 2, 'quadratic': 3, 'cubic': 4}[dealiasing]
    longitude_nodes = order * longitude_wavenumbers + 1
    latitude_nodes = math.ceil(longitude_nodes / 2)
    return cls(longitude_wavenumbers, longitude_wavenumbers + 1, longitude_nodes,
               latitude_nodes, latitude_spacing, longitude_offset, radius,
               spherical_harmonics_impl, spmd_mesh)

  @classmethod
  def construct(cls, max_wavenumber, gaussian_nodes, latitude_spacing='gauss',
                longitude_offset=0.0, radius=None,
                spherical_harmonics_impl=RealSphericalHarmonics, spmd_mesh=None):
    return cls(max_wavenumber + 1, max_wavenumber + 2, 4 * gaussian_nodes,
               2 * gaussian_nodes, latitude_spacing, longitude_offset, radius,
               spherical_harmonics_impl, spmd_mesh)

  # pylint:disable=invalid-name
  @classmethod
  def T21(cls, **kw): return cls.construct(21, 16, **kw)
  @classmethod
  def T31(cls, **kw): return cls.construct(31, 24, **kw)
  @classmethod
  def T42(cls, **kw): return cls.construct(42, 32, **kw)
  @classmethod
  def T85(cls, **kw): return cls.construct(85, 64, **kw)
  @classmethod
  def T106(cls, **kw): return cls.construct(106, 80, **kw)
  @classmethod
  def T119(cls, **kw): return cls.construct(119, 90, **kw)
  @classmethod
  def T170(cls, **kw): return cls.construct(170, 128, **kw)
  @classmethod
  def T213(cls, **kw): return cls.construct(213, 160, **kw)
  @classmethod
  def T340(cls, **kw): return cls.construct(340, 256, **kw)
  @classmethod
  def T425(cls, **kw): return cls.construct(425, 320, **kw)
  @classmethod
  def TL31(cls, **kw): return cls.construct(31, 16, **kw)
  @classmethod
  def TL47(cls, **kw): return cls.construct(47, 24, **kw)
  @classmethod
  def TL63(cls, **kw): return cls.construct(63, 32, **kw)
  @classmethod
  def TL95(cls, **kw): return cls.construct(95, 48, **kw)
  @classmethod
  def TL127(cls, **kw): return cls.construct(127, 64, **kw)
  @classmethod
  def TL159(cls, **kw): return cls.construct(159, 80, **kw)
  @classmethod
  def TL179(cls, **kw): return cls.construct(179, 90, **kw)
  @classmethod
  def TL255(cls, **kw): return cls.construct(255, 128, **kw)
  @classmethod
  def TL639(cls, **kw): return cls.construct(639, 320, **kw)
  @classmethod
  def TL1279(cls, **kw): return cls.construct(1279, 640, **kw)
  # pylint:enable=invalid-name

  def asdict(self):
    items = {f.name: getattr(self, f.name) for f in dataclasses.fields(self)}
    impl = self.spherical_harmonics_impl
    items['spherical_harmonics_impl'] = getattr(impl, '__name__', getattr(
        getattr(impl, 'func', None), '__name__', str(impl)))
    items['spmd_mesh'] = ''
    return items

  @functools.cached_property
  def spherical_harmonics(self) -> RealSphericalHarmonics:
    return self.spherical_harmonics_impl(
        longitude_wavenumbers=self.longitude_wavenumbers,
        total_wavenumbers=self.total_wavenumbers,
        longitude_nodes=self.longitude_nodes,
        latitude_nodes=self.latitude_nodes,
        latitude_spacing=self.latitude_spacing)

  @property
  def longitudes(self):
    return self.nodal_axes[0]

  @property
  def latitudes(self):
    return np.arcsin(self.nodal_axes[1])

  @functools.cached_property
  def nodal_axes(self):
    lon, sin_lat = self.spherical_harmonics.nodal_axes
    return lon + self.longitude_offset, sin_lat

  @functools.cached_property
  def nodal_shape(self):
    return self.spherical_harmonics.nodal_shape

  @functools.cached_property
  def nodal_padding(self):
    return self.spherical_harmonics.nodal_padding

  @functools.cached_property
  def nodal_mesh(self):
    return np.meshgrid(*self.nodal_axes, indexing='ij')

  @functools.cached_property
  def modal_axes(self):
    return self.spherical_harmonics.modal_axes

  @functools.cached_property
  def modal_shape(self):
    return self.spherical_harmonics.modal_shape

  @functools.cached_property
  def modal_padding(self):
    return self.spherical_harmonics.modal_padding

  @functools.cached_property
  def mask(self):
    return self.spherical_harmonics.mask

  @functools.cached_property
  def modal_mesh(self):
    return np.meshgrid(*self.spherical_harmonics.modal_axes, indexing='ij')

  @functools.cached_property
  def cos_lat(self):
    _, sin_lat = self.nodal_axes
    return np.sqrt(1 - sin_lat ** 2)

  @functools.cached_property
  def sec2_lat(self):
    _, sin_lat = self.nodal_axes
    return 1 / (1 - sin_lat ** 2)

  @functools.cached_property
  def laplacian_eigenvalues(self):
    _, l = self.modal_axes
    return -l * (l + 1) / (self.radius ** 2)

  # ---- device-side constants
  def const(self, name: str, device, dtype) -> torch.Tensor:
    """Cached device tensor of a host metadata array (e.g. 'sec2_lat')."""
    return _grid_const(self, name, torch.device(device), dtype)

  # ---- transforms
  def to_nodal(self, x):
    """Maps modal -> nodal over every array leaf (spherical_harmonic.py:1017-1023)."""
    sh = self.spherical_harmonics

    def f(leaf):
      t, restore = _as_tensor(leaf)
      return restore(sh.inverse_transform(t))
    return tree_map_over_nonscalars(f, x)

  def to_modal(self, z):
    """Maps nodal -> modal over every array leaf (spherical_harmonic.py:1025-1031)."""
    sh = self.spherical_harmonics

    def f(leaf):
      t, restore = _as_tensor(leaf)
      return restore(sh.transform(t))
    return tree_map_over_nonscalars(f, z)

  def to_nodal_batched(self, leaves):
    """One launch per stage for a list of modal tensors sharing dtype/device."""
    return _batched(self, leaves, inverse=True)

  def to_modal_batched(self, leaves):
    return _batched(self, leaves, inverse=False)

  # ---- operators
  def _wrap(self, x, fn):
    t, restore = _as_tensor(x)
    return restore(fn(t))

  def laplacian(self, x):
    return self._wrap(x, lambda t: t * self.const('laplacian_eigenvalues', t.device, t.dtype))

  def inverse_laplacian(self, x):
    return self._wrap(x, lambda t: t * self.const('inverse_laplacian', t.device, t.dtype))

  def clip_wavenumbers(self, x, n: int = 1):
    if n <= 0:
      raise ValueError(f'`n` must be >= 0; got {n}.')

    def clip(leaf):
      def fn(t):
        mask = torch.ones(self.modal_shape[-1], dtype=t.dtype, device=t.device)
        mask[-(n + self.modal_padding[-1]):] = 0
        return t * mask
      return self._wrap(leaf, fn)
    return tree_map_over_nonscalars(clip, x)

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
    return self._wrap(x, self.spherical_harmonics.longitudinal_derivative)

  def cos_lat_d_dlat(self, x):
    def fn(t):
      up = self.const('cdlat_up', t.device, t.dtype)
      dn = self.const('cdlat_dn', t.device, t.dtype)
      return _shift_l(up * t, -1) + _shift_l(dn * t, +1)
    return self._wrap(x, fn)

  def sec_lat_d_dlat_cos2(self, x):
    def fn(t):
      up = self.const('scdlat_up', t.device, t.dtype)
      dn = self.const('scdlat_dn', t.device, t.dtype)
      return _shift_l(up * t, -1) + _shift_l(dn * t, +1)
    return self._wrap(x, fn)

  def cos_lat_grad(self, x, clip: bool = True):
    raw = (self.d_dlon(x) / self.radius, self.cos_lat_d_dlat(x) / self.radius)
    return self.clip_wavenumbers(raw) if clip else raw

  def k_cross(self, v):
    return -v[1], v[0]

  def div_cos_lat(self, v, clip: bool = True):
    raw = (self.d_dlon(v[0]) + self.sec_lat_d_dlat_cos2(v[1])) / self.radius
    return self.clip_wavenumbers(raw) if clip else raw

  def curl_cos_lat(self, v, clip: bool = True):
    raw = (self.d_dlon(v[1]) - self.sec_lat_d_dlat_cos2(v[0])) / self.radius
    return self.clip_wavenumbers(raw) if clip else raw

  @property
  def quadrature_weights(self):
    _, w = get_latitude_nodes(self.latitude_nodes, self.latitude_spacing)
    return np.broadcast_to((2 * np.pi / self.longitude_nodes) * w, self.nodal_shape)

  def integrate(self, z):
    def fn(t):
      w = self.const('quadrature_weights_lat', t.device, t.dtype) * self.radius ** 2
      return torch.einsum('y,...xy->...', w, t)
    return self._wrap(z, fn)


def _grid_host_const(grid: Grid, name: str) -> np.ndarray:
  if name == 'laplacian_eigenvalues':
    return grid.laplacian_eigenvalues
  if name == 'inverse_laplacian':
    with np.errstate(divide='ignore', invalid='ignore'):
      inv = 1 / grid.laplacian_eigenvalues
    inv[0] = 0
    inv[grid.total_wavenumbers:] = 0
    return inv
  if name in ('cdlat_up', 'cdlat_dn', 'scdlat_up', 'scdlat_dn'):
    _, l = grid.modal_mesh
    a, b = grid._derivative_recurrence_weights
    return {'cdlat_up': (l + 1) * a, 'cdlat_dn': -l * b,
            'scdlat_up': (l - 1) * a, 'scdlat_dn': -(l + 2) * b}[name]
  if name in ('sec2_lat', 'cos_lat'):
    return getattr(grid, name)
  if name == 'sin_lat':
    return grid.nodal_axes[1]
  if name == 'quadrature_weights_lat':
    _, w = get_latitude_nodes(grid.latitude_nodes, grid.latitude_spacing)
    return (2 * np.pi / grid.longitude_nodes) * w
  raise KeyError(name)


_const_cache: dict = {}


def _grid_const(grid, name, device, dtype):
  key = (grid, name, device, dtype)
  t = _const_cache.get(key)
  if t is None:
    t = torch.as_tensor(np.asarray(_grid_host_const(grid, name), np.float64),
                        device=device).to(dtype)
    _const_cache[key] = t
  return t


def _batched(grid: Grid, leaves, inverse: bool):
  """Concatenates leaves (..., a, b) along a flat batch, one transform, split."""
  sh = grid.spherical_harmonics
  tail = sh.modal_shape if inverse else sh.nodal_shape
  out_tail = sh.nodal_shape if inverse else sh.modal_shape
  flat = []
  leads = []
  for t in leaves:
    if tuple(t.shape[-2:]) != tail:
      raise ValueError(f'expected trailing shape {tail}, got {tuple(t.shape)}')
    leads.append(t.shape[:-2])
    flat.append(t.reshape((-1,) + tail))
  cat = torch.cat(flat, 0) if len(flat) > 1 else flat[0]
  out = sh.inverse_transform(cat) if inverse else sh.transform(cat)
  sizes = [f.shape[0] for f in flat]
  parts = torch.split(out, sizes, 0)
  return [p.reshape(lead + out_tail) for p, lead in zip(parts, leads)]


_CONSTANT_NORMALIZATION_FACTOR = 3.5449077


def add_constant(x, c):
  out = x.clone()
  out[..., 0, 0] += _CONSTANT_NORMALIZATION_FACTOR * c
  return out


def get_cos_lat_vector(vorticity, divergence, grid: Grid, clip: bool = True):
  """u cos(lat) from vorticity and divergence (spherical_harmonic.py:1164-1178)."""
  stream_function = grid.inverse_laplacian(vorticity)
  velocity_potential = grid.inverse_laplacian(divergence)
  a = grid.cos_lat_grad(velocity_potential, clip=clip)
  b = grid.k_cross(grid.cos_lat_grad(stream_function, clip=clip))
  return (a[0] + b[0], a[1] + b[1])


def uv_nodal_to_vor_div_modal(grid: Grid, u_nodal, v_nodal, clip: bool = True):
  """spherical_harmonic.py:1181-1195."""
  u, restore = _as_tensor(u_nodal)
  v, _ = _as_tensor(v_nodal)
  cos = grid.const('cos_lat', u.device, u.dtype)
  uc, vc = grid.to_modal_batched([u / cos, v / cos])
  return (restore(grid.curl_cos_lat((uc, vc), clip=clip)),
          restore(grid.div_cos_lat((uc, vc), clip=clip)))


def vor_div_to_uv_nodal(grid: Grid, vorticity, divergence, clip: bool = True):
  """spherical_harmonic.py:1198-1210."""
  z, restore = _as_tensor(vorticity)
  d, _ = _as_tensor(divergence)
  u, v = get_cos_lat_vector(z, d, grid, clip=clip)
  un, vn = grid.to_nodal_batched([u, v])
  cos = grid.const('cos_lat', un.device, un.dtype)
  return restore(un / cos), restore(vn / cos)
