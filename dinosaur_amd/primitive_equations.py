"""Primitive equations on sigma levels, for MI355X.

Drop-in for the sigma path of dinosaur/primitive_equations.py: ``State``
(:61-84), the vertical weight matrices (:283-541), ``div_sec_lat`` (:601-623),
``truncated_modal_orography`` (:626-636), ``PrimitiveEquationsSigma``
(:835-1458) with ``explicit_terms`` / ``implicit_terms`` / ``implicit_inverse``
and the deprecated aliases ``PrimitiveEquations``, ``MoistPrimitiveEquations``
and ``MoistPrimitiveEquationsWithCloudMoisture`` (:2534-2630).

Design.  One ``explicit_terms`` evaluation is one batched inverse transform of
every field it needs (vorticity, divergence, temperature, tracers, u cos,
v cos, grad ln ps, and the humidity extras), the nodal nonlinear terms, one
batched forward transform of every nodal product, and a modal epilogue.  The
reference instead transforms pytree leaf by leaf (pytree_utils.py:86-98).
``implicit_terms`` and ``implicit_inverse`` run FP64 GPU kernels;
``implicit_inverse`` factorises once per (levels, eta) through the vertical
modes of A = G H + R T_ref dsigma^T (exact: the reference's per-l block
inverse reduces to (I - eta^2 lambda_l A)^{-1} on the divergence).

States may carry leading batch (ensemble member) dimensions: leaves are
(..., K, 2M-1, L) tensors and every operation broadcasts over the leading
dimensions.
"""
from __future__ import annotations

import ctypes
import dataclasses
import functools
from typing import Any, Callable, Mapping

import numpy as np
import torch

from . import _lib
from . import coordinate_systems
from . import sigma_coordinates
from . import spherical_harmonic
from . import units as units_lib
from ._lib import StateShapeError

Array = Any


# ------------------------------------------------------------------ State


@dataclasses.dataclass
class State:
  """Spectral state of the primitive equations (primitive_equations.py:61-84).

  A state may be backed by one contiguous (fields, 2M-1, L) buffer whose
  views are the leaves, in the order vorticity, divergence,
  temperature_variation, log_surface_pressure, tracers (each member-major).
  Arithmetic on two such states is one tensor operation on the buffers;
  states built leaf by leaf work the same way, one operation per leaf.
  """
  vorticity: Array
  divergence: Array
  temperature_variation: Array
  log_surface_pressure: Array
  tracers: Mapping[str, Array] = dataclasses.field(default_factory=dict)
  sim_time: float | None = None
  _flat: Any = dataclasses.field(default=None, repr=False, compare=False)

  _ARRAYS = ('vorticity', 'divergence', 'temperature_variation', 'log_surface_pressure')

  @classmethod
  def from_flat(cls, buf, lead, K, names, sim_time=None):
    """State whose leaves are views of buf (fields, R, L)."""
    E = int(np.prod(lead)) if lead else 1
    tail = tuple(buf.shape[1:])
    sizes = [E * K] * 3 + [E] + [E * K] * len(names)
    parts = torch.split(buf, sizes, 0)
    lev = [K, K, K, 1] + [K] * len(names)
    views = [pp.view(tuple(lead) + (kk,) + tail) for pp, kk in zip(parts, lev)]
    st = cls(views[0], views[1], views[2], views[3], dict(zip(names, views[4:])), sim_time)
    st._flat = (buf, tuple(lead), K, tuple(names))
    return st

  def flat_buffer(self):
    """The backing buffer if every leaf is still its view, else None."""
    if self._flat is None:
      return None
    buf, lead, K, names = self._flat
    if tuple(self.tracers.keys()) != names:
      return None
    E = int(np.prod(lead)) if lead else 1
    step = buf.stride(0) * buf.element_size()
    base = buf.data_ptr()
    offs = [0, E * K, 2 * E * K, 3 * E * K]
    offs += [3 * E * K + E + i * E * K for i in range(len(names))]
    leaves = [getattr(self, f) for f in self._ARRAYS] + [self.tracers[n] for n in names]
    for leaf, off in zip(leaves, offs):
      if leaf.data_ptr() != base + off * step or not leaf.is_contiguous():
        return None
    return buf

  def as_flat(self):
    """This state backed by one (fields, R, L) buffer (see to_flat)."""
    return to_flat(self)

  def tree_map_arrays(self, fn):
    kw = {f: fn(getattr(self, f)) for f in self._ARRAYS}
    kw['tracers'] = {k: fn(v) for k, v in self.tracers.items()}
    kw['sim_time'] = self.sim_time
    return State(**kw)

  def _combine(self, other, op):
    if isinstance(other, State):
      st = None
      if self.sim_time is not None and other.sim_time is not None:
        st = op(self.sim_time, other.sim_time)
      a, b = self.flat_buffer(), other.flat_buffer()
      if a is not None and b is not None and self._flat[1:] == other._flat[1:]:
        return State.from_flat(op(a, b), *self._flat[1:], sim_time=st)
      kw = {f: op(getattr(self, f), getattr(other, f)) for f in self._ARRAYS}
      kw['tracers'] = {k: op(v, other.tracers[k]) for k, v in self.tracers.items()}
      kw['sim_time'] = st
      return State(**kw)
    st = None if self.sim_time is None else op(self.sim_time, other)
    a = self.flat_buffer()
    if a is not None:
      return State.from_flat(op(a, other), *self._flat[1:], sim_time=st)
    kw = {f: op(getattr(self, f), other) for f in self._ARRAYS}
    kw['tracers'] = {k: op(v, other) for k, v in self.tracers.items()}
    kw['sim_time'] = st
    return State(**kw)

  def __add__(self, other):
    return self._combine(other, lambda a, b: a + b)

  def __sub__(self, other):
    return self._combine(other, lambda a, b: a - b)

  def __mul__(self, c):
    return self._combine(c, lambda a, b: a * b)

  __rmul__ = __mul__

  def __neg__(self):
    return self * -1.0

  def asdict(self):
    out = {f: getattr(self, f) for f in self._ARRAYS}
    out['tracers'] = dict(self.tracers)
    if self.sim_time is not None:
      out['sim_time'] = self.sim_time
    return out

  @property
  def fields(self):
    return [f for f in dataclasses.fields(self) if not f.name.startswith('_')]


def flat_layout(state: State):
  """(lead, K, names) of a state's leaves."""
  lead = tuple(state.vorticity.shape[:-3])
  return lead, state.vorticity.shape[-3], tuple(state.tracers.keys())


def to_flat(state: State) -> State:
  """Flat-backed version of `state` (a copy unless it already is one)."""
  if state.flat_buffer() is not None:
    return state
  lead, K, names = flat_layout(state)
  R, L = state.vorticity.shape[-2:]
  leaves = [getattr(state, f) for f in State._ARRAYS] + [state.tracers[n] for n in names]
  buf = torch.cat([t.reshape(-1, R, L) for t in leaves], 0)
  return State.from_flat(buf, lead, K, names, sim_time=state.sim_time)


def validate_state_shape(state: State, coords: coordinate_systems.CoordinateSystem):
  """primitive_equations.py:91-120."""
  for name in ('vorticity', 'divergence', 'temperature_variation'):
    shape = tuple(getattr(state, name).shape[-3:])
    if shape != coords.modal_shape:
      raise StateShapeError(f'Expected {name} shape {coords.modal_shape}; got shape {shape}.')
  shape = tuple(state.log_surface_pressure.shape[-3:])
  if shape != coords.surface_modal_shape:
    raise StateShapeError(f'Expected log_surface_pressure shape {coords.surface_modal_shape}; '
                          f'got shape {shape}.')
  for tracer_name, array in state.tracers.items():
    if tuple(array.shape[-3:]) != coords.modal_shape:
      raise StateShapeError(f'Expected tracer {tracer_name} shape {coords.modal_shape}; '
                            f'got shape {tuple(array.shape)}.')


# ------------------------------------------------------------------ vertical weights


def get_sigma_ratios(coordinates: sigma_coordinates.SigmaCoordinates) -> np.ndarray:
  """alpha[j] = log(s[j+1] / s[j]) / 2, alpha[-1] = -log(s[-1]) (:283-307)."""
  alpha = np.diff(np.log(coordinates.centers), append=0) / 2
  alpha[-1] = -np.log(coordinates.centers[-1])
  return alpha


def get_geopotential_weights_sigma(coordinates, ideal_gas_constant) -> np.ndarray:
  """Upper-triangular G / R with alpha[j] on the diagonal and
  alpha[k] + alpha[k-1] above it (:310-336)."""
  alpha = get_sigma_ratios(coordinates)
  k = coordinates.layers
  # alpha[k] + alpha[k-1] in alpha's own dtype (float32 for float32 levels),
  # as the reference's scalar loop does
  pair = np.concatenate([[0.0], (alpha[1:] + alpha[:-1]).astype(np.float64)])
  upper = np.triu(np.ones((k, k)), 1) * pair[None, :]
  return ideal_gas_constant * (upper + np.diag(alpha.astype(np.float64)))


def get_temperature_implicit_weights_sigma(coordinates, reference_temperature, kappa) -> np.ndarray:
  """H (:418-487): kappa T_ref alpha terms minus the reference-lapse terms K."""
  ref = np.asarray(reference_temperature)
  if ref.ndim != 1 or ref.shape[-1] != coordinates.layers:
    raise ValueError('`reference_temp` must be a vector of length `coordinates.layers`; '
                     f'got shape {ref.shape} and {coordinates.layers} layers.')
  k = coordinates.layers
  lower = np.tril(np.ones((k, k)))
  alpha = get_sigma_ratios(coordinates)[:, None]
  pa = lower * alpha
  pa_prev = np.zeros_like(pa)
  pa_prev[1:] = pa[:-1]
  ds = coordinates.layer_thickness
  h0 = kappa * ref[:, None] * (pa + pa_prev) / ds[:, None]
  k0 = np.concatenate((np.diff(ref) / (ds[:-1] + ds[1:]), [0]))[:, None]
  kk = k0 * (lower - np.cumsum(ds)[:, None])
  kk_prev = np.zeros_like(kk)
  kk_prev[1:] = kk[:-1]
  return (h0 - kk - kk_prev) * ds


def _level_matrix(a: np.ndarray, x: torch.Tensor) -> torch.Tensor:
  """'gh,...hml->...gml'."""
  at = torch.as_tensor(np.asarray(a, np.float64), device=x.device).to(x.dtype)
  return torch.einsum('gh,...hml->...gml', at, x)


def get_geopotential_diff_sigma(temperature, coordinates, ideal_gas_constant, method='dense',
                                sharding=None):
  del method, sharding
  return _level_matrix(get_geopotential_weights_sigma(coordinates, ideal_gas_constant),
                       temperature)


def get_temperature_implicit_sigma(divergence, coordinates, reference_temperature, kappa,
                                   method='dense', sharding=None):
  del method, sharding
  return _level_matrix(
      -get_temperature_implicit_weights_sigma(coordinates, reference_temperature, kappa),
      divergence)


def _get_implicit_term_matrix_sigma(eta, coords, reference_temperature, kappa,
                                    ideal_gas_constant) -> np.ndarray:
  """(L, 2K+1, 2K+1) FP64 matrix of (I - eta G) (:544-598); used for checks."""
  k = coords.vertical.layers
  lam = coords.horizontal.laplacian_eigenvalues
  l = coords.horizontal.modal_shape[1]
  g = get_geopotential_weights_sigma(coords.vertical, ideal_gas_constant)
  h = get_temperature_implicit_weights_sigma(coords.vertical, reference_temperature, kappa)
  t = np.asarray(reference_temperature, np.float64)
  ds = coords.vertical.layer_thickness.astype(np.float64)
  out = np.zeros((l, 2 * k + 1, 2 * k + 1))
  out[:, :k, :k] = np.eye(k)
  out[:, :k, k:2 * k] = eta * lam[:, None, None] * g[None]
  out[:, :k, 2 * k] = eta * ideal_gas_constant * lam[:, None] * t[None, :]
  out[:, k:2 * k, :k] = eta * h[None]
  out[:, k:2 * k, k:2 * k] = np.eye(k)
  out[:, 2 * k, :k] = eta * ds[None, :]
  out[:, 2 * k, 2 * k] = 1.0
  return out


# ------------------------------------------------------------------ helpers


def div_sec_lat(m_component, n_component, grid: spherical_harmonic.Grid):
  """H(M, N) in the modal basis (:601-623)."""
  sec2 = grid.const('sec2_lat', m_component.device, m_component.dtype)
  mm, nn = grid.to_modal_batched([m_component * sec2, n_component * sec2])
  return grid.div_cos_lat((mm, nn), clip=False)


def truncated_modal_orography(orography, coords, wavenumbers_to_clip: int = 1):
  """Modal orography with the top wavenumbers removed (:626-636)."""
  grid = coords.horizontal
  if tuple(orography.shape) != grid.nodal_shape:
    raise ValueError(f'Expected nodal orography with shape={grid.nodal_shape}')
  return grid.clip_wavenumbers(grid.to_modal(orography), n=wavenumbers_to_clip)


def _vadv(w, x, dsig_c2c_inv, dtype_like):
  """centered_vertical_advection along axis -3 with zero boundary fluxes."""
  xd = torch.diff(x, dim=-3) * dsig_c2c_inv
  wx = w * xd
  wxp = torch.nn.functional.pad(wx, (0, 0, 0, 0, 1, 1))
  return -0.5 * (wxp[..., 1:, :, :] + wxp[..., :-1, :, :])


# ------------------------------------------------------------------ implicit factor


class _ImplicitOperator:
  """Holds the FP64 level matrices and per-eta GPU factor objects."""

  def __init__(self, coords, reference_temperature, kappa, r_gas, matrices=None):
    """matrices: None for sigma levels, else (G, H, c, w) of another vertical
    discretisation whose implicit block matrix has the same pattern
    [[I, eta lam G, eta lam c], [eta H, I, 0], [eta w^T, 0, 1]]: c takes the
    place of R T_ref and w that of dsigma."""
    v = coords.vertical
    self.K = v.layers
    self.L = coords.horizontal.modal_shape[1]
    self.R = coords.horizontal.modal_shape[0]
    self.r_gas = float(r_gas)
    if matrices is None:
      G = get_geopotential_weights_sigma(v, r_gas)
      H = get_temperature_implicit_weights_sigma(v, reference_temperature, kappa)
      tref = np.asarray(reference_temperature, np.float64)
      dsig = v.layer_thickness
    else:
      G, H, c, dsig = matrices
      tref = np.asarray(c, np.float64) / self.r_gas
    self.G = np.ascontiguousarray(G, np.float64)
    self.H = np.ascontiguousarray(H, np.float64)
    self.tref = np.ascontiguousarray(tref, np.float64)
    self.dsig = np.ascontiguousarray(np.asarray(dsig, np.float64))
    self.lam = np.ascontiguousarray(coords.horizontal.laplacian_eigenvalues.astype(np.float64))
    a = self.G @ self.H + self.r_gas * np.outer(self.tref, self.dsig)
    d, vecs = np.linalg.eig(a)
    if np.abs(d.imag).max() > 1e-12 * max(1.0, np.abs(d.real).max()):
      raise ValueError('vertical structure matrix has complex eigenvalues; the reference '
                       'temperature profile is outside the supported (stable) regime')
    vecs = vecs.real
    if np.linalg.cond(vecs) > 1e8:
      raise ValueError('vertical modes are ill-conditioned (cond > 1e8)')
    self.V = np.ascontiguousarray(vecs)
    self.Vinv = np.ascontiguousarray(np.linalg.inv(vecs))
    self.d = np.ascontiguousarray(d.real)
    self._factors = {}

  def factor(self, eta: float, device) -> ctypes.c_void_p:
    key = (float(eta), torch.device(device).index)
    f = self._factors.get(key)
    if f is None:
      denom = 1.0 - eta * eta * self.lam[:, None] * self.d[None, :]
      if np.abs(denom).min() < 1e-12:
        raise ValueError(f'implicit matrix is singular for step_size={eta}')
      h = ctypes.c_void_p()
      dp = _lib.dptr
      with torch.cuda.device(device):
        rc = _lib.lib().dino_implicit_create(
            self.K, self.L, self.R, dp(self.G), dp(self.H), dp(self.tref), dp(self.dsig),
            dp(self.lam), dp(self.V), dp(self.Vinv), dp(self.d), self.r_gas, float(eta),
            ctypes.byref(h))
      _lib.check(rc, 'dino_implicit_create')
      f = h
      self._factors[key] = f
    return f

  def __del__(self):
    try:
      if _lib._lib is None:
        return
      for h in getattr(self, '_factors', {}).values():
        _lib.release('dino_implicit_destroy', h)
    except Exception:  # pylint: disable=broad-except  (interpreter shutdown)
      pass

  def apply_into(self, fn_name, eta, members, div, temp, lnps, od, ot, ol):
    f = self.factor(eta, div.device)
    code = _lib.DINO_F64 if div.dtype == torch.float64 else _lib.DINO_F32
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    stream = ctypes.c_void_p(torch.cuda.current_stream(div.device).cuda_stream)
    with torch.cuda.device(div.device):
      rc = getattr(_lib.lib(), fn_name)(f, p(div), p(temp), p(lnps), p(od), p(ot), p(ol),
                                        members, code, stream)
    _lib.check(rc, fn_name)
    if _lib._active is not None:
      self._tally(fn_name, members, div)

  def _tally(self, fn_name, members, like):
    """Algorithmic work of one implicit kernel: the (2K+1) fields of each
    member read and written once on the triangle (read twice when
    accumulating); FP64 flops of the per-degree (2K+1)^2 block for the solve
    and of the G, R T_ref, H and dsigma products for the terms."""
    K = self.K
    n_real = _lib.triangle_entries(self.R, self.L)
    cols = n_real * members
    passes = 3 if fn_name.endswith('accumulate') else 2
    nbytes = like.element_size() * cols * (2 * K + 1) * passes
    if 'solve' in fn_name:
      flops = 2.0 * (2 * K + 1) ** 2 * cols
    else:
      flops = 2.0 * (2 * K * K + 2 * K) * cols
    _lib.tally(fn_name[len('dino_'):].replace('_accumulate', ''), nbytes, flops, 'f64')

  def apply(self, fn_name, eta, div, temp, lnps):
    div = div.contiguous()
    temp = temp.contiguous()
    lnps = lnps.contiguous()
    members = int(np.prod(div.shape[:-3])) if div.ndim > 3 else 1
    f = self.factor(eta, div.device)
    outs = [torch.empty_like(div), torch.empty_like(temp), torch.empty_like(lnps)]
    code = _lib.DINO_F64 if div.dtype == torch.float64 else _lib.DINO_F32
    if div.dtype not in (torch.float32, torch.float64):
      raise TypeError(f'unsupported dtype {div.dtype}')
    p = lambda t: ctypes.c_void_p(t.data_ptr())
    stream = ctypes.c_void_p(torch.cuda.current_stream(div.device).cuda_stream)
    with torch.cuda.device(div.device):
      rc = getattr(_lib.lib(), fn_name)(f, p(div), p(temp), p(lnps), p(outs[0]), p(outs[1]),
                                        p(outs[2]), members, code, stream)
    _lib.check(rc, fn_name)
    return outs


class _ImplicitSolve(torch.autograd.Function):
  """(I - eta G)^{-1} on (div, temp, lnps); the backward is the transposed
  solve kernel (dino_implicit_solve_adjoint)."""

  @staticmethod
  def forward(ctx, div, temp, lnps, op, eta):
    ctx.op, ctx.eta = op, eta
    return tuple(op.apply('dino_implicit_solve', eta, div, temp, lnps))

  @staticmethod
  def backward(ctx, gd, gt, gl):
    return tuple(ctx.op.apply('dino_implicit_solve_adjoint', ctx.eta, gd, gt, gl)) + (
        None, None)


class _ImplicitTerms(torch.autograd.Function):
  """Implicit tendencies; the backward is dino_implicit_terms_adjoint."""

  @staticmethod
  def forward(ctx, div, temp, lnps, op, eta):
    ctx.op, ctx.eta = op, eta
    return tuple(op.apply('dino_implicit_terms', eta, div, temp, lnps))

  @staticmethod
  def backward(ctx, gd, gt, gl):
    return tuple(ctx.op.apply('dino_implicit_terms_adjoint', ctx.eta, gd, gt, gl)) + (
        None, None)


# ------------------------------------------------------------------ fused nodal terms


class _NodalHandle:
  def __init__(self, handle):
    self.handle = handle

  def __del__(self):
    try:
      if self.handle.value:
        _lib.release('dino_nodal_destroy', self.handle)
    except Exception:  # pylint: disable=broad-except  (interpreter shutdown)
      pass


class _NodalFused(torch.autograd.Function):
  """(fields_in, I, J) nodal inputs -> (fields_out, I, J) nodal products, with
  the hand-written adjoint kernel as backward."""

  @staticmethod
  def forward(ctx, nodal, h, members):
    nodal = nodal.contiguous()
    lib = _lib.lib()
    n_out = lib.dino_nodal_fields(h.handle, members, 1)
    n_in = lib.dino_nodal_fields(h.handle, members, 0)
    if nodal.shape[0] != n_in:
      raise ValueError(f'fused nodal kernel expects {n_in} input fields, got {nodal.shape[0]}')
    out = torch.empty((n_out,) + tuple(nodal.shape[1:]), dtype=nodal.dtype, device=nodal.device)
    code = _lib.DINO_F64 if nodal.dtype == torch.float64 else _lib.DINO_F32
    stream = ctypes.c_void_p(torch.cuda.current_stream(nodal.device).cuda_stream)
    if not _lib.skip_values():  # the backward reads only the saved inputs
      with torch.cuda.device(nodal.device):
        rc = lib.dino_nodal_explicit(h.handle, ctypes.c_void_p(nodal.data_ptr()),
                                     ctypes.c_void_p(out.data_ptr()), members, code, stream)
      _lib.check(rc, 'dino_nodal_explicit')
      _lib.tally('nodal', nodal.element_size() * nodal[0].numel() * (n_in + n_out))
    ctx.save_for_backward(nodal)
    ctx.h, ctx.members, ctx.code = h, members, code
    return out

  @staticmethod
  def backward(ctx, g):
    (nodal,) = ctx.saved_tensors
    g = g.contiguous()
    gin = torch.empty_like(nodal)
    stream = ctypes.c_void_p(torch.cuda.current_stream(nodal.device).cuda_stream)
    with torch.cuda.device(nodal.device):
      rc = _lib.lib().dino_nodal_explicit_adjoint(
          ctx.h.handle, ctypes.c_void_p(nodal.data_ptr()), ctypes.c_void_p(g.data_ptr()),
          ctypes.c_void_p(gin.data_ptr()), ctx.members, ctx.code, stream)
    _lib.check(rc, 'dino_nodal_explicit_adjoint')
    _lib.tally('nodal_adjoint', g.element_size() * nodal[0].numel()
               * (2 * nodal.shape[0] + g.shape[0]))
    return gin, None, None


class _Handle:
  """Owns a library object created by dino_*_create."""

  def __init__(self, handle, destroy: str):
    self.handle = handle
    self._destroy = destroy

  def __del__(self):
    try:
      if self.handle.value:
        _lib.release(self._destroy, self.handle)
    except Exception:  # pylint: disable=broad-except  (interpreter shutdown)
      pass


def _launch(name, h, a, b, out, members, like):
  code = _lib.DINO_F64 if like.dtype == torch.float64 else _lib.DINO_F32
  stream = ctypes.c_void_p(torch.cuda.current_stream(like.device).cuda_stream)
  ptr = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else None
  with torch.cuda.device(like.device):
    rc = getattr(_lib.lib(), name)(h.handle, ptr(a), ptr(b), ptr(out), members, code, stream)
  _lib.check(rc, name)
  if _lib._active is not None:
    n_real = _lib.triangle_entries(*out.shape[-2:])
    fields = a.shape[0] + (1 if b is not None else 0) + out.shape[0]
    _lib.tally(name[len('dino_'):], like.element_size() * n_real * fields)


class _ModalPrep(torch.autograd.Function):
  """State buffer -> inverse-transform input buffer (adjoint: transposed
  stencils)."""

  @staticmethod
  def forward(ctx, buf, h, members):
    buf = buf.contiguous()
    n = _lib.lib().dino_modal_fields(h.handle, members, 1)
    out = torch.empty((n,) + tuple(buf.shape[1:]), dtype=buf.dtype, device=buf.device)
    _launch('dino_modal_prep', h, buf, None, out, members, buf)
    ctx.h, ctx.members, ctx.shape = h, members, tuple(buf.shape)
    return out

  @staticmethod
  def backward(ctx, g):
    g = g.contiguous()
    gin = torch.empty(ctx.shape, dtype=g.dtype, device=g.device)
    _launch('dino_modal_prep_adjoint', ctx.h, g, None, gin, ctx.members, g)
    return gin, None, None


class _ModalEpilogue(torch.autograd.Function):
  """Forward-transform output buffer -> state-shaped tendency buffer."""

  @staticmethod
  def forward(ctx, x, h, members, orography):
    x = x.contiguous()
    n = _lib.lib().dino_modal_fields(h.handle, members, 0)
    out = torch.empty((n,) + tuple(x.shape[1:]), dtype=x.dtype, device=x.device)
    if not _lib.skip_values():  # linear: the backward reads nothing saved
      _launch('dino_modal_epilogue', h, x, orography, out, members, x)
    ctx.h, ctx.members, ctx.shape = h, members, tuple(x.shape)
    return out

  @staticmethod
  def backward(ctx, g):
    g = g.contiguous()
    gin = torch.empty(ctx.shape, dtype=g.dtype, device=g.device)
    _launch('dino_modal_epilogue_adjoint', ctx.h, g, None, gin, ctx.members, g)
    return gin, None, None, None


def _state_views(buf, members, K, nt):
  """(div, temp, lnps) views of a flat state buffer."""
  E = members
  return buf[E * K:2 * E * K], buf[2 * E * K:3 * E * K], buf[3 * E * K:3 * E * K + E]


class _ImplicitFlat(torch.autograd.Function):
  """Implicit terms (solve=False) or (I - eta G)^{-1} (solve=True) on a flat
  state buffer; FP64 GPU kernels, transposed kernels as backward."""

  @staticmethod
  def _run(buf, op, eta, members, nt, solve, adjoint):
    K, E = op.K, members
    out = torch.empty_like(buf)
    d, t, l = _state_views(buf, members, K, nt)
    od, ot, ol = _state_views(out, members, K, nt)
    if solve:  # vorticity and tracers pass through
      out[:E * K].copy_(buf[:E * K])
      out[3 * E * K + E:].copy_(buf[3 * E * K + E:])
    else:  # they have no implicit tendency
      out[:E * K].zero_()
      out[3 * E * K + E:].zero_()
    name = 'dino_implicit_solve' if solve else 'dino_implicit_terms'
    op.apply_into(name + ('_adjoint' if adjoint else ''), eta, members, d, t, l, od, ot, ol)
    return out

  @staticmethod
  def forward(ctx, buf, op, eta, members, nt, solve):
    ctx.op, ctx.eta, ctx.members, ctx.nt, ctx.solve = op, eta, members, nt, solve
    return _ImplicitFlat._run(buf.contiguous(), op, eta, members, nt, solve, False)

  @staticmethod
  def backward(ctx, g):
    gin = _ImplicitFlat._run(g.contiguous(), ctx.op, ctx.eta, ctx.members, ctx.nt, ctx.solve,
                             True)
    return gin, None, None, None, None, None


class _FlatImexOps:
  """Flat-buffer operations of one state layout for the fused IMEX step
  (time_integration._fused_imex_step): explicit terms (differentiable),
  implicit terms stored only over the divergence / temperature / ln ps range
  (vorticity and tracers have none, primitive_equations.py:1230-1249), their
  accumulating adjoint, and the implicit solve and its adjoint in place.

  A compact implicit buffer holds flat elements [go, ie) of a state, where
  [ib, ie) is the divergence..ln ps range and go = ib rounded down to a
  multiple of 4, so that 16-byte vectors of full and compact buffers line up.
  """

  def __init__(self, eq, layout, members):
    self.eq = eq
    self.layout = layout
    lead, K, names = layout
    self.members = members
    self.K = K
    self.names = names
    R, L = eq.coords.horizontal.modal_shape
    self.RL = R * L
    E = members
    self.ib = E * K * self.RL
    self.ie = (3 * E * K + E) * self.RL
    self.go = self.ib - self.ib % 4
    self.n = (3 * E * K + E + E * K * len(names)) * self.RL
    self.op = eq._implicit
    # Real layout: the stage kernels skip the all-zero part l < m of each field
    real = isinstance(eq.coords.horizontal.spherical_harmonics,
                      spherical_harmonic.RealSphericalHarmonics)
    self.field_size = self.RL if real else 0
    self.degrees = L if real else 0
    # Stage values may leave the entries below the triangle unwritten when
    # every reader skips them: the column-local implicit kernels, and the
    # explicit terms when they run the fused modal kernels (which read
    # l >= m only); the tensor-op path's derivative recurrences would
    # multiply those entries by zero coefficients.
    self.stage_triangle_only = (real and eq.fused_nodal and
                                type(eq).explicit_terms is PrimitiveEquationsSigma.explicit_terms)

  def state(self, buf):
    return State.from_flat(buf, *self.layout)

  def explicit(self, buf):
    out = self.eq.explicit_terms(self.state(buf))
    fb = out.flat_buffer()
    return fb if fb is not None else to_flat(out).flat_buffer()

  def _views(self, full):
    E, K, RL = self.members, self.K, self.RL
    d = full[E * K:2 * E * K]
    t = full[2 * E * K:3 * E * K]
    l = full[3 * E * K:3 * E * K + E]
    return d, t, l

  def _compact_views(self, comp):
    E, K, RL = self.members, self.K, self.RL
    o = self.ib - self.go
    d = comp[o:o + E * K * RL]
    t = comp[o + E * K * RL:o + 2 * E * K * RL]
    l = comp[o + 2 * E * K * RL:o + (2 * E * K + E) * RL]
    return d, t, l

  def new_compact(self, like):
    return torch.empty(self.ie - self.go, dtype=like.dtype, device=like.device)

  def terms(self, buf):
    comp = self.new_compact(buf)
    self.op.apply_into('dino_implicit_terms', 0.0, self.members, *self._views(buf),
                       *self._compact_views(comp))
    return comp

  def terms_adjoint_into(self, gcomp, ybar):
    """ybar[divergence..ln ps] = (implicit terms)^T gcomp; the rest of ybar
    is not written."""
    self.op.apply_into('dino_implicit_terms_adjoint', 0.0, self.members,
                       *self._compact_views(gcomp), *self._views(ybar))

  def terms_adjoint_acc(self, gcomp, ybar):
    """ybar[divergence..ln ps] += (implicit terms)^T gcomp."""
    self.op.apply_into('dino_implicit_terms_adjoint_accumulate', 0.0, self.members,
                       *self._compact_views(gcomp), *self._views(ybar))

  def solve_(self, buf, eta):
    v = self._views(buf)
    self.op.apply_into('dino_implicit_solve', float(eta), self.members, *v, *v)

  def solve_adjoint_(self, buf, eta):
    v = self._views(buf)
    self.op.apply_into('dino_implicit_solve_adjoint', float(eta), self.members, *v, *v)


# ------------------------------------------------------------------ equations


@dataclasses.dataclass
class PrimitiveEquationsSigma:
  """Primitive equations on sigma coordinates (primitive_equations.py:835-1458).

  implicit_inverse_method, vertical_matmul_method: accepted for API
  compatibility; every method gives the same operator here.
  fused_nodal: True runs the nodal nonlinear terms as one fused HIP kernel
  (with a hand-written adjoint); False runs them as separate tensor
  operations, the reference's formulation.
  """
  reference_temperature: np.ndarray
  orography: Array
  coords: coordinate_systems.CoordinateSystem
  physics_specs: units_lib.SimUnits
  vertical_matmul_method: str | None = dataclasses.field(default=None, kw_only=True)
  implicit_inverse_method: str = dataclasses.field(default='split', kw_only=True)
  vertical_advection: Callable | None = dataclasses.field(default=None, kw_only=True)
  include_vertical_advection: bool = dataclasses.field(default=True, kw_only=True)
  humidity_key: str | None = dataclasses.field(default=None, kw_only=True)
  cloud_keys: tuple | None = dataclasses.field(default=None, kw_only=True)
  fused_nodal: bool = dataclasses.field(default=True, kw_only=True)

  def __post_init__(self):
    if not np.allclose(self.coords.horizontal.radius, self.physics_specs.radius, rtol=1e-5):
      raise ValueError('inconsistent radius between coordinates and constants: '
                       f'{self.coords.horizontal.radius=} != {self.physics_specs.radius=}')
    if self.cloud_keys is not None and self.humidity_key is None:
      raise ValueError('cloud_keys requires humidity_key to be set.')
    if self.implicit_inverse_method not in ('split', 'stacked', 'blockwise'):
      raise ValueError(f'invalid implicit_inverse_method {self.implicit_inverse_method}')
    self.reference_temperature = np.asarray(self.reference_temperature)
    self._consts = {}

  # ---- static data on the device
  @functools.cached_property
  def _implicit(self) -> _ImplicitOperator:
    return _ImplicitOperator(self.coords, self.reference_temperature,
                             self.physics_specs.kappa, self.physics_specs.R)

  def _c(self, name, like: torch.Tensor):
    key = (name, like.device, like.dtype)
    t = self._consts.get(key)
    if t is None:
      t = torch.as_tensor(np.asarray(self._host_const(name), np.float64),
                          device=like.device).to(like.dtype)
      self._consts[key] = t
    return t

  def _host_const(self, name):
    grid = self.coords.horizontal
    v = self.coords.vertical
    if name == 'T_ref':
      return self.reference_temperature[:, None, None]
    if name == 'sec2':
      return grid.sec2_lat
    if name == 'coriolis':
      return 2 * self.physics_specs.angular_velocity * grid.nodal_axes[1]
    if name == 'dsig':
      return v.layer_thickness[:, None, None]
    if name == 'c2c_inv':
      return (1 / v.center_to_center)[:, None, None]
    if name == 'sum_sigma':
      return np.cumsum(v.layer_thickness)[:-1, None, None]
    if name == 'alpha':
      return get_sigma_ratios(v)[:, None, None]
    if name == 'geo_weights':
      return get_geopotential_weights_sigma(v, self.physics_specs.R)
    raise KeyError(name)

  @property
  def T_ref(self):
    return self.reference_temperature[..., np.newaxis, np.newaxis]

  @property
  def coriolis_parameter(self):
    _, sin_lat = self.coords.horizontal.nodal_mesh
    return 2 * self.physics_specs.angular_velocity * sin_lat

  # ---- explicit terms
  def _nodal_handle(self, device, tracer_names):
    key = (torch.device(device).index, tuple(tracer_names))
    cache = self.__dict__.setdefault('_nodal_handles', {})
    h = cache.get(key)
    if h is not None:
      return h
    grid = self.coords.horizontal
    v = self.coords.vertical
    ps = self.physics_specs
    K = v.layers
    f64 = lambda a: np.ascontiguousarray(np.asarray(a, np.float64))
    sec2 = f64(grid.sec2_lat)
    cor = f64(2 * ps.angular_velocity * grid.nodal_axes[1])
    dsig = f64(v.layer_thickness)
    c2c = np.zeros(K)
    c2c[:K - 1] = 1 / v.center_to_center
    ss = np.zeros(K)
    ss[:K - 1] = np.cumsum(v.layer_thickness)[:K - 1]
    alpha = f64(get_sigma_ratios(v))
    tref = f64(self.reference_temperature)
    gw = get_geopotential_weights_sigma(v, ps.R)
    geo_off = f64(np.concatenate([[0.0], gw[0, 1:]]))
    geo_diag = f64(np.diag(gw))
    hum = tracer_names.index(self.humidity_key) if self.humidity_key is not None else -1
    clouds = [tracer_names.index(k) for k in (self.cloud_keys or ())]
    cl = (ctypes.c_int * 4)(*(clouds + [0] * (4 - len(clouds))))
    handle = ctypes.c_void_p()
    dp = _lib.dptr
    with torch.cuda.device(device):
      rc = _lib.lib().dino_nodal_create(
          K, grid.longitude_nodes, grid.latitude_nodes, len(tracer_names), hum, cl, len(clouds),
          int(self.include_vertical_advection), dp(sec2), dp(cor), dp(dsig), dp(f64(c2c)),
          dp(f64(ss)), dp(alpha), dp(tref), dp(geo_off), dp(geo_diag), float(ps.R),
          float(ps.kappa), float(ps.R_vapor), float(ps.Cp), float(ps.Cp_vapor),
          ctypes.byref(handle))
    _lib.check(rc, 'dino_nodal_create')
    h = _NodalHandle(handle)
    cache[key] = h
    return h

  def _modal_handle(self, device, names):
    key = (torch.device(device).index, tuple(names))
    cache = self.__dict__.setdefault('_modal_handles', {})
    h = cache.get(key)
    if h is not None:
      return h
    grid = self.coords.horizontal
    R_, L_ = grid.modal_shape
    f64 = lambda a: np.ascontiguousarray(np.asarray(a, np.float64))
    il = f64(spherical_harmonic._grid_host_const(grid, 'inverse_laplacian'))
    ev = f64(grid.laplacian_eigenvalues)
    mask = np.ones(L_)
    mask[-1] = 0.0
    cst = [f64(spherical_harmonic._grid_host_const(grid, n))
           for n in ('cdlat_up', 'cdlat_dn', 'scdlat_up', 'scdlat_dn')]
    hum = names.index(self.humidity_key) if self.humidity_key is not None else -1
    handle = ctypes.c_void_p()
    dp = _lib.dptr
    with torch.cuda.device(device):
      rc = _lib.lib().dino_modal_create(
          R_, L_, self.coords.vertical.layers, len(names), hum, float(grid.radius),
          float(self.physics_specs.g), dp(il), dp(ev), dp(f64(mask)), *(dp(c) for c in cst),
          ctypes.byref(handle))
    _lib.check(rc, 'dino_modal_create')
    h = _Handle(handle, 'dino_modal_destroy')
    cache[key] = h
    return h

  def _check_tracers(self, state):
    names = list(state.tracers.keys())
    if self.humidity_key is not None and self.humidity_key not in state.tracers:
      raise ValueError(f'`{self.humidity_key}` is not found in tracers: {names}.')
    for key in self.cloud_keys or ():
      if key not in state.tracers:
        raise ValueError(f'`{key}` is not found in tracers: {names}.')
    return names

  def _inverse_inputs(self, state, names):
    """Modal fields for the one batched inverse transform, in the order of the
    fused kernel's input groups (spherical_harmonic.py:1164-1178, clip=False)."""
    grid = self.coords.horizontal
    a = grid.radius
    zeta, delta = state.vorticity, state.divergence
    lnps = state.log_surface_pressure
    psi = grid.inverse_laplacian(zeta)
    chi = grid.inverse_laplacian(delta)
    ucos_m = (grid.d_dlon(chi) - grid.cos_lat_d_dlat(psi)) / a
    vcos_m = (grid.cos_lat_d_dlat(chi) + grid.d_dlon(psi)) / a
    inv = [zeta, delta, state.temperature_variation] + [state.tracers[k] for k in names]
    inv += [ucos_m, vcos_m, grid.d_dlon(lnps) / a, grid.cos_lat_d_dlat(lnps) / a]
    if self.humidity_key is not None:
      q_m = state.tracers[self.humidity_key]
      inv += [grid.laplacian(lnps), grid.d_dlon(q_m) / a, grid.cos_lat_d_dlat(q_m) / a]
    return inv

  def _nodal_unfused(self, nod, names, x):
    """The nodal nonlinear terms as separate tensor operations (the
    reference's formulation, primitive_equations.py:168-225, 778-805,
    861-1124); returns the forward-transform inputs in output-group order."""
    ps = self.physics_specs
    moist = self.humidity_key is not None
    zn, dn, tn = nod[0], nod[1], nod[2]
    trn = dict(zip(names, nod[3:3 + len(names)]))
    i0 = 3 + len(names)
    un, vn, gxn, gyn = nod[i0:i0 + 4]
    if moist:
      lap_lnps_n, dqx_n, dqy_n = nod[i0 + 4:i0 + 7]
    sec2 = self._c('sec2', x)
    dsig = self._c('dsig', x)
    c2c = self._c('c2c_inv', x)
    tref = self._c('T_ref', x)
    alpha = self._c('alpha', x)
    udg = (un * gxn + vn * gyn) * sec2
    f_expl = torch.cumsum(udg * dsig, dim=-3)
    f_full = torch.cumsum((dn + udg) * dsig, dim=-3)
    ss = self._c('sum_sigma', x)
    sd_expl = ss * f_expl[..., -1:, :, :] - f_expl[..., :-1, :, :]
    sd_full = ss * f_full[..., -1:, :, :] - f_full[..., :-1, :, :]

    def vadv(w, field):
      return _vadv(w, field, c2c, x)

    total_vort = zn + self._c('coriolis', x)
    nvu = -vn * total_vort * sec2
    nvv = un * total_vort * sec2
    if self.include_vertical_advection:
      sdu = -vadv(sd_full, un)
      sdv = -vadv(sd_full, vn)
    else:
      sdu = sdv = 0.0
    if moist:
      q = trn[self.humidity_key]
      adj = 1 + (ps.R_vapor / ps.R - 1) * q
      for key in self.cloud_keys or ():
        adj = adj - trn[key]
    else:
      adj = 1.0
    rt = ps.R * tn * adj
    comb_u = nvu + (sdu + rt * gxn) * sec2
    comb_v = nvv + (sdv + rt * gyn) * sec2
    ke = (un * un + vn * vn) * sec2 / 2
    dT_v = vadv(sd_full, tn) if self.include_vertical_advection else 0.0
    if np.unique(self.reference_temperature.ravel()).size > 1:
      dT_v = dT_v + vadv(sd_expl, tref)

    def t_omega(tfield, g_term):
      f = torch.cumsum(g_term * dsig, dim=-3)
      af = alpha * f
      af_prev = torch.nn.functional.pad(af, (0, 0, 0, 0, 1, 0))[..., :-1, :, :]
      return tfield * (udg - (af + af_prev) / dsig)

    mean_t = t_omega(tref, udg)
    if not moist:
      dT_a = ps.kappa * (mean_t + t_omega(tn, udg + dn))
    else:
      gcr = ps.R_vapor / ps.R
      hcr = ps.Cp_vapor / ps.Cp
      vtc = tn * ((1 + (gcr - 1) * q) / (1 + (hcr - 1) * q))
      hrc = tref * (((gcr - hcr) * q) / (1 + (hcr - 1) * q))
      dT_a = ps.kappa * (mean_t + t_omega(vtc + hrc, udg + dn))
    lsp_n = -(udg * dsig).sum(dim=-3, keepdim=True)
    temp_n = tn * dn + dT_v + dT_a
    fwd = [comb_u, comb_v, ke, temp_n, un * tn * sec2, vn * tn * sec2, lsp_n]
    for k in names:
      s = trn[k]
      tr_v = vadv(sd_full, s) if self.include_vertical_advection else 0.0
      fwd += [tr_v + s * dn, un * s * sec2, vn * s * sec2]
    if moist:
      coef = tref * (ps.R_vapor - ps.R)
      curl_term = coef * sec2 * (gxn * dqy_n - gyn * dqx_n)
      corr = q * lap_lnps_n * tref * (ps.R_vapor - ps.R)
      dot_term = coef * sec2 * (dqx_n * gxn + dqy_n * gyn)
      tdiff = q * (tn + tref) * (ps.R_vapor / ps.R - 1)
      geo = torch.einsum('gh,...hij->...gij', self._c('geo_weights', x), tdiff)
      fwd += [curl_term, geo, dot_term + corr]
    return fwd

  def _orography_modal(self, x):
    key = ('orography', x.device, x.dtype)
    t = self._consts.get(key)
    if t is None:
      oro = self.orography
      t = (oro.to(x.device, x.dtype) if isinstance(oro, torch.Tensor)
           else torch.as_tensor(np.asarray(oro), device=x.device).to(x.dtype))
      self._consts[key] = t
    return t

  def explicit_terms(self, state: State) -> State:
    """Explicit tendencies (primitive_equations.py:1126-1194)."""
    validate_state_shape(state, self.coords)
    names = self._check_tracers(state)
    # the fused modal kernels address the Real layout's rows (m = 0, +1, -1, ...)
    if self.fused_nodal and isinstance(self.coords.horizontal.spherical_harmonics,
                                       spherical_harmonic.RealSphericalHarmonics):
      return self._explicit_fused(state, names)
    return self._explicit_unfused(state, names)

  def _explicit_fused(self, state: State, names) -> State:
    """prep kernel -> one inverse transform -> fused nodal kernel -> one
    forward transform -> epilogue kernel, on flat (fields, R, L) buffers."""
    grid = self.coords.horizontal
    sh = grid.spherical_harmonics
    flat = to_flat(state)
    buf = flat.flat_buffer()
    lead, K, _ = flat._flat[1:]
    members = int(np.prod(lead)) if lead else 1
    mh = self._modal_handle(buf.device, names)
    prep = _ModalPrep.apply(buf, mh, members)
    # the modal buffers between the transforms and the prep / epilogue kernels
    # are read only on the triangle l >= m: written triangle-only
    own = (type(sh).inverse_transform is spherical_harmonic.RealSphericalHarmonics.inverse_transform
           and type(sh).transform is spherical_harmonic.RealSphericalHarmonics.transform)
    kw_inv = {'_triangle_only_adjoint': True} if own else {}
    kw_fwd = {'_triangle_only': True} if own else {}
    nodal = sh.inverse_transform(prep, **kw_inv)
    nout = _NodalFused.apply(nodal, self._nodal_handle(buf.device, names), members)
    mod = sh.transform(nout, **kw_fwd)
    tend = _ModalEpilogue.apply(mod, mh, members, self._orography_modal(buf).contiguous())
    return State.from_flat(tend, lead, K, names,
                           sim_time=None if state.sim_time is None else 1.0)

  @property
  def _modal_moist(self) -> bool:
    """Whether the humidity curl/divergence terms enter the modal assembly."""
    return self.humidity_key is not None

  def _explicit_unfused(self, state: State, names) -> State:
    grid = self.coords.horizontal
    ps = self.physics_specs
    a = grid.radius
    x = state.vorticity
    moist = self._modal_moist
    inv = self._inverse_inputs(state, names)
    nod = grid.to_nodal_batched(inv)
    mod = grid.to_modal_batched(self._nodal_unfused(nod, names, x))
    cu, cv, ke_m, temp_m, ut_m, vt_m, lsp_m = mod[:7]

    def div_cos(mm, nn):
      return (grid.d_dlon(mm) + grid.sec_lat_d_dlat_cos2(nn)) / a

    def curl_cos(mm, nn):
      return (grid.d_dlon(nn) - grid.sec_lat_d_dlat_cos2(mm)) / a

    vort_t = -curl_cos(cu, cv)
    div_t = -div_cos(cu, cv) - grid.laplacian(ke_m)
    div_t = div_t - ps.g * grid.laplacian(self._orography_modal(x))
    temp_t = temp_m - div_cos(ut_m, vt_m)
    tracers_t = {}
    for n, k in enumerate(names):
      tm, um, vm = mod[7 + 3 * n: 10 + 3 * n]
      tracers_t[k] = tm - div_cos(um, vm)
    if moist:
      j0 = 7 + 3 * len(names)
      curl_m, geo_m, dc_m = mod[j0:j0 + 3]
      vort_t = vort_t + curl_m
      div_t = div_t + (-grid.laplacian(geo_m) - dc_m)
    tend = State(
        vorticity=vort_t,
        divergence=div_t,
        temperature_variation=temp_t,
        log_surface_pressure=lsp_m,
        tracers=tracers_t,
        sim_time=None if state.sim_time is None else 1.0,
    )
    return grid.clip_wavenumbers(tend)

  # ---- implicit terms
  def implicit_terms(self, state: State) -> State:
    """Implicit tendencies (primitive_equations.py:1196-1249), FP64 kernel."""
    validate_state_shape(state, self.coords)
    op = self._implicit
    buf = state.flat_buffer()
    if buf is not None:
      lead, K, names = state._flat[1:]
      members = int(np.prod(lead)) if lead else 1
      out = _ImplicitFlat.apply(buf, op, 0.0, members, len(names), False)
      return State.from_flat(out, lead, K, names,
                             sim_time=None if state.sim_time is None else 0.0)
    d, t, l = _ImplicitTerms.apply(state.divergence, state.temperature_variation,
                                   state.log_surface_pressure, op, 0.0)
    return State(
        vorticity=torch.zeros_like(state.vorticity),
        divergence=d,
        temperature_variation=t,
        log_surface_pressure=l,
        tracers={k: torch.zeros_like(v) for k, v in state.tracers.items()},
        sim_time=None if state.sim_time is None else 0.0,
    )

  def _flat_imex_ops(self, state):
    """Flat-buffer hooks for the fused IMEX step, or None when `state` is not
    flat or a subclass replaces the implicit methods."""
    cls = type(self)
    if (cls.implicit_terms is not PrimitiveEquationsSigma.implicit_terms or
        cls.implicit_inverse is not PrimitiveEquationsSigma.implicit_inverse):
      return None
    buf = state.flat_buffer()
    if buf is None or not buf.is_cuda:
      return None
    validate_state_shape(state, self.coords)
    lead, K, names = state._flat[1:]
    members = int(np.prod(lead)) if lead else 1
    return _FlatImexOps(self, (lead, K, names), members)

  def implicit_inverse(self, state: State, step_size: float) -> State:
    """(1 - step_size * implicit_terms)^{-1} (primitive_equations.py:1251-1458)."""
    if isinstance(step_size, torch.Tensor):
      raise TypeError(
          f'`step_size` must be concrete but a tensor was passed: {step_size}. '
          'Pass a Python float: the inverse is factorised in FP64 for each static step.')
    validate_state_shape(state, self.coords)
    op = self._implicit
    buf = state.flat_buffer()
    if buf is not None:
      lead, K, names = state._flat[1:]
      members = int(np.prod(lead)) if lead else 1
      out = _ImplicitFlat.apply(buf, op, float(step_size), members, len(names), True)
      return State.from_flat(out, lead, K, names, sim_time=state.sim_time)
    d, t, l = _ImplicitSolve.apply(state.divergence, state.temperature_variation,
                                   state.log_surface_pressure, op, float(step_size))
    return State(state.vorticity, d, t, l, dict(state.tracers), sim_time=state.sim_time)

  # ---- reference helper methods kept for API compatibility
  def kinetic_energy_tendency(self, aux_state):
    grid = self.coords.horizontal
    u, v = aux_state.cos_lat_u
    ke = (u * u + v * v) * self._c('sec2', u) / 2
    return -grid.laplacian(grid.to_modal(ke))

  def orography_tendency(self):
    grid = self.coords.horizontal
    oro = self.orography
    return -self.physics_specs.g * grid.laplacian(oro)


class PrimitiveEquations(PrimitiveEquationsSigma):
  """Dry sigma-coordinate equations (deprecated alias, :2534-2562)."""

  def __init__(self, reference_temperature, orography, coords, physics_specs, *,
               vertical_matmul_method=None, implicit_inverse_method='split',
               vertical_advection=None, include_vertical_advection=True):
    super().__init__(reference_temperature, orography, coords, physics_specs,
                     vertical_matmul_method=vertical_matmul_method,
                     implicit_inverse_method=implicit_inverse_method,
                     vertical_advection=vertical_advection,
                     include_vertical_advection=include_vertical_advection,
                     humidity_key=None, cloud_keys=None)


PrimitiveEquationsWithTime = PrimitiveEquations


class MoistPrimitiveEquations(PrimitiveEquationsSigma):
  """Moist equations with specific humidity (deprecated alias, :2568-2596)."""

  def __init__(self, reference_temperature, orography, coords, physics_specs, *,
               vertical_matmul_method=None, implicit_inverse_method='split',
               vertical_advection=None, include_vertical_advection=True):
    super().__init__(reference_temperature, orography, coords, physics_specs,
                     vertical_matmul_method=vertical_matmul_method,
                     implicit_inverse_method=implicit_inverse_method,
                     vertical_advection=vertical_advection,
                     include_vertical_advection=include_vertical_advection,
                     humidity_key='specific_humidity', cloud_keys=None)


class MoistPrimitiveEquationsWithCloudMoisture(PrimitiveEquationsSigma):
  """Moist equations with cloud water species (deprecated alias, :2599-2630)."""

  def __init__(self, reference_temperature, orography, coords, physics_specs, *,
               vertical_matmul_method=None, implicit_inverse_method='split',
               vertical_advection=None, include_vertical_advection=True):
    super().__init__(reference_temperature, orography, coords, physics_specs,
                     vertical_matmul_method=vertical_matmul_method,
                     implicit_inverse_method=implicit_inverse_method,
                     vertical_advection=vertical_advection,
                     include_vertical_advection=include_vertical_advection,
                     humidity_key='specific_humidity',
                     cloud_keys=('specific_cloud_liquid_water_content',
                                 'specific_cloud_ice_water_content'))


# ------------------------------------------------------------------ hybrid levels


def _hybrid_coefficients(a, b, ps, xp=np):
  """Simmons-Burridge coefficients on interfaces p = a + b ps
  (primitive_equations.py:1483-1533): alpha_k = 1 - p_{k-1/2} / dp_k *
  ln(p_{k+1/2} / p_{k-1/2}) and that logarithm, pressures floored at 1e-6 ps
  (the model top has p = 0).  ps: a scalar (a, b 1-D) or a (..., 1, I, J)
  tensor (a, b (K+1, 1, 1) tensors)."""
  p_half = a + b * ps
  if xp is np:
    upper, lower = p_half[:-1], p_half[1:]
    floor = 1e-6 * ps
    log_ratio = np.log(np.maximum(lower, floor) / np.maximum(upper, floor))
  else:
    upper, lower = p_half[..., :-1, :, :], p_half[..., 1:, :, :]
    floor = 1e-6 * ps
    log_ratio = torch.log(torch.maximum(lower, floor) / torch.maximum(upper, floor))
  dp = lower - upper
  return 1 - (upper / dp) * log_ratio, log_ratio


def get_geopotential_weights_hybrid(coordinates, ideal_gas_constant, p_s_ref) -> np.ndarray:
  """R [alpha_k on the diagonal, ln(p ratio)_j at j > k], the coefficients at
  the reference surface pressure (primitive_equations.py:1605-1626)."""
  alpha, log_ratio = _hybrid_coefficients(coordinates.a_boundaries, coordinates.b_boundaries,
                                          p_s_ref)
  k = coordinates.layers
  w = np.triu(np.ones((k, k)), 1) * log_ratio[None, :] + np.diag(alpha)
  return ideal_gas_constant * w


def get_temperature_implicit_weights_hybrid(coordinates, reference_temperature, kappa,
                                            p_s_ref) -> np.ndarray:
  """H of the hybrid implicit temperature tendency (primitive_equations.py:
  1736-1783): kappa T_ref times alpha-weighted lower-triangular sums over
  dp_ref (rows), minus the reference-lapse advection (driven by sigma
  thicknesses), times dp_ref of the divergence level (columns)."""
  tref = np.asarray(reference_temperature)
  if tref.ndim != 1 or tref.shape[-1] != coordinates.layers:
    raise ValueError('`reference_temp` must be a vector of length `coordinates.layers`; '
                     f'got shape {tref.shape} and {coordinates.layers} layers.')
  k = coordinates.layers
  lower = np.tril(np.ones((k, k)))
  alpha, _ = _hybrid_coefficients(coordinates.a_boundaries, coordinates.b_boundaries, p_s_ref)
  pa = lower * alpha[:, None]
  pa_above = np.concatenate([np.zeros((1, k)), pa[:-1]], axis=0)
  dp_ref = coordinates.pressure_thickness + coordinates.sigma_thickness * p_s_ref
  h0 = kappa * tref[:, None] * (pa + pa_above) / dp_ref[:, None]
  lapse = np.append(np.diff(tref) / (dp_ref[:-1] + dp_ref[1:]), 0.0)[:, None]
  adv = lapse * (lower - np.cumsum(coordinates.sigma_thickness)[:, None])
  adv_above = np.concatenate([np.zeros((1, k)), adv[:-1]], axis=0)
  # the last factor scales columns (the divergence level j), as the reference's
  # broadcast of the 1-D thickness does
  return (h0 - adv - adv_above) * dp_ref[None, :]


def _pgf_lps_reference(coordinates, reference_temperature, ideal_gas_constant, p_s_ref):
  """Coefficient of grad ln ps in the pressure-gradient force at the
  reference state, R T_ref / dp [ln(p ratio) b_{k-1/2} + alpha db] ps_ref
  (primitive_equations.py:1859-1888)."""
  alpha, log_ratio = _hybrid_coefficients(coordinates.a_boundaries, coordinates.b_boundaries,
                                          p_s_ref)
  dp = np.maximum(coordinates.layer_thickness(p_s_ref), 1e-6 * p_s_ref)
  weight = log_ratio * coordinates.b_boundaries[:-1] + alpha * coordinates.sigma_thickness
  return ideal_gas_constant * np.asarray(reference_temperature, np.float64) / dp * weight * p_s_ref


def hybrid_vertical_advection(mass_flux, x, pressure_thickness):
  """-(M_{k+1/2} (x_{k+1} - x_k) + M_{k-1/2} (x_k - x_{k-1})) / (2 dp_k), the
  mass flux given at the K - 1 internal interfaces (primitive_equations.py:
  1941-1957, whose flux array also carries the zero end values)."""
  flux = mass_flux * torch.diff(x, dim=-3)
  padded = torch.nn.functional.pad(flux, (0, 0, 0, 0, 1, 1))
  return -0.5 * (padded[..., 1:, :, :] + padded[..., :-1, :, :]) / pressure_thickness


def get_geopotential_diff_hybrid(temperature, coordinates, ideal_gas_constant, surface_pressure):
  """R [sum_{j>k} T_j ln(p ratio)_j + alpha_k T_k] with the coefficients of
  the nodal surface pressure (primitive_equations.py:1629-1674); `coordinates`
  in the units of `surface_pressure`."""
  like = temperature
  vec = lambda v: torch.as_tensor(np.asarray(v, np.float64), device=like.device).to(
      like.dtype).reshape(-1, 1, 1)
  alpha, log_ratio = _hybrid_coefficients(vec(coordinates.a_boundaries),
                                          vec(coordinates.b_boundaries), surface_pressure, torch)
  w = temperature * log_ratio
  below = torch.flip(torch.cumsum(torch.flip(w, (-3,)), dim=-3), (-3,)) - w
  return ideal_gas_constant * (below + alpha * temperature)


def get_geopotential_on_hybrid(temperature, surface_pressure, specific_humidity=None,
                               clouds=None, *, nodal_orography, coordinates,
                               gravity_acceleration, ideal_gas_constant,
                               water_vapor_gas_constant=None):
  """Nodal geopotential from nodal temperature (and moisture) on hybrid levels
  (primitive_equations.py:1677-1733)."""
  if specific_humidity is not None:
    if water_vapor_gas_constant is None:
      raise ValueError('Must provide `water_vapor_gas_constant` with `specific_humidity`.')
    ratio = water_vapor_gas_constant / ideal_gas_constant
    cloud = 0.0 if clouds is None else clouds
    temperature = temperature * (1 + (ratio - 1) * specific_humidity - cloud)
  return nodal_orography * gravity_acceleration + get_geopotential_diff_hybrid(
      temperature, coordinates, ideal_gas_constant, surface_pressure)


class _NodalHybridFused(torch.autograd.Function):
  """(fields_in, I, J) nodal inputs -> (fields_out, I, J) products on hybrid
  levels, by the fused HIP kernel; the backward differentiates the same terms
  written as tensor operations (PrimitiveEquationsHybrid._nodal_terms)."""

  @staticmethod
  def forward(ctx, nodal, h, members, eq, names):
    nodal = nodal.contiguous()
    lib = _lib.lib()
    n_out = lib.dino_nodal_hybrid_fields(h.handle, members, 1)
    n_in = lib.dino_nodal_hybrid_fields(h.handle, members, 0)
    if nodal.shape[0] != n_in:
      raise ValueError(f'hybrid nodal kernel expects {n_in} input fields, got {nodal.shape[0]}')
    out = torch.empty((n_out,) + tuple(nodal.shape[1:]), dtype=nodal.dtype, device=nodal.device)
    code = _lib.DINO_F64 if nodal.dtype == torch.float64 else _lib.DINO_F32
    stream = ctypes.c_void_p(torch.cuda.current_stream(nodal.device).cuda_stream)
    if not _lib.skip_values():  # the backward reads only the saved inputs
      with torch.cuda.device(nodal.device):
        rc = lib.dino_nodal_hybrid_explicit(h.handle, ctypes.c_void_p(nodal.data_ptr()),
                                            ctypes.c_void_p(out.data_ptr()), members, code, stream)
      _lib.check(rc, 'dino_nodal_hybrid_explicit')
      _lib.tally('nodal_hybrid', nodal.element_size() * nodal[0].numel() * (n_in + n_out))
    ctx.save_for_backward(nodal)
    ctx.members, ctx.eq, ctx.names = members, eq, names
    return out

  @staticmethod
  def backward(ctx, g):
    (nodal,) = ctx.saved_tensors
    with torch.enable_grad():
      x = nodal.detach().requires_grad_(True)
      y = ctx.eq._nodal_buffer_terms(x, ctx.members, ctx.names)
      (gin,) = torch.autograd.grad(y, x, g)
    return gin, None, None, None, None


@dataclasses.dataclass
class PrimitiveEquationsHybrid(PrimitiveEquationsSigma):
  """Primitive equations on hybrid sigma-pressure levels
  (primitive_equations.py:1960-2518).

  coords.vertical is a HybridCoordinates with a in hPa (`hpa_quantity`);
  reference_surface_pressure is a pressure quantity (default 101325 Pa).
  Explicit terms: the modal prep kernel, one inverse transform (which also
  carries ln ps, for the columns' surface pressure), the fused hybrid nodal
  kernel, one forward transform and the sigma modal epilogue.  Implicit terms
  and inverse: the FP64 vertical-mode kernels of the sigma equations with
  the hybrid matrices (the block pattern of the implicit operator is the
  same, :1890-1938).  Moisture enters through the virtual temperature only
  (the sigma humidity tendencies are zero here, :2491-2518).
  """
  reference_surface_pressure: Any = dataclasses.field(
      default=101325.0 * units_lib.scales.units.pascal, kw_only=True)
  hpa_quantity: Any = dataclasses.field(default=units_lib.scales.units.hPa, kw_only=True)

  def __post_init__(self):
    super().__post_init__()
    ps = self.physics_specs
    p_ref = ps.nondimensionalize(self.reference_surface_pressure)
    if p_ref <= 0:
      raise ValueError('`reference_surface_pressure` must be positive.')
    self._p_s_ref = float(p_ref)
    from . import hybrid_coordinates  # pylint: disable=import-outside-toplevel
    v = self.coords.vertical
    if not isinstance(v, hybrid_coordinates.HybridCoordinates):
      raise TypeError(f'PrimitiveEquationsHybrid needs HybridCoordinates, got {type(v)}')
    a = np.asarray(ps.nondimensionalize(self.hpa_quantity * np.asarray(v.a_boundaries,
                                                                        np.float64)))
    self.nondim_levels = hybrid_coordinates.HybridCoordinates(a, np.asarray(v.b_boundaries,
                                                                            np.float64))
    self.nondim_coords = coordinate_systems.CoordinateSystem(self.coords.horizontal,
                                                             self.nondim_levels)

  @property
  def p_s_ref(self) -> float:
    return self._p_s_ref

  @property
  def _modal_moist(self) -> bool:
    return False

  @functools.cached_property
  def _implicit(self) -> _ImplicitOperator:
    lev = self.nondim_levels
    ps = self.physics_specs
    G = get_geopotential_weights_hybrid(lev, ps.R, self.p_s_ref)
    H = get_temperature_implicit_weights_hybrid(lev, self.reference_temperature, ps.kappa,
                                                self.p_s_ref)
    c = _pgf_lps_reference(lev, self.reference_temperature, ps.R, self.p_s_ref)
    w = lev.pressure_thickness / self.p_s_ref + lev.sigma_thickness
    return _ImplicitOperator(self.coords, self.reference_temperature, ps.kappa, ps.R,
                             matrices=(G, H, c, w))

  # ---- explicit terms
  def _nodal_handle(self, device, tracer_names):
    key = (torch.device(device).index, tuple(tracer_names))
    cache = self.__dict__.setdefault('_nodal_handles', {})
    h = cache.get(key)
    if h is not None:
      return h
    grid = self.coords.horizontal
    lev = self.nondim_levels
    ps = self.physics_specs
    f64 = lambda a: np.ascontiguousarray(np.asarray(a, np.float64))
    hum = tracer_names.index(self.humidity_key) if self.humidity_key is not None else -1
    clouds = [tracer_names.index(k) for k in (self.cloud_keys or ())]
    cl = (ctypes.c_int * 4)(*(clouds + [0] * (4 - len(clouds))))
    handle = ctypes.c_void_p()
    dp = _lib.dptr
    sec2 = f64(grid.sec2_lat)
    cor = f64(2 * ps.angular_velocity * grid.nodal_axes[1])
    a, b, tref = f64(lev.a_boundaries), f64(lev.b_boundaries), f64(self.reference_temperature)
    with torch.cuda.device(device):
      rc = _lib.lib().dino_nodal_hybrid_create(
          lev.layers, grid.longitude_nodes, grid.latitude_nodes, len(tracer_names), hum, cl,
          len(clouds), int(self.include_vertical_advection), dp(sec2), dp(cor), dp(a), dp(b),
          dp(tref), float(self.p_s_ref), float(ps.R), float(ps.kappa), float(ps.R_vapor),
          float(ps.Cp), float(ps.Cp_vapor), ctypes.byref(handle))
    _lib.check(rc, 'dino_nodal_hybrid_create')
    h = _Handle(handle, 'dino_nodal_hybrid_destroy')
    cache[key] = h
    return h

  def _modal_handle(self, device, names):
    key = (torch.device(device).index, tuple(names))
    cache = self.__dict__.setdefault('_modal_handles', {})
    h = cache.get(key)
    if h is not None:
      return h
    grid = self.coords.horizontal
    R_, L_ = grid.modal_shape
    f64 = lambda a: np.ascontiguousarray(np.asarray(a, np.float64))
    il = f64(spherical_harmonic._grid_host_const(grid, 'inverse_laplacian'))
    ev = f64(grid.laplacian_eigenvalues)
    mask = np.ones(L_)
    mask[-1] = 0.0
    cst = [f64(spherical_harmonic._grid_host_const(grid, n))
           for n in ('cdlat_up', 'cdlat_dn', 'scdlat_up', 'scdlat_dn')]
    hum = names.index(self.humidity_key) if self.humidity_key is not None else -1
    handle = ctypes.c_void_p()
    dp = _lib.dptr
    with torch.cuda.device(device):
      rc = _lib.lib().dino_modal_create_ex(
          R_, L_, self.coords.vertical.layers, len(names), hum, float(grid.radius),
          float(self.physics_specs.g), dp(il), dp(ev), dp(f64(mask)), *(dp(c) for c in cst),
          _lib.MODAL_HYBRID, ctypes.byref(handle))
    _lib.check(rc, 'dino_modal_create_ex')
    h = _Handle(handle, 'dino_modal_destroy')
    cache[key] = h
    return h

  def _explicit_fused(self, state: State, names) -> State:
    """As the sigma path, with the hybrid nodal kernel between the transforms."""
    sh = self.coords.horizontal.spherical_harmonics
    flat = to_flat(state)
    buf = flat.flat_buffer()
    lead, K, _ = flat._flat[1:]
    members = int(np.prod(lead)) if lead else 1
    mh = self._modal_handle(buf.device, names)
    prep = _ModalPrep.apply(buf, mh, members)
    own = (type(sh).inverse_transform is spherical_harmonic.RealSphericalHarmonics.inverse_transform
           and type(sh).transform is spherical_harmonic.RealSphericalHarmonics.transform)
    kw_inv = {'_triangle_only_adjoint': True} if own else {}
    kw_fwd = {'_triangle_only': True} if own else {}
    nodal = sh.inverse_transform(prep, **kw_inv)
    nout = _NodalHybridFused.apply(nodal, self._nodal_handle(buf.device, names), members, self,
                                   tuple(names))
    mod = sh.transform(nout, **kw_fwd)
    tend = _ModalEpilogue.apply(mod, mh, members, self._orography_modal(buf).contiguous())
    return State.from_flat(tend, lead, K, names,
                           sim_time=None if state.sim_time is None else 1.0)

  def _inverse_inputs(self, state, names):
    """The dry sigma inputs plus ln ps itself (order of the nodal kernel's
    input groups)."""
    grid = self.coords.horizontal
    a = grid.radius
    zeta, delta = state.vorticity, state.divergence
    lnps = state.log_surface_pressure
    psi = grid.inverse_laplacian(zeta)
    chi = grid.inverse_laplacian(delta)
    ucos_m = (grid.d_dlon(chi) - grid.cos_lat_d_dlat(psi)) / a
    vcos_m = (grid.cos_lat_d_dlat(chi) + grid.d_dlon(psi)) / a
    inv = [zeta, delta, state.temperature_variation] + [state.tracers[k] for k in names]
    inv += [ucos_m, vcos_m, grid.d_dlon(lnps) / a, grid.cos_lat_d_dlat(lnps) / a, lnps]
    return inv

  def _nodal_unfused(self, nod, names, x):
    i0 = 3 + len(names)
    return self._nodal_terms(nod[0], nod[1], nod[2], dict(zip(names, nod[3:i0])), *nod[i0:i0 + 5])

  def _nodal_buffer_terms(self, nodal, members, names):
    """_nodal_terms on a (fields, I, J) buffer with the fused kernel's groups;
    returns the kernel's output buffer."""
    E, K = members, self.coords.vertical.layers
    I, J = nodal.shape[-2:]
    sizes = [K] * (3 + len(names)) + [K, K, 1, 1, 1]
    parts = torch.split(nodal, [E * n for n in sizes])
    parts = [t.reshape(E, n, I, J) for t, n in zip(parts, sizes)]
    nt = len(names)
    outs = self._nodal_terms(parts[0], parts[1], parts[2], dict(zip(names, parts[3:3 + nt])),
                             *parts[3 + nt:])
    return torch.cat([o.reshape(-1, I, J) for o in outs])

  def _nodal_terms(self, zn, dn, tn, trn, un, vn, gxn, gyn, lnn):
    """Nodal products of the hybrid explicit terms as tensor operations, in
    the fused kernel's output-group order (primitive_equations.py:1536-1602,
    2032-2362): CU CV KE TEMP UT VT LSP (TRN TU TV) per tracer."""
    ps_ = self.physics_specs
    lev = self.nondim_levels
    K = lev.layers
    like = tn
    vec = lambda v: torch.as_tensor(np.asarray(v, np.float64), device=like.device).to(
        like.dtype).reshape(-1, 1, 1)
    sec2 = self._c('sec2', like)
    cor = self._c('coriolis', like)
    tref = self._c('T_ref', like)
    ah, bh = vec(lev.a_boundaries), vec(lev.b_boundaries)
    da, db = vec(lev.pressure_thickness), vec(lev.sigma_thickness)
    aref_np, lref_np = _hybrid_coefficients(lev.a_boundaries, lev.b_boundaries, self.p_s_ref)
    aref, lref = vec(aref_np), vec(lref_np)
    pref = vec(_pgf_lps_reference(lev, self.reference_temperature, ps_.R, self.p_s_ref))
    ps = torch.exp(lnn)
    dp = da + db * ps
    alpha, lr = _hybrid_coefficients(ah, bh, ps, torch)
    udg = (un * gxn + vn * gyn) * sec2
    adv = db * ps * udg
    PA = torch.cumsum(adv, dim=-3)
    PF = torch.cumsum(dp * dn + adv, dim=-3)
    b_int = bh[1:K]
    m_full = b_int * PF[..., -1:, :, :] - PF[..., :-1, :, :]
    m_expl = b_int * PA[..., -1:, :, :] - PA[..., :-1, :, :]

    def vadv(m, field):
      return hybrid_vertical_advection(m, field, dp)

    iva = self.include_vertical_advection
    if self.humidity_key is not None:
      q = trn[self.humidity_key]
      adj = 1 + (ps_.R_vapor / ps_.R - 1) * q
      for key in self.cloud_keys or ():
        adj = adj - trn[key]
    else:
      adj = 1.0
    weight = lr * bh[:-1] + alpha * db
    dpf = torch.maximum(dp, 1e-6 * ps)
    coef = ps_.R * (tn + tref) / dpf * weight * ps * adj - pref
    total_vort = zn + cor
    cu = -vn * total_vort * sec2 + ((-vadv(m_full, un) if iva else 0.0) + coef * gxn) * sec2
    cv = un * total_vort * sec2 + ((-vadv(m_full, vn) if iva else 0.0) + coef * gyn) * sec2
    tv = (tn + tref) * adj
    w = tv * lr
    geo_full = ps_.R * (torch.flip(torch.cumsum(torch.flip(w, (-3,)), -3), (-3,)) - w
                        + alpha * tv)
    lt = lref * tn
    geo_lin = ps_.R * (torch.flip(torch.cumsum(torch.flip(lt, (-3,)), -3), (-3,)) - lt
                       + aref * tn)
    ke = (un * un + vn * vn) * sec2 / 2 + (geo_full - geo_lin)
    vgl = udg * ps * weight / dpf

    def t_omega(tfield, prefix):
      af = aref * prefix
      af_above = torch.nn.functional.pad(af, (0, 0, 0, 0, 1, 0))[..., :-1, :, :]
      return tfield * (vgl - (af + af_above) / dp)

    if self.humidity_key is None:
      tvar = tn
    else:
      gcr, hcr = ps_.R_vapor / ps_.R, ps_.Cp_vapor / ps_.Cp
      den = 1 + (hcr - 1) * q
      tvar = tn * ((1 + (gcr - 1) * q) / den) + tref * (((gcr - hcr) * q) / den)
    temp = tn * dn + ps_.kappa * (t_omega(tref, PA) + t_omega(tvar, PF))
    if iva:
      temp = temp + vadv(m_full, tn)
    if np.unique(self.reference_temperature.ravel()).size > 1:
      temp = temp + vadv(m_expl, tref.expand_as(tn))
    lsp = -((dn * da).sum(dim=-3, keepdim=True) * (1 / ps - 1 / self.p_s_ref)
            + (udg * db).sum(dim=-3, keepdim=True))
    outs = [cu, cv, ke, temp, un * tn * sec2, vn * tn * sec2, lsp]
    for k, s in trn.items():
      trn_t = s * dn + (vadv(m_full, s) if iva else 0.0)
      outs += [trn_t, un * s * sec2, vn * s * sec2]
    return outs
