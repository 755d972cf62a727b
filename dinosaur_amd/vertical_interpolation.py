"""Vertical regridding between hybrid and sigma levels on the GPU (mirrors
dinosaur/vertical_interpolation.py:303-440): conservative regridding in
sigma (regrid_hybrid_to_sigma) or pressure (regrid_hybrid_to_hybrid) and
linear interpolation between layer centres (interp_hybrid_to_sigma), by the
per-column HIP kernel of csrc/regrid.hip.  Host-side interval helpers
(interval_overlap, conservative_regrid_weights) keep the reference's NumPy
forms for callers that want the weight matrices.

Leaves are nodal tensors (..., K, X, Y); scalars and leaves of fewer than
three dimensions pass through.  surface_pressure is (X, Y), (1, X, Y) or
carries the leaves' leading dimensions, in the units of the hybrid
coordinates' a (hPa for the reference's level tables).
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib
from . import hybrid_coordinates
from . import sigma_coordinates


def interval_overlap(source_bounds, target_bounds) -> np.ndarray:
  """(target, source) overlap lengths of the cells (:329-341)."""
  s = np.asarray(source_bounds, np.float64)
  t = np.asarray(target_bounds, np.float64)
  return np.maximum(np.minimum(t[1:, None], s[None, 1:]) - np.maximum(t[:-1, None], s[None, :-1]),
                    0)


def conservative_regrid_weights(source_bounds, target_bounds) -> np.ndarray:
  """Overlaps normalised per target cell, rows without overlap zero (:344-364)."""
  w = interval_overlap(source_bounds, target_bounds)
  with np.errstate(invalid='ignore', divide='ignore'):
    return np.nan_to_num(w / w.sum(axis=1, keepdims=True))


class _Regridder:
  """Owns a dino_vregrid of one (mode, source levels, target levels)."""

  def __init__(self, mode, sa, sb, ta, tb):
    f64 = lambda v: np.ascontiguousarray(np.asarray(v, np.float64))
    self.sa, self.sb, self.ta, self.tb = f64(sa), f64(sb), f64(ta), f64(tb)
    self.Ks, self.Kt = len(self.sa) - 1, len(self.ta) - 1
    self.mode = mode
    self.handle = ctypes.c_void_p()
    dp = _lib.dptr
    _lib.check(_lib.lib().dino_vregrid_create(mode, self.Ks, dp(self.sa), dp(self.sb), self.Kt,
                                              dp(self.ta), dp(self.tb),
                                              ctypes.byref(self.handle)), 'dino_vregrid_create')

  def __del__(self):
    try:
      if self.handle.value:
        _lib.release('dino_vregrid_destroy', self.handle)
    except Exception:  # pylint: disable=broad-except  (interpreter shutdown)
      pass

  def __call__(self, x: torch.Tensor, surface_pressure) -> torch.Tensor:
    if x.shape[-3] != self.Ks:
      raise ValueError(f'Source has {self.Ks} layers, but field has {x.shape[-3]}')
    if not x.is_cuda:
      raise ValueError('vertical regridding runs on the GPU: pass device tensors')
    if x.dtype not in (torch.float32, torch.float64):
      raise TypeError(f'unsupported dtype {x.dtype}')
    lead, cols_shape = x.shape[:-3], x.shape[-2:]
    cols = int(np.prod(cols_shape))
    n_fields = int(np.prod(lead)) if lead else 1
    ps = torch.as_tensor(surface_pressure, device=x.device).to(x.dtype)
    if ps.shape[-2:] != cols_shape:
      ps = ps.expand(cols_shape)
    ps = ps.reshape(-1, cols).contiguous()
    n_ps = ps.shape[0]
    if n_fields % n_ps:
      raise ValueError(f'surface pressure with {n_ps} slices does not match leading shape {lead}')
    src = x.contiguous()
    out = torch.empty(lead + (self.Kt,) + tuple(cols_shape), dtype=x.dtype, device=x.device)
    code = _lib.DINO_F64 if x.dtype == torch.float64 else _lib.DINO_F32
    stream = ctypes.c_void_p(torch.cuda.current_stream(x.device).cuda_stream)
    per = n_fields // n_ps  # fields per surface-pressure slice
    # the kernel's grid takes <= 65535 fields; a chunk covers whole slices
    chunk = 65535 if n_ps == 1 else (65535 // per) * per
    if chunk == 0:
      raise ValueError(f'{per} fields per surface-pressure slice exceed 65535')
    esz = x.element_size()
    with torch.cuda.device(x.device):
      for f0 in range(0, n_fields, chunk):
        f1 = min(n_fields, f0 + chunk)
        rc = _lib.lib().dino_vregrid_apply(
            self.handle, ctypes.c_void_p(src.data_ptr() + f0 * self.Ks * cols * esz),
            ctypes.c_void_p(ps.data_ptr() + (0 if n_ps == 1 else f0 // per) * cols * esz),
            ctypes.c_void_p(out.data_ptr() + f0 * self.Kt * cols * esz),
            f1 - f0, 1 if n_ps == 1 else (f1 - f0) // per, cols, code, stream)
        _lib.check(rc, 'dino_vregrid_apply')
    return out


def _map_nonscalar(fn, tree):
  if isinstance(tree, dict):
    return {k: _map_nonscalar(fn, v) for k, v in tree.items()}
  if isinstance(tree, (list, tuple)):
    return type(tree)(_map_nonscalar(fn, v) for v in tree)
  if isinstance(tree, torch.Tensor) and tree.ndim >= 3:
    return fn(tree)
  return tree


def regrid_hybrid_to_sigma(fields, hybrid_coords: hybrid_coordinates.HybridCoordinates,
                           sigma_coords: sigma_coordinates.SigmaCoordinates, surface_pressure):
  """Conservative regridding from hybrid to sigma levels in sigma = a / ps + b
  (vertical_interpolation.py:366-401)."""
  r = _Regridder(_lib.VREGRID_CONSERVATIVE_SIGMA, hybrid_coords.a_boundaries,
                 hybrid_coords.b_boundaries, np.zeros(sigma_coords.layers + 1),
                 sigma_coords.boundaries)
  return _map_nonscalar(lambda x: r(x, surface_pressure), fields)


def regrid_hybrid_to_hybrid(fields, source_coords: hybrid_coordinates.HybridCoordinates,
                            target_coords: hybrid_coordinates.HybridCoordinates,
                            surface_pressure):
  """Conservative regridding between hybrid levels in pressure
  (vertical_interpolation.py:404-440)."""
  r = _Regridder(_lib.VREGRID_CONSERVATIVE_PRESSURE, source_coords.a_boundaries,
                 source_coords.b_boundaries, target_coords.a_boundaries,
                 target_coords.b_boundaries)
  return _map_nonscalar(lambda x: r(x, surface_pressure), fields)


def interp_hybrid_to_sigma(fields, hybrid_coords: hybrid_coordinates.HybridCoordinates,
                           sigma_coords: sigma_coordinates.SigmaCoordinates, surface_pressure):
  """Linear interpolation from hybrid layer centres to sigma layer centres,
  one extrapolated cell at each end, NaN beyond (vertical_interpolation.py:
  303-326)."""
  # the target centres as the sigma coordinates hold them (float32 midpoints)
  centres = np.append(np.asarray(sigma_coords.centers, np.float64), 0.0)
  r = _Regridder(_lib.VREGRID_LINEAR_SIGMA, hybrid_coords.a_boundaries,
                 hybrid_coords.b_boundaries, centres, sigma_coords.boundaries)
  return _map_nonscalar(lambda x: r(x, surface_pressure), fields)
