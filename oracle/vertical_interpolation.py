"""Vertical regridding between hybrid and sigma levels (oracle restatement;
test infrastructure only).

Restates dinosaur/vertical_interpolation.py:
  _extrapolate_left/_right/_both, _linear_interp_with_safe_extrap  :82-102
  interp_hybrid_to_sigma                                           :303-326
  _interval_overlap, conservative_regrid_weights                  :329-364
  regrid_hybrid_to_sigma                                           :366-401
  regrid_hybrid_to_hybrid                                          :404-440
Column by column with NumPy (small cases only).

Pinned in tests/test_oracle_vertical_interpolation.py by the reference's
known answers (vertical_interpolation_test.py:168-238), row sums of the
weights, conservation of the column integral, identity regridding onto the
same levels, and planted mistakes.
"""
from __future__ import annotations

import numpy as np


def interval_overlap(source_bounds, target_bounds):
  """(target, source) lengths of the overlaps of the cells (:329-341)."""
  s = np.asarray(source_bounds, np.float64)
  t = np.asarray(target_bounds, np.float64)
  upper = np.minimum(t[1:, None], s[None, 1:])
  lower = np.maximum(t[:-1, None], s[None, :-1])
  return np.maximum(upper - lower, 0)


def conservative_regrid_weights(source_bounds, target_bounds):
  """Overlaps normalised per target cell; rows with no overlap are zero
  (nan_to_num, :344-364)."""
  w = interval_overlap(source_bounds, target_bounds)
  with np.errstate(invalid='ignore', divide='ignore'):
    w = np.nan_to_num(w / w.sum(axis=1, keepdims=True))
  return w


def _columns(field, surface_pressure):
  f = np.asarray(field, np.float64)
  ps = np.broadcast_to(np.asarray(surface_pressure, np.float64), f.shape[-2:])
  return f, ps


def regrid_hybrid_to_sigma(field, hybrid_coords, sigma_coords, surface_pressure):
  """field (..., K_hybrid, X, Y) -> (..., K_sigma, X, Y) by conservative
  regridding in sigma = a / ps + b (:366-401)."""
  f, ps = _columns(field, surface_pressure)
  if f.shape[-3] != hybrid_coords.layers:
    raise ValueError(f'Source has {hybrid_coords.layers} layers, but field has {f.shape[-3]}')
  out = np.zeros(f.shape[:-3] + (sigma_coords.layers,) + f.shape[-2:])
  a = np.asarray(hybrid_coords.a_boundaries, np.float64)
  b = np.asarray(hybrid_coords.b_boundaries, np.float64)
  tb = np.asarray(sigma_coords.boundaries, np.float64)
  for x in range(f.shape[-2]):
    for y in range(f.shape[-1]):
      w = conservative_regrid_weights(a / ps[x, y] + b, tb)
      out[..., :, x, y] = np.einsum('ab,...b->...a', w, f[..., :, x, y])
  return out


def regrid_hybrid_to_hybrid(field, source_coords, target_coords, surface_pressure):
  """Conservative regridding in pressure a + b ps (:404-440)."""
  f, ps = _columns(field, surface_pressure)
  if f.shape[-3] != source_coords.layers:
    raise ValueError(f'Source has {source_coords.layers} layers, but field has {f.shape[-3]}')
  out = np.zeros(f.shape[:-3] + (target_coords.layers,) + f.shape[-2:])
  sa, sb = (np.asarray(v, np.float64) for v in (source_coords.a_boundaries,
                                                source_coords.b_boundaries))
  ta, tb = (np.asarray(v, np.float64) for v in (target_coords.a_boundaries,
                                                target_coords.b_boundaries))
  for x in range(f.shape[-2]):
    for y in range(f.shape[-1]):
      w = conservative_regrid_weights(sa + sb * ps[x, y], ta + tb * ps[x, y])
      out[..., :, x, y] = np.einsum('ab,...b->...a', w, f[..., :, x, y])
  return out


def _extrapolate_both(y):
  y = np.concatenate([[y[0] - (y[1] - y[0])], y])
  return np.concatenate([y, [y[-1] + (y[-1] - y[-2])]])


def linear_interp_with_safe_extrap(x, xp, fp, n=1):
  """np.interp after extending (xp, fp) linearly by n cells at each end; NaN
  beyond (:96-102)."""
  xp = np.asarray(xp, np.float64)
  fp = np.asarray(fp, np.float64)
  for _ in range(n):
    xp = _extrapolate_both(xp)
    fp = _extrapolate_both(fp)
  return np.interp(x, xp, fp, left=np.nan, right=np.nan)


def interp_hybrid_to_sigma(field, hybrid_coords, sigma_coords, surface_pressure):
  """Linear interpolation from hybrid-level centres (in sigma at the column's
  ps) to sigma-level centres (:303-326)."""
  f, ps = _columns(field, surface_pressure)
  out = np.zeros(f.shape[:-3] + (sigma_coords.layers,) + f.shape[-2:])
  a = np.asarray(hybrid_coords.a_boundaries, np.float64)
  b = np.asarray(hybrid_coords.b_boundaries, np.float64)
  tc = np.asarray(sigma_coords.centers, np.float64)
  for x in range(f.shape[-2]):
    for y in range(f.shape[-1]):
      s = a / ps[x, y] + b
      sc = (s[1:] + s[:-1]) / 2
      for idx in np.ndindex(f.shape[:-3]):
        out[idx + (slice(None), x, y)] = linear_interp_with_safe_extrap(
            tc, sc, f[idx + (slice(None), x, y)])
  return out
