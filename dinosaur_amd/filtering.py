"""Spectral filters (mirrors dinosaur/filtering.py:27-100)."""
from __future__ import annotations

import numpy as np
import torch

from . import spherical_harmonic


def _preserves_shape(target_shape, scaling_shape):
  return tuple(target_shape) == np.broadcast_shapes(tuple(target_shape), scaling_shape)


def _make_filter_fn(scaling: np.ndarray):
  cache = {}

  def rescale(x):
    if x is None or isinstance(x, (int, float)):
      return x
    if not _preserves_shape(x.shape, scaling.shape):
      return x
    if isinstance(x, torch.Tensor):
      key = (x.device, x.dtype)
      s = cache.get(key)
      if s is None:
        s = torch.as_tensor(scaling, device=x.device).to(x.dtype)
        cache[key] = s
      return x * s
    return (x * scaling).astype(x.dtype)

  def apply(tree):
    flat = getattr(tree, 'flat_buffer', None)
    buf = flat() if flat is not None else None
    if buf is not None and buf.shape[-1] == scaling.shape[-1] and scaling.ndim == 1:
      # every leaf of a flat-backed state ends in the (.., L) axis the
      # scaling broadcasts over: one operation on the whole buffer
      return type(tree).from_flat(rescale(buf), *tree._flat[1:], sim_time=tree.sim_time)
    return spherical_harmonic.tree_map_over_nonscalars(rescale, tree)

  # 1-D (per total wavenumber) scalings may be fused into the producing pass
  apply._degree_scaling = scaling if scaling.ndim == 1 else None

  return apply


def exponential_filter(grid, attenuation=16, order=18, cutoff=0):
  """Damps modes with k = l / l_max > cutoff by
  exp(-attenuation ((k - cutoff) / (1 - cutoff))^(2 order)) (filtering.py:38-89)."""
  _, total_wavenumber = grid.modal_axes
  k = total_wavenumber / total_wavenumber.max()
  a, c, p = attenuation, cutoff, order
  scaling = np.exp((k > c) * (-a * (((k - c) / (1 - c)) ** (2 * p))))
  return _make_filter_fn(scaling)


def horizontal_diffusion_filter(grid, scale, order=1):
  """filtering.py:92-100."""
  eigenvalues = grid.laplacian_eigenvalues
  scaling = np.exp(-scale * (-eigenvalues) ** order)
  return _make_filter_fn(scaling)
