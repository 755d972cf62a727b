"""Horizontal x vertical discretisation (mirrors
dinosaur/coordinate_systems.py:84-199, without SPMD sharding)."""
from __future__ import annotations

import dataclasses
from typing import Any

from . import sigma_coordinates
from . import spherical_harmonic


@dataclasses.dataclass(frozen=True)
class CoordinateSystem:
  horizontal: spherical_harmonic.Grid
  vertical: sigma_coordinates.SigmaCoordinates
  spmd_mesh: Any = None

  def __post_init__(self):
    if self.spmd_mesh is not None:
      raise ValueError('dinosaur_amd runs one coordinate system per GPU; no spmd_mesh')

  @property
  def nodal_shape(self):
    return (self.vertical.layers,) + self.horizontal.nodal_shape

  @property
  def modal_shape(self):
    return (self.vertical.layers,) + self.horizontal.modal_shape

  @property
  def surface_nodal_shape(self):
    return (1,) + self.horizontal.nodal_shape

  @property
  def surface_modal_shape(self):
    return (1,) + self.horizontal.modal_shape

  @property
  def dycore_sharding(self):
    return None

  def asdict(self):
    out = dict(self.horizontal.asdict())
    out.update(self.vertical.asdict())
    out['horizontal_grid_type'] = type(self.horizontal).__name__
    out['vertical_grid_type'] = type(self.vertical).__name__
    return out
