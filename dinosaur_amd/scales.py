"""Physical constants and the default nondimensionalisation scale.

Mirrors dinosaur/scales.py:38-77 (constants) and :206-228 (DEFAULT_SCALE =
NEURALGCM_V1: length = Earth radius, time = 1 / (2 Omega), mass = 1 kg,
temperature = 1 K).  pint is not available here, so quantities carry their
dimensions as integer exponents of (length, time, mass, temperature).
"""
from __future__ import annotations

import dataclasses

import numpy as np


@dataclasses.dataclass(frozen=True)
class Quantity:
  """A value with SI dimensions length^L time^T mass^M temperature^K."""
  magnitude: object
  length: int = 0
  time: int = 0
  mass: int = 0
  temperature: int = 0

  def __mul__(self, other):
    if isinstance(other, Quantity):
      return Quantity(self.magnitude * other.magnitude, self.length + other.length,
                      self.time + other.time, self.mass + other.mass,
                      self.temperature + other.temperature)
    return dataclasses.replace(self, magnitude=self.magnitude * other)

  __rmul__ = __mul__

  def __truediv__(self, other):
    if isinstance(other, Quantity):
      return Quantity(self.magnitude / other.magnitude, self.length - other.length,
                      self.time - other.time, self.mass - other.mass,
                      self.temperature - other.temperature)
    return dataclasses.replace(self, magnitude=self.magnitude / other)

  def __rtruediv__(self, other):
    return Quantity(other / self.magnitude, -self.length, -self.time, -self.mass,
                    -self.temperature)

  def __pow__(self, p):
    return Quantity(self.magnitude ** p, self.length * p, self.time * p,
                    self.mass * p, self.temperature * p)


class _Units:
  m = meter = Quantity(1.0, length=1)
  s = second = Quantity(1.0, time=1)
  kg = kilogram = Quantity(1.0, mass=1)
  degK = K = Quantity(1.0, temperature=1)
  J = Quantity(1.0, length=2, time=-2, mass=1)
  pascal = Pa = Quantity(1.0, length=-1, time=-2, mass=1)
  hPa = Quantity(100.0, length=-1, time=-2, mass=1)
  day = Quantity(86400.0, time=1)
  hour = Quantity(3600.0, time=1)
  minute = Quantity(60.0, time=1)
  dimensionless = Quantity(1.0)


units = _Units()

RADIUS = 6.37122e6 * units.m
ANGULAR_VELOCITY = OMEGA = 7.292e-5 / units.s
GRAVITY_ACCELERATION = 9.80616 * units.m / units.s ** 2
ISOBARIC_HEAT_CAPACITY = 1004 * units.J / units.kilogram / units.degK
WATER_VAPOR_CP = 1859 * units.J / units.kilogram / units.degK
KAPPA = 2 / 7 * units.dimensionless
IDEAL_GAS_CONSTANT = ISOBARIC_HEAT_CAPACITY * KAPPA
IDEAL_GAS_CONSTANT_H20 = 461.0 * units.J / units.kilogram / units.degK


@dataclasses.dataclass(frozen=True)
class Scale:
  """Scales for (length, time, mass, temperature) in SI base units."""
  length: float
  time: float
  mass: float
  temperature: float

  def _factor(self, q: Quantity) -> float:
    return (self.length ** q.length) * (self.time ** q.time) * (
        self.mass ** q.mass) * (self.temperature ** q.temperature)

  def nondimensionalize(self, quantity):
    if not isinstance(quantity, Quantity):
      return quantity
    return quantity.magnitude / self._factor(quantity)

  def dimensionalize(self, value, unit: Quantity) -> Quantity:
    q = Quantity(np.asarray(value) * self._factor(unit) / unit.magnitude, unit.length,
                 unit.time, unit.mass, unit.temperature)
    return q


NEURALGCM_V1_SCALE = Scale(RADIUS.magnitude, (1 / 2 / OMEGA).magnitude, 1.0, 1.0)
SI_SCALE = Scale(1.0, 1.0, 1.0, 1.0)
DEFAULT_SCALE = NEURALGCM_V1_SCALE
