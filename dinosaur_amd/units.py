"""Nondimensional simulation constants (mirrors dinosaur/units.py:92-199)."""
from __future__ import annotations

import dataclasses

from . import scales


@dataclasses.dataclass(frozen=True)
class SimUnits:
  radius: float
  angular_velocity: float
  gravity_acceleration: float
  ideal_gas_constant: float
  water_vapor_gas_constant: float
  water_vapor_isobaric_heat_capacity: float
  kappa: float
  scale: scales.Scale

  @property
  def R(self) -> float:
    return self.ideal_gas_constant

  @property
  def R_vapor(self) -> float:
    return self.water_vapor_gas_constant

  @property
  def g(self) -> float:
    return self.gravity_acceleration

  @property
  def Cp(self) -> float:
    return self.ideal_gas_constant / self.kappa

  @property
  def Cp_vapor(self) -> float:
    return self.water_vapor_isobaric_heat_capacity

  def nondimensionalize(self, quantity):
    return self.scale.nondimensionalize(quantity)

  def dimensionalize(self, value, unit):
    return self.scale.dimensionalize(value, unit)

  @classmethod
  def from_si(cls,
              radius_si=scales.RADIUS,
              angular_velocity_si=scales.ANGULAR_VELOCITY,
              gravity_acceleration_si=scales.GRAVITY_ACCELERATION,
              ideal_gas_constant_si=scales.IDEAL_GAS_CONSTANT,
              water_vapor_gas_constant_si=scales.IDEAL_GAS_CONSTANT_H20,
              water_vapor_isobaric_heat_capacity_si=scales.WATER_VAPOR_CP,
              kappa_si=scales.KAPPA,
              scale=scales.DEFAULT_SCALE) -> 'SimUnits':
    nd = scale.nondimensionalize
    return cls(nd(radius_si), nd(angular_velocity_si), nd(gravity_acceleration_si),
               nd(ideal_gas_constant_si), nd(water_vapor_gas_constant_si),
               nd(water_vapor_isobaric_heat_capacity_si), nd(kappa_si), scale)
