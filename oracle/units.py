"""Physical constants and nondimensionalisation (oracle restatement).

Restates dinosaur/scales.py:38-77 (constants) and :206-228 (DEFAULT_SCALE =
NEURALGCM_V1: length = Earth radius, time = 1/(2 Omega), mass = 1 kg,
temperature = 1 K), and dinosaur/units.py:92-199 (SimUnits.from_si).  pint is
absent here, so dimensions are given explicitly as exponents of
(length, time, mass, temperature).

Pinned by: dt(20 min) = 0.175008 and a = dt/tau = 16.0 (time_integration.py:
430, 446-447), Omega_nondim = 0.5 exactly (definition of the time unit).
"""
from __future__ import annotations

import dataclasses

RADIUS_M = 6.37122e6                 # scales.py:38
OMEGA_PER_S = 7.292e-5               # scales.py:41
GRAVITY_SI = 9.80616                 # scales.py:44
CP_SI = 1004.0                       # scales.py:47
CP_VAPOR_SI = 1859.0                 # scales.py:52
KAPPA = 2 / 7                        # scales.py:61
R_SI = CP_SI * KAPPA                 # scales.py:70
R_VAPOR_SI = 461.0                   # scales.py:74

@dataclasses.dataclass(frozen=True)
class Scale:
  """Units of length, time, mass and temperature (scales.py:186-228)."""
  length: float = RADIUS_M
  time: float = 1 / 2 / OMEGA_PER_S      # scales.py:206-211
  mass: float = 1.0
  temperature: float = 1.0

  def factor(self, length=0, time=0, mass=0, temperature=0) -> float:
    return ((self.length ** length) * (self.time ** time) * (self.mass ** mass)
            * (self.temperature ** temperature))


DEFAULT_SCALE = Scale()
LENGTH_UNIT = DEFAULT_SCALE.length
TIME_UNIT = DEFAULT_SCALE.time
MASS_UNIT = DEFAULT_SCALE.mass
TEMPERATURE_UNIT = DEFAULT_SCALE.temperature


def nondim(value: float, length=0, time=0, mass=0, temperature=0,
           scale: Scale = DEFAULT_SCALE) -> float:
  """value_SI / scale**dims (scales.py:168-181)."""
  return value / scale.factor(length, time, mass, temperature)


def dim(value, length=0, time=0, mass=0, temperature=0, scale: Scale = DEFAULT_SCALE):
  return value * scale.factor(length, time, mass, temperature)


def seconds(s):
  return nondim(s, time=1)


def pascal(p):
  # Pa = kg m^-1 s^-2
  return nondim(p, length=-1, time=-2, mass=1)


def meters(m):
  return nondim(m, length=1)


def kelvin(t):
  return nondim(t, temperature=1)


def per_day(n_days):
  return nondim(1.0 / (n_days * 86400.0), time=-1)


@dataclasses.dataclass(frozen=True)
class SimUnits:
  """units.py:92-199 (nondimensional constants)."""
  radius: float
  angular_velocity: float
  gravity_acceleration: float
  ideal_gas_constant: float
  water_vapor_gas_constant: float
  water_vapor_isobaric_heat_capacity: float
  kappa: float

  @property
  def R(self):
    return self.ideal_gas_constant

  @property
  def R_vapor(self):
    return self.water_vapor_gas_constant

  @property
  def g(self):
    return self.gravity_acceleration

  @property
  def Cp(self):
    return self.ideal_gas_constant / self.kappa

  @property
  def Cp_vapor(self):
    return self.water_vapor_isobaric_heat_capacity

  scale: Scale = DEFAULT_SCALE

  def nondim(self, value, **dims):
    return nondim(value, scale=self.scale, **dims)

  def dim(self, value, **dims):
    return dim(value, scale=self.scale, **dims)

  @classmethod
  def from_si(cls, ideal_gas_constant_si=None, kappa_si=None, scale: Scale = DEFAULT_SCALE):
    kappa = KAPPA if kappa_si is None else kappa_si
    r_si = (CP_SI * KAPPA) if ideal_gas_constant_si is None else ideal_gas_constant_si
    nd = lambda v, **d: nondim(v, scale=scale, **d)  # noqa: E731
    return cls(
        radius=nd(RADIUS_M, length=1),
        angular_velocity=nd(OMEGA_PER_S, time=-1),
        gravity_acceleration=nd(GRAVITY_SI, length=1, time=-2),
        ideal_gas_constant=nd(r_si, length=2, time=-2, temperature=-1),
        water_vapor_gas_constant=nd(R_VAPOR_SI, length=2, time=-2, temperature=-1),
        water_vapor_isobaric_heat_capacity=nd(CP_VAPOR_SI, length=2, time=-2, temperature=-1),
        kappa=kappa,
        scale=scale,
    )
