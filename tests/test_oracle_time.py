"""Pins the oracle's SIL3 stepper and filters with closed forms
(time_integration_test.py:60-206, filtering_test.py:38-85)."""
import numpy as np

from oracle import sht, time_integration as ti


class _Linear:
  """du/dt = a u (explicit) + b u (implicit); exact solution exp((a+b)t)."""

  def __init__(self, a, b):
    self.a, self.b = a, b

  def explicit_terms(self, u):
    return self.a * u

  def implicit_terms(self, u):
    return self.b * u

  def implicit_inverse(self, u, eta):
    return u / (1 - eta * self.b)


def test_sil3_accuracy_and_order():
  eq = _Linear(-0.7 + 1.3j, -2.0)
  errs = []
  for n in (20, 40, 80):
    dt = 1.0 / n
    step = ti.imex_rk_sil3(eq, dt)
    u = np.array(1.0 + 0j)
    for _ in range(n):
      u = step(u)
    errs.append(abs(u - np.exp(-0.7 + 1.3j - 2.0)))
  # second order overall (implicit part), consistent to < 1e-3 at dt = 1/20
  assert errs[0] < 1e-3
  assert errs[0] / errs[1] > 3.5 and errs[1] / errs[2] > 3.5


def test_sil3_explicit_only_is_third_order():
  eq = _Linear(-1.0 + 2.0j, 0.0)
  errs = []
  for n in (16, 32):
    step = ti.imex_rk_sil3(eq, 1.0 / n)
    u = np.array(1.0 + 0j)
    for _ in range(n):
      u = step(u)
    errs.append(abs(u - np.exp(-1.0 + 2.0j)))
  assert errs[0] / errs[1] > 7.0


def test_exponential_filter_kat():
  # filtering_test.py:38-43 (TL127)
  g = sht.Grid.TL127()
  s = ti.exponential_filter_scaling(g)
  assert s[0] == 1
  assert abs(s[-1] - 1.125e-7) < 1e-9
  f = ti.exponential_filter(g)
  x = np.ones(g.modal_shape)
  np.testing.assert_allclose(f(x)[3], s)
  # scalar leaves (e.g. sim_time) pass through unscaled (filtering.py:27-35)
  y = np.float64(3.0)
  assert f(y) == 3.0
  y = np.ones((1, 1))
  assert f(y) is y
  # step filter at dt = 20 min equals attenuation 16
  step = ti.exponential_step_filter(g, 0.175008)
  np.testing.assert_allclose(step(None, x)[0], np.exp(-16 * (np.arange(129) / 128) ** 36), rtol=1e-6)
