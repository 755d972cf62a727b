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


class _Oscillator:
  """x0 grows linearly, (x1, x2) is a harmonic oscillator of period 2 pi."""

  def explicit_terms(self, x):
    return np.array([0.1, x[2], -x[1]])

  def implicit_terms(self, x):
    return np.zeros_like(x)

  def implicit_inverse(self, x, eta):
    return x


def test_dfi_known_answers():
  """The reference's DFI known answers (time_integration_test.py:305-335)."""
  acc = ti.accumulate_repeated(lambda x: 2 * x, np.arange(4), 1)
  assert acc == 0 * 2 + 1 * 4 + 2 * 8 + 3 * 16
  w = ti.dfi_lanczos_weights(10, 10, 0.1)
  assert w.size == 50 and (np.diff(w) < 0).all()
  assert abs(w[0] - 1.0) < 5e-3 and abs(w[-1]) < 5e-3
  dfi = ti.digital_filter_initialization(_Oscillator(), ti.imex_rk_sil3, [], 20.0, 20.0, 0.01)
  np.testing.assert_allclose(dfi(np.array([1.0, 1.0, 1.0])), [1.0, 0.0, 0.0], atol=1e-3)


def test_dfi_filter_response_closed_form():
  """On the oscillator the DFI output is the filter's discrete response
  H(w) = w0 + 2 sum_n w_n cos(n w dt) times the input (the time-symmetric
  weights cancel the sine parts), to the SIL3 phase error; the linear mode
  is unchanged.  A planted sign error in the reversed equation fails."""
  dt, span, cutoff = 0.05, 6.0, 6.0
  dfi = ti.digital_filter_initialization(_Oscillator(), ti.imex_rk_sil3, [], span, cutoff, dt)
  out = dfi(np.array([2.0, 1.0, 0.0]))
  w = ti.dfi_lanczos_weights(span, cutoff, dt)
  total = 1 + 2 * w.sum()
  n = np.arange(1, w.size + 1)
  h = (1 + 2 * (w * np.cos(n * dt)).sum()) / total
  np.testing.assert_allclose(out, [2.0, h, 0.0], atol=5e-6)
  orig = ti.TimeReversed.explicit_terms
  try:
    ti.TimeReversed.explicit_terms = lambda self, x: self.forward_eq.explicit_terms(x)
    bad = dfi(np.array([2.0, 1.0, 0.0]))
  finally:
    ti.TimeReversed.explicit_terms = orig
  assert abs(bad[0] - 2.0) > 1e-2 or abs(bad[1] - h) > 1e-3


def test_horizontal_diffusion_filter_closed_form():
  """The top degree decays by exp(-dt / tau) per step, degree 0 is kept."""
  from oracle import sht
  g = sht.Grid.T21()
  f = ti.horizontal_diffusion_step_filter(g, dt=0.3, tau=2.0, order=2)
  x = np.ones(g.modal_shape)
  y = f(None, x)
  np.testing.assert_allclose(y[:, -1], np.exp(-0.3 / 2.0), rtol=1e-12)
  np.testing.assert_allclose(y[:, 0], 1.0)
