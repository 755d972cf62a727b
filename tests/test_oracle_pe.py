"""Pins the oracle's primitive-equation terms.

Composed explicit_terms have no golden vectors anywhere in the reference, so
they are pinned by physics invariants chosen so a dropped or mis-signed term
fails one of them: JW steady-state balance of the divergence budget, tracer
advection by the JW zonal jet against the closed-form -u . grad q, q = 0 moist
equals dry, the implicit-inverse identity, entrywise G/H/alpha values, and
Held-Suarez closed forms.
"""
import numpy as np
import pytest

from oracle import pe, sht, sigma, states, units

PS = units.SimUnits.from_si()


def _coords(grid, k):
  return pe.Coords(grid, sigma.SigmaCoordinates.equidistant(k))


def test_sigma_ratios_and_geopotential_weights():
  # primitive_equations_test.py:122-136 and :288-315
  c = sigma.SigmaCoordinates.equidistant(10)
  alpha = pe.get_sigma_ratios(c)
  s = c.centers.astype(np.float64)
  for j in range(10):
    exp = -np.log(s[j]) if j == 9 else (np.log(s[j + 1]) - np.log(s[j])) / 2
    assert abs(alpha[j] - exp) < 1e-6  # alpha inherits float32 centres
  g = pe.get_geopotential_weights_sigma(c, 12.3)
  for j in range(10):
    for k in range(10):
      if k < j:
        exp = 0
      elif k == j:
        exp = alpha[j]
      else:
        exp = alpha[k] + alpha[k - 1]
      assert abs(g[j, k] - 12.3 * exp) < 1e-9


def test_temperature_implicit_weights_entrywise():
  # primitive_equations_test.py:345-400: H entries from the docstring formula
  c = sigma.SigmaCoordinates.equidistant(6)
  t = np.linspace(250, 300, 6)
  kappa = 0.3
  h = pe.get_temperature_implicit_weights_sigma(c, t, kappa)
  alpha = pe.get_sigma_ratios(c)
  ds = c.layer_thickness.astype(np.float64)
  n = 6

  def P(i):
    return 1.0 if i >= 0 else 0.0

  def K(r, s):
    if r < 0 or r == n - 1:
      return 0.0
    return (t[r + 1] - t[r]) / (ds[r + 1] + ds[r]) * (P(r - s) - ds[:r + 1].sum())

  for r in range(n):
    for s in range(n):
      first = kappa * t[r] * (P(r - s) * alpha[r] + (P(r - s - 1) * alpha[r - 1] if r > 0 else 0)) / ds[r]
      exp = (first - K(r, s) - K(r - 1, s)) * ds[s]
      assert abs(h[r, s] - exp) < 1e-7 * max(1, abs(exp)), (r, s)


@pytest.mark.parametrize('method_eta', [0.3, 0.1])
def test_implicit_inverse_identity(method_eta):
  # primitive_equations_test.py:578-658
  grid = sht.Grid.with_wavenumbers(16)
  coords = _coords(grid, 5)
  ps = units.SimUnits.from_si(ideal_gas_constant_si=33.0, kappa_si=1.4)
  rs = np.random.RandomState(0)
  scale = (grid.total_wavenumbers + 1) ** -2
  st = pe.State(*(scale * rs.normal(size=s) for s in
                  [coords.modal_shape] * 3 + [coords.surface_modal_shape]))
  eq = pe.PrimitiveEquationsSigma(np.linspace(100, 200, 5),
                                  np.zeros(grid.modal_shape), coords, ps)
  im = eq.implicit_terms(st)
  inv = eq.implicit_inverse(st - method_eta * im, method_eta)
  for f in pe.State.FIELDS:
    np.testing.assert_allclose(getattr(inv, f), getattr(st, f), atol=1e-12)
  with pytest.raises(TypeError):
    eq.implicit_inverse(st, np.array(0.3))


def _jw(grid, k):
  coords = _coords(grid, k)
  st, oro, reft, _ = states.steady_state_jw(coords, PS)
  morog = pe.truncated_modal_orography(oro, coords)
  return coords, st, morog, reft


def test_jw_steady_state_divergence_budget_balances():
  # primitive_equations_test.py:186-276 (stationarity).  The explicit
  # divergence tendency is O(0.1); explicit + implicit must cancel it to the
  # vertical discretisation error.  Vorticity, temperature and log-pressure
  # tendencies of the zonally symmetric, non-divergent state vanish.
  coords, st, morog, reft = _jw(sht.Grid.with_wavenumbers(42), 26)
  eq = pe.PrimitiveEquationsSigma(reft, morog, coords, PS)
  ex = eq.explicit_terms(st)
  tot = ex + eq.implicit_terms(st)
  assert np.abs(ex.divergence).max() > 0.05
  assert np.abs(tot.divergence).max() < 1e-3
  for f in ('vorticity', 'temperature_variation', 'log_surface_pressure'):
    assert np.abs(getattr(tot, f)).max() < 1e-11
  # dropping the orography term breaks the balance by O(0.1)
  eq_flat = pe.PrimitiveEquationsSigma(reft, np.zeros_like(morog), coords, PS)
  bad = eq_flat.explicit_terms(st) + eq_flat.implicit_terms(st)
  assert np.abs(bad.divergence).max() > 1e-2


def test_tracer_advection_by_jw_jet_matches_closed_form():
  grid = sht.Grid.with_wavenumbers(85)
  coords, st, morog, reft = _jw(grid, 4)
  q = states.gaussian_scalar(coords, PS, perturbation_radius=0.3)
  st.tracers = {'q': q}
  eq = pe.PrimitiveEquationsSigma(reft, morog, coords, PS)
  tend = eq.explicit_terms(st).tracers['q']
  # closed form: dq/dt = -u/(a cos) dq/dlon with the JW zonal wind
  aux = pe.compute_diagnostic_state_sigma(st, coords)
  ucos = aux.cos_lat_u[0]
  qn = grid.to_nodal(q)
  dq_dlon = grid.to_nodal(grid.d_dlon(q))
  exact = -ucos * dq_dlon / (grid.cos_lat ** 2) / PS.radius
  got = grid.to_nodal(tend)
  scale = np.abs(exact).max()
  assert scale > 1e-3
  assert np.abs(got - exact).max() / scale < 2e-3


def test_moist_with_zero_humidity_equals_dry():
  # primitive_equations_test.py:660-703
  coords, st, morog, reft = _jw(sht.Grid.T21(), 4)
  st = st + states.baroclinic_perturbation_jw(coords, PS)
  st.tracers = {'specific_humidity': states.gaussian_scalar(coords, PS, amplitude=0.0)}
  a = pe.PrimitiveEquationsSigma(reft, morog, coords, PS).explicit_terms(st)
  b = pe.PrimitiveEquationsSigma(reft, morog, coords, PS,
                                 humidity_key='specific_humidity').explicit_terms(st)
  for f in pe.State.FIELDS:
    np.testing.assert_allclose(getattr(a, f), getattr(b, f), atol=1e-12)
  # and with humidity the terms do change the tendencies
  st.tracers = {'specific_humidity': states.gaussian_scalar(coords, PS, amplitude=0.01)}
  c = pe.PrimitiveEquationsSigma(reft, morog, coords, PS,
                                 humidity_key='specific_humidity').explicit_terms(st)
  assert np.abs(c.divergence - a.divergence).max() > 1e-6


def test_held_suarez_closed_forms():
  # held_suarez.py:99-167 and held_suarez_test.py:168-196
  grid = sht.Grid.T21()
  coords = _coords(grid, 8)
  st, oro, reft = states.isothermal_rest_atmosphere(coords, PS)
  hs = pe.HeldSuarezForcingSigma(coords, PS, reft)
  s = coords.vertical.centers.astype(np.float64)
  kv = hs.kv()[:, 0, 0]
  np.testing.assert_allclose(kv, units.per_day(1) * np.maximum(0, (s - 0.7) / 0.3), rtol=1e-6)
  out = hs.explicit_terms(st)
  assert np.all(out.log_surface_pressure == 0)
  # at rest: only Newtonian cooling toward T_eq acts
  assert np.abs(out.vorticity).max() < 1e-14
  lat = hs.lat
  p = s[:, None, None] * units.pascal(1e5) / units.pascal(1e5)
  teq = np.maximum(200., p ** PS.kappa * (315 - 60 * np.sin(lat) ** 2 - 10 * np.log(p) * np.cos(lat) ** 2))
  exp = grid.to_modal(-hs.kt() * (288.0 - teq))
  np.testing.assert_allclose(out.temperature_variation, exp, atol=1e-12)
