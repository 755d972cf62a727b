"""Pins the oracle's hybrid sigma-pressure coordinates and primitive equations.

* the reference's own known answers for the coordinate constructors and the
  vertical integrals and difference (hybrid_coordinates_test.py:142-365):
  second-order convergence to closed-form integrals;
* the Simmons-Burridge column balance: for an isothermal column the
  geopotential's response to ln ps plus the pressure-gradient coefficient is
  R T exactly, so a resting isothermal atmosphere over balanced orography
  feels no force (pins alpha_k, the log-pressure ratios, the geopotential
  weights and the pressure-gradient coefficient);
* hybrid levels built from sigma levels reproduce the sigma equations'
  vorticity and log-pressure tendencies to round-off, and their divergence
  tendency differs from sigma's by exactly the Laplacian of the difference of
  the two geopotential discretisations (primitive_equations_test.py:785-854
  asserts closeness only);
* the explicit temperature tendency of a uniform-T' column driven by a
  single-harmonic divergence converges to the continuous value;
* the log-pressure tendency is the column mass continuity;
* the implicit-inverse identity (primitive_equations_test.py:856-895).
Planted mistakes in the oracle fail the identities.
"""
import numpy as np
import pytest

from oracle import hybrid, pe, sht, sigma, states, units

PS = units.SimUnits.from_si()


# ---------------------------------------------------------------- coordinates


def test_construction():
  with pytest.raises(ValueError, match='got 2 and 3'):
    hybrid.HybridCoordinates(np.array([1, 2]), np.array([1, 2, 3]))
  lv = hybrid.HybridCoordinates.from_coefficients([0.0, 0.1, 0.2], [0.5, 0.6, 0.7], 1000.0)
  np.testing.assert_allclose(lv.a_boundaries, [0.0, 100.0, 200.0])
  np.testing.assert_allclose(lv.b_boundaries, [0.5, 0.6, 0.7])
  lv = hybrid.HybridCoordinates.analytic_levels(10, p_top=10.0, p0=1000.0)
  assert len(lv.a_boundaries) == 11 and lv.layers == 10
  assert abs(lv.a_boundaries[0] - 10.0) < 1e-12 and lv.b_boundaries[0] == 0.0
  assert lv.a_boundaries[-1] == 0.0 and lv.b_boundaries[-1] == 1.0
  s = sigma.SigmaCoordinates.equidistant(10)
  lv = hybrid.HybridCoordinates.from_sigma_levels(s)
  np.testing.assert_allclose(lv.a_boundaries, 0.0)
  np.testing.assert_allclose(lv.b_boundaries, s.boundaries)
  # pressure at the boundaries of a standard column
  np.testing.assert_allclose(lv.pressure_boundaries(1000.0), 1000.0 * s.boundaries, rtol=1e-7)


def _scaling_ok(layers, errors, order=2):
  ratios = np.diff(np.log(errors))
  expected = np.diff(np.log(1.0 / np.asarray(layers, float) ** order))
  return np.allclose(ratios, expected, atol=0.1)


def _quadratic(p, lon, lat):
  return p ** 2 * (1 + np.cos(lon) * np.cos(lat))


def _quadratic_int(p, lon, lat):
  return p ** 3 / 3 * (1 + np.cos(lon) * np.cos(lat))


def _exponential(p, lon, lat):
  return np.exp(p / 1000) * np.cos(lon) * np.sin(lat)


def _exponential_int(p, lon, lat):
  return 1000 * np.exp(p / 1000) * np.cos(lon) * np.sin(lat)


@pytest.mark.parametrize('fn,integral,layers,res', [
    (_quadratic, _quadratic_int, [10, 20, 40, 80, 160, 320], 8),
    (_exponential, _exponential_int, [10, 20, 40, 80, 160], 16),
])
def test_pressure_integrals_converge(fn, integral, layers, res):
  """hybrid_coordinates_test.py:190-282: downward / upward cumulative and
  definite integrals over pressure on analytic levels, surface pressure
  1000 + 100 cos(lon) cos(lat) hPa."""
  grid = sht.Grid.with_wavenumbers(res)
  lon, lat = np.meshgrid(grid.nodal_axes[0], grid.nodal_axes[1])
  ps = 1000.0 + 100 * np.cos(lon) * np.cos(lat)
  errs = {'down': [], 'up': [], 'total': []}
  for n in layers:
    lv = hybrid.HybridCoordinates.analytic_levels(n)
    x = fn(lv.pressure_centers(ps), lon, lat)
    ind = integral(lv.pressure_boundaries(ps), lon, lat)
    down = hybrid.cumulative_integral_over_pressure(x, ps, lv)
    up = hybrid.cumulative_integral_over_pressure(x, ps, lv, downward=False)
    total = hybrid.integral_over_pressure(x, ps, lv, keepdims=False)
    errs['down'].append(np.abs((ind[-1] - ind[0]) - down[-1]).max())
    errs['up'].append(np.abs(-(ind[0] - ind[-1]) - up[0]).max())
    errs['total'].append(np.abs((ind[-1] - ind[0]) - total).max())
  for k, e in errs.items():
    assert _scaling_ok(layers, e), (k, e)
  np.testing.assert_allclose(total, ind[-1] - ind[0], rtol=1e-2)


def _sigma_quadratic(s, lon, lat):
  s, lon, lat = s[:, None, None], lon[None, :, None], lat[None, None, :]
  return s ** 2 * (1 + np.cos(lon) * np.cos(lat))


def _sigma_quadratic_int(s, lon, lat):
  s, lon, lat = s[:, None, None], lon[None, :, None], lat[None, None, :]
  return s ** 3 / 3 * (1 + np.cos(lon) * np.cos(lat))


def _sigma_quadratic_d(s, lon, lat):
  s, lon, lat = s[:, None, None], lon[None, :, None], lat[None, None, :]
  return 2 * s * (1 + np.cos(lon) * np.cos(lat))


def test_sigma_integrals_and_difference_converge():
  """hybrid_coordinates_test.py:283-365 on linear levels (b = eta)."""
  grid = sht.Grid.with_wavenumbers(8)
  lon, lat = grid.nodal_axes
  layers = [10, 20, 40, 80, 160, 320]
  e_cum, e_tot = [], []
  for n in layers:
    lv = hybrid.HybridCoordinates.analytic_levels(n, stretch_exponent=1.0, sigma_exponent=1.0)
    x = _sigma_quadratic(lv.b_centers, lon, lat)
    ind = _sigma_quadratic_int(lv.b_boundaries, lon, lat)
    e_cum.append(np.abs((ind[-1] - ind[0]) - hybrid.cumulative_integral_over_sigma(x, lv)[-1]).max())
    e_tot.append(np.abs((ind[-1] - ind[0]) - hybrid.integral_over_sigma(x, lv, keepdims=False)).max())
  assert _scaling_ok(layers, e_cum) and _scaling_ok(layers, e_tot)
  lv = hybrid.HybridCoordinates.analytic_levels(320, sigma_exponent=1.0, stretch_exponent=1.0)
  d = hybrid.centered_difference(_sigma_quadratic(lv.b_centers, lon, lat), lv)
  np.testing.assert_allclose(d, _sigma_quadratic_d(lv.b_boundaries[1:-1], lon, lat), atol=1e-3)


# ---------------------------------------------------------------- Simmons-Burridge balance


def _column_balance_residual(levels, t0=250.0, r_gas=287.0, ps=95000.0, eps=1e-4):
  """max_k |d Phi_k / d ln ps + C_k - R T0| / (R T0) for an isothermal column
  (a in Pa), by central differences in ln ps."""
  lv = hybrid.HybridCoordinates(levels.a_boundaries * 100.0, levels.b_boundaries)
  temp = np.full((lv.layers, 1, 1), t0)

  def phi(p):
    return hybrid.get_geopotential_diff_hybrid(temp, lv, r_gas, np.full((1, 2, 2), p))[:, 0, 0]
  dphi = (phi(ps * np.exp(eps)) - phi(ps * np.exp(-eps))) / (2 * eps)
  coef = hybrid.pgf_lps_coefficient(np.full(lv.layers, t0), ps, lv, r_gas)
  return np.abs(dphi + coef - r_gas * t0).max() / (r_gas * t0)


@pytest.mark.parametrize('levels', [
    hybrid.HybridCoordinates.analytic_levels(20, p_top=10.0),
    hybrid.HybridCoordinates.analytic_levels(37),
    hybrid.HybridCoordinates.from_sigma_levels(sigma.SigmaCoordinates.equidistant(20)),
])
def test_isothermal_column_is_in_balance(levels):
  assert _column_balance_residual(levels) < 1e-8


def test_column_balance_fails_for_planted_mistakes(monkeypatch):
  levels = hybrid.HybridCoordinates.analytic_levels(20, p_top=10.0)
  orig = hybrid.vertical_coefficients

  def half_alpha(coords, ps):
    alpha, log_ratio = orig(coords, ps)
    return 0.5 * alpha, log_ratio
  monkeypatch.setattr(hybrid, 'vertical_coefficients', half_alpha)
  assert _column_balance_residual(levels) > 1e-3
  monkeypatch.undo()

  def upper_b(temperature, ps, coords, r_gas):
    # B of the lower boundary where the upper one belongs
    alpha, log_ratio = hybrid.vertical_coefficients(coords, ps)
    dp = coords.layer_thickness(ps)
    w = log_ratio * coords.b_boundaries[1:] + alpha * np.diff(coords.b_boundaries)
    return r_gas * temperature / dp * w * ps
  monkeypatch.setattr(hybrid, 'pgf_lps_coefficient', upper_b)
  assert _column_balance_residual(levels) > 1e-3


# ---------------------------------------------------------------- equations


def _jw_pair(k=12, res=21):
  grid = sht.Grid.with_wavenumbers(res)
  sl = sigma.SigmaCoordinates.equidistant(k, dtype=np.float64)
  cs = pe.Coords(grid, sl)
  ch = pe.Coords(grid, hybrid.HybridCoordinates.from_sigma_levels(sl))
  st, oro, reft, _ = states.steady_state_jw(cs, PS)
  # vertically varying divergence and a non-uniform ln ps, so the mass flux,
  # its vertical advection and the pressure-gradient terms are all exercised
  rs = np.random.RandomState(3)
  _, l = grid.modal_mesh

  def pert(shape, amp):
    return amp * rs.normal(size=shape) * grid.mask * (l <= 10) / (l + 1) ** 2
  st = st + pe.State(pert(cs.modal_shape, 1e-2), pert(cs.modal_shape, 1e-2),
                     pert(cs.modal_shape, 1.0), pert(cs.surface_modal_shape, 1e-2))
  morog = pe.truncated_modal_orography(oro, cs)
  return (grid, sl, st, pe.PrimitiveEquationsSigma(reft, morog, cs, PS),
          hybrid.PrimitiveEquationsHybrid(reft, morog, ch, PS))


def _sigma_like_mismatch():
  grid, sl, st, es, eh = _jw_pair()
  a = es.explicit_terms(st) + es.implicit_terms(st)
  b = eh.explicit_terms(st) + eh.implicit_terms(st)
  g_s = pe.get_geopotential_weights_sigma(sl, PS.R)
  g_h = hybrid.get_geopotential_weights_hybrid(eh.levels, PS.R, eh.p_s_ref)
  pred = -grid.laplacian(pe.vertical_matvec(g_h - g_s, st.temperature_variation))
  top = (Ellipsis, slice(0, -1))
  div = np.abs((b.divergence - a.divergence)[top] - pred[top]).max() / np.abs(pred[top]).max()
  vort = np.abs(b.vorticity - a.vorticity).max() / np.abs(a.vorticity).max()
  lnps = np.abs(b.log_surface_pressure - a.log_surface_pressure).max() / np.abs(
      a.log_surface_pressure).max()
  return div, vort, lnps, np.abs(pred).max() / np.abs(a.divergence).max()


def test_sigma_like_hybrid_reproduces_sigma_equations():
  div, vort, lnps, pred_size = _sigma_like_mismatch()
  assert pred_size > 0.1  # the two geopotential discretisations differ visibly
  assert div < 1e-7 and vort < 1e-12 and lnps < 1e-12, (div, vort, lnps)


def test_sigma_like_comparison_fails_for_planted_mistakes(monkeypatch):
  orig = hybrid.hybrid_vertical_advection
  monkeypatch.setattr(hybrid, 'hybrid_vertical_advection', lambda m, x, dp: 2 * orig(m, x, dp))
  _, vort, _, _ = _sigma_like_mismatch()
  assert vort > 1e-3
  monkeypatch.undo()
  orig_lsp = hybrid.PrimitiveEquationsHybrid.nodal_log_pressure_tendency
  monkeypatch.setattr(hybrid.PrimitiveEquationsHybrid, 'nodal_log_pressure_tendency',
                      lambda self, aux, ps: 0.5 * orig_lsp(self, aux, ps))
  _, _, lnps, _ = _sigma_like_mismatch()
  assert lnps > 1e-3


def _temperature_profile_errors(layers=(8, 16, 32, 64, 128)):
  """As the sigma test: T'(sigma) = sigma^2, divergence (1 - 2 sigma) Y on
  sigma-like hybrid levels; continuous tendency (2 - kappa) sigma^2 (1 - sigma) Y."""
  grid = sht.Grid.T21()
  n_l = grid.modal_shape[1]
  row, deg = 1, 3
  errs = []
  for k in layers:
    sl = sigma.SigmaCoordinates.equidistant(k, dtype=np.float64)
    coords = pe.Coords(grid, hybrid.HybridCoordinates.from_sigma_levels(sl))
    s = sl.centers
    div = np.zeros(coords.modal_shape)
    div[:, row, deg] = 1 - 2 * s
    temp = np.zeros(coords.modal_shape)
    temp[:, 0, 0] = s ** 2 * np.sqrt(4 * np.pi)
    lnps = np.zeros(coords.surface_modal_shape)
    lnps[0, 0, 0] = np.log(units.pascal(1e5)) * np.sqrt(4 * np.pi)
    st = pe.State(np.zeros(coords.modal_shape), div, temp, lnps)
    eq = hybrid.PrimitiveEquationsHybrid(np.full(k, 250.0), np.zeros(grid.modal_shape), coords, PS)
    tend = eq.explicit_terms(st).temperature_variation
    others = np.delete(tend.reshape(k, -1), row * n_l + deg, axis=1)
    assert np.abs(others).max() < 1e-11
    errs.append(np.abs(tend[:, row, deg] - (2 - PS.kappa) * s ** 2 * (1 - s)).max())
  return np.array(errs)


def test_adiabatic_and_vertical_advection_converge_to_closed_form():
  """The reference's hybrid T omega/p sums the alpha-weighted running
  integrals of the level above and below (primitive_equations.py:2032-2056),
  which converges at first order in the interior: error ratios > 2 per
  halving, 2e-4 at 128 layers."""
  errs = _temperature_profile_errors()
  assert errs[-1] < 2.5e-4 and np.all(errs[:-1] / errs[1:] > 2.0), errs


def test_closed_form_fails_for_planted_mistake(monkeypatch):
  orig = hybrid.PrimitiveEquationsHybrid._t_omega_over_p
  monkeypatch.setattr(hybrid.PrimitiveEquationsHybrid, '_t_omega_over_p',
                      lambda self, t, g, v, ps: -orig(self, t, g, v, ps))
  errs = _temperature_profile_errors((32, 64, 128))
  assert errs[-1] > 1e-2


def test_log_pressure_tendency_is_column_mass_continuity():
  """d ln ps / dt = -sum_k (dp_k / ps div_k + dB_k v . grad ln ps) on
  analytic levels with a pressure part (explicit + implicit)."""
  grid = sht.Grid.with_wavenumbers(21)
  coords = pe.Coords(grid, hybrid.HybridCoordinates.analytic_levels(12, p_top=5.0))
  rs = np.random.RandomState(2)
  _, l = grid.modal_mesh

  def pert(shape, amp):
    return amp * rs.normal(size=shape) * grid.mask * (l <= 8) / (l + 1)
  lnps = pert(coords.surface_modal_shape, 1e-2)
  lnps[..., 0, 0] += np.log(units.pascal(1e5)) * np.sqrt(4 * np.pi)
  st = pe.State(pert(coords.modal_shape, 1e-2), pert(coords.modal_shape, 1e-2),
                pert(coords.modal_shape, 1.0), lnps)
  eq = hybrid.PrimitiveEquationsHybrid(np.full(12, 260.0), np.zeros(grid.modal_shape), coords, PS)
  tot = eq.explicit_terms(st) + eq.implicit_terms(st)
  aux = hybrid.compute_diagnostic_state_hybrid(st, eq.nondim_coords)
  ps = np.exp(grid.to_nodal(st.log_surface_pressure))
  cont = -np.sum(eq.levels.layer_thickness(ps) / ps * aux.divergence
                 + eq.levels.sigma_thickness[:, None, None] * aux.u_dot_grad_log_sp, axis=0)
  want = grid.to_modal(cont)
  top = (Ellipsis, slice(0, -1))
  err = np.abs(tot.log_surface_pressure[0][top] - want[top]).max() / np.abs(want).max()
  assert err < 1e-12, err
  # the mass flux vanishes at the model top and the surface
  assert np.abs(aux.mass_flux_full[0]).max() == 0.0
  assert np.abs(aux.mass_flux_full[-1]).max() < 1e-12 * np.abs(aux.mass_flux_full).max()


@pytest.mark.parametrize('levels', [
    hybrid.HybridCoordinates.from_sigma_levels(sigma.SigmaCoordinates.equidistant(5)),
    hybrid.HybridCoordinates.analytic_levels(8, p_top=5.0),
])
def test_implicit_inverse_identity(levels):
  """primitive_equations_test.py:856-895: (1 - eta G)^-1 (x - eta G x) = x."""
  grid = sht.Grid.with_wavenumbers(16)
  coords = pe.Coords(grid, levels)
  k = levels.layers
  eq = hybrid.PrimitiveEquationsHybrid(np.full(k, 280.0), np.zeros(grid.modal_shape), coords, PS)
  rs = np.random.RandomState(0)
  sc = (grid.total_wavenumbers + 1) ** -2
  st = pe.State(*(sc * rs.normal(size=s) for s in [coords.modal_shape] * 3
                  + [coords.surface_modal_shape]))
  inv = eq.implicit_inverse(st - 0.1 * eq.implicit_terms(st), 0.1)
  for f in pe.State.FIELDS:
    np.testing.assert_allclose(getattr(inv, f), getattr(st, f), atol=1e-12)
  with pytest.raises(TypeError):
    eq.implicit_inverse(st, np.array(0.1))


@pytest.mark.parametrize('layers', [4, 10])
def test_explicit_terms_shapes(layers):
  """primitive_equations_test.py:741-785."""
  grid = sht.Grid.with_wavenumbers(32)
  coords = pe.Coords(grid, hybrid.HybridCoordinates.analytic_levels(layers))
  ms, ss = coords.modal_shape, coords.surface_modal_shape
  st = pe.State(np.ones(ms), np.ones(ms), np.ones(ms), np.ones(ss))
  eq = hybrid.PrimitiveEquationsHybrid(300 * np.ones(layers), np.zeros(grid.modal_shape),
                                       coords, PS)
  for out in (eq.explicit_terms(st), eq.implicit_terms(st)):
    for f in pe.State.FIELDS:
      assert getattr(out, f).shape == getattr(st, f).shape, f


def _hybrid_noise_case():
  """The GPU parity case of the hybrid explicit terms: T21, 12 analytic levels
  (p_top = 5 hPa), seeded fields, ln ps around ln(1e5 Pa) nondimensional."""
  og = sht.Grid.T21()
  lev = hybrid.HybridCoordinates.analytic_levels(12, p_top=5.0)
  ps_ = units.SimUnits.from_si()
  rs = np.random.RandomState(3)
  _, l = og.modal_mesh
  oro = rs.normal(size=og.modal_shape) * og.mask * 1e-4 / (l + 1) ** 2
  eq = hybrid.PrimitiveEquationsHybrid(np.full(12, 260.0), oro, pe.Coords(og, lev), ps_)

  def field(n, amp):
    return rs.normal(size=(n,) + og.modal_shape) * og.mask * amp / (l + 1) ** 2

  lnps = field(1, 0.05)
  lnps[..., 0, 0, 0] += np.sqrt(4 * np.pi) * np.log(ps_.nondim(1e5, length=-1, time=-2, mass=1))
  return og, eq, [field(12, 1e-2), field(12, 1e-2), field(12, 1.0), lnps]


def test_hybrid_divergence_rounding_floor():
  """The hybrid explicit terms are ill-conditioned in ln ps: the geopotential
  and pressure-gradient terms use ps = exp(ln ps) with nondimensional
  ln ps ~ 45, and the Laplacian amplifies the round-off a transform leaves
  at high degrees.  Sending ln ps once through an FP64 nodal round trip
  (to_nodal, to_modal) moves the divergence tendency by ~4e-9 of its
  maximum (temperature ~1e-9, ln ps ~1e-10, vorticity ~2e-12).  This floor
  is why the GPU parity test compares the FP64 divergence at 2e-9 (two
  independent FP64 pipelines differ there by 4.4e-10) and the other fields
  at 1e-10 (they differ by < 4e-13)."""
  og, eq, arrs = _hybrid_noise_case()
  ref = eq.explicit_terms(pe.State(*arrs))
  pert = list(arrs)
  pert[3] = og.to_modal(og.to_nodal(arrs[3]))
  # the round trip changes ln ps only at round-off
  assert np.abs(pert[3] - arrs[3]).max() < 1e-13 * np.abs(arrs[3]).max()
  out = eq.explicit_terms(pe.State(*pert))
  rel = lambda f: (np.abs(getattr(out, f) - getattr(ref, f)).max() /
                   np.abs(getattr(ref, f)).max())
  assert 1e-9 < rel('divergence') < 2e-8
  assert 1e-10 < rel('temperature_variation') < 1e-8
  assert rel('vorticity') < 1e-11
