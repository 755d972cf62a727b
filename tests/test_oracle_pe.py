"""Pins the oracle's primitive-equation terms.

Composed explicit_terms have no golden vectors anywhere in the reference, so
they are pinned by physics invariants chosen so a dropped or mis-signed term
fails one of them: JW steady-state balance of the divergence budget, tracer
advection by the JW zonal jet against the closed-form -u . grad q, q = 0 moist
equals dry, the implicit-inverse identity, entrywise G/H/alpha values, and
Held-Suarez closed forms.  The humidity and adiabatic terms are pinned by
three identities the continuous equations fix: uniform humidity is a dry
atmosphere with virtual gas constants, the split humidity terms sum to the
unsplit virtual-temperature pressure gradient, and the discrete T omega/p and
sigma-dot advection converge at second order to their closed forms.  The
tendencies in SI units do not depend on the nondimensionalisation scale.
Mutation checks plant a mistake in the oracle and show each identity fails.
"""
import dataclasses

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


# ---------------------------------------------------------------- humidity and adiabatic terms


def _perturbed_jw(grid, k, seed=3):
  coords, st, morog, reft = _jw(grid, k)
  rs = np.random.RandomState(seed)
  _, l = grid.modal_mesh

  def pert(shape, amp):
    return amp * rs.normal(size=shape) * grid.mask * (l <= 10) / (l + 1) ** 2
  st = st + pe.State(pert(coords.modal_shape, 1e-2), pert(coords.modal_shape, 1e-2),
                     pert(coords.modal_shape, 1.0), pert(coords.surface_modal_shape, 1e-2))
  return coords, st, morog, reft


def _uniform_q_mismatch(reft_kind='jw', q0=0.01):
  """Relative mismatch between the moist total tendency (explicit +
  implicit) of a state with spatially uniform humidity q0 and no clouds and
  the dry total tendency with R -> R (1 + (R_v/R - 1) q0) and
  kappa -> kappa (1 + (R_v/R - 1) q0) / (1 + (C_pv/C_p - 1) q0): the moist
  equations with uniform q are the dry equations of a gas with those
  constants.  Compared below the top degree (the reference clips explicit
  tendencies there and not implicit ones)."""
  grid = sht.Grid.T21()
  coords, st, morog, reft = _perturbed_jw(grid, 12)
  if reft_kind == 'linear':
    reft = np.linspace(220.0, 300.0, 12)
  qm = grid.to_modal(np.full(coords.nodal_shape, q0))
  zeros = np.zeros_like(qm)
  moist_state = pe.State(st.vorticity, st.divergence, st.temperature_variation,
                         st.log_surface_pressure,
                         tracers={'q': qm, 'c1': zeros, 'c2': zeros.copy()})
  moist = pe.PrimitiveEquationsSigma(reft, morog, coords, PS, humidity_key='q',
                                     cloud_keys=('c1', 'c2'))
  tot_m = moist.explicit_terms(moist_state) + moist.implicit_terms(moist_state)
  c = PS.R_vapor / PS.R - 1
  hc = PS.Cp_vapor / PS.Cp - 1
  ps_eff = units.SimUnits.from_si(ideal_gas_constant_si=units.R_SI * (1 + c * q0),
                                  kappa_si=units.KAPPA * (1 + c * q0) / (1 + hc * q0))
  dry = pe.PrimitiveEquationsSigma(reft, morog, coords, ps_eff)
  dry_state = pe.State(st.vorticity, st.divergence, st.temperature_variation,
                       st.log_surface_pressure)
  tot_d = dry.explicit_terms(dry_state) + dry.implicit_terms(dry_state)
  worst = 0.0
  for f in pe.State.FIELDS:
    a, b = getattr(tot_m, f)[..., :-1], getattr(tot_d, f)[..., :-1]
    worst = max(worst, np.abs(a - b).max() / np.abs(b).max())
  return worst


@pytest.mark.parametrize('reft_kind', ['jw', 'linear'])
def test_uniform_humidity_is_dry_gas_with_virtual_constants(reft_kind):
  """Pins the humidity divergence term (virtual-temperature geopotential and
  the q lap(ln ps) correction), the virtual-temperature factor of the
  pressure-gradient term and the moist adiabatic factor: at q0 = 0.01 they
  shift the tendencies by ~1e-3 relative, and the identity holds to 1e-9."""
  assert _uniform_q_mismatch(reft_kind) < 1e-9


def test_uniform_humidity_identity_fails_for_planted_mistakes(monkeypatch):
  assert _uniform_q_mismatch() < 1e-9
  orig = pe.PrimitiveEquationsSigma.divergence_tendency_due_to_humidity
  monkeypatch.setattr(pe.PrimitiveEquationsSigma, 'divergence_tendency_due_to_humidity',
                      lambda self, s, a: -orig(self, s, a))
  assert _uniform_q_mismatch() > 1e-4
  monkeypatch.undo()
  orig_adj = pe.PrimitiveEquationsSigma.nodal_temperature_adiabatic_tendency
  # the moist adiabatic factor computed with the dry kappa only
  monkeypatch.setattr(pe.PrimitiveEquationsSigma, 'nodal_temperature_adiabatic_tendency',
                      lambda self, aux: orig_adj(dataclasses.replace(self, humidity_key=None,
                                                                     cloud_keys=None),
                                                 aux))
  assert _uniform_q_mismatch() > 1e-4


def _unsplit_mismatch(grid_name='T21', lmax=7, sign=1.0):
  """Relative mismatch between the total (explicit + implicit) vorticity and
  divergence tendencies of a resting moist atmosphere with non-uniform q,
  T' and ln ps (all of degree <= lmax, so their products are exact on the
  grid) and the unsplit form: zeta_t = -k . curl(R T_v grad ln ps),
  delta_t = -div(R T_v grad ln ps) - lap(Phi(T_v) + g h) with the full
  virtual temperature T_v = (T_ref + T')(1 + (R_v/R - 1) q).  The reference
  splits this into the dry terms, the humidity dot/curl terms in grad q and
  the q lap(ln ps) correction; `sign` multiplies the humidity terms (a
  planted mistake when -1)."""
  grid = getattr(sht.Grid, grid_name)()
  k = 6
  coords = _coords(grid, k)
  rs = np.random.RandomState(4)
  _, l = grid.modal_mesh

  def pert(shape, amp):
    return amp * rs.normal(size=shape) * grid.mask * (l <= lmax) / (l + 1)
  zero = np.zeros(coords.modal_shape)
  tprime = pert(coords.modal_shape, 3.0)
  lnps = pert(coords.surface_modal_shape, 1e-2)
  lnps[..., 0, 0] = np.log(units.pascal(1e5)) * np.sqrt(4 * np.pi)
  q = pert(coords.modal_shape, 3e-3)
  q[..., 0, 0] += 0.008 * np.sqrt(4 * np.pi)
  oro = pert(grid.modal_shape, 1e-4)
  tref = np.full(k, 250.0)
  st = pe.State(zero, zero.copy(), tprime, lnps, tracers={'q': q})
  eq = pe.PrimitiveEquationsSigma(tref, oro, coords, PS, humidity_key='q')
  aux = pe.compute_diagnostic_state_sigma(st, coords)
  ex = eq.explicit_terms(st)
  if sign != 1.0:
    hd = eq.divergence_tendency_due_to_humidity(st, aux)
    hv = eq.vorticity_tendency_due_to_humidity(st, aux)
    ex.divergence = ex.divergence + (sign - 1) * grid.clip_wavenumbers(hd)
    ex.vorticity = ex.vorticity + (sign - 1) * grid.clip_wavenumbers(hv)
  tot = ex + eq.implicit_terms(st)
  c = PS.R_vapor / PS.R - 1
  tv = (tref[:, None, None] + aux.temperature_variation) * (1 + c * grid.to_nodal(q))
  gx, gy = aux.cos_lat_grad_log_sp
  force = (grid.to_modal(PS.R * tv * gx * grid.sec2_lat),
           grid.to_modal(PS.R * tv * gy * grid.sec2_lat))
  g_w = pe.get_geopotential_weights_sigma(coords.vertical, PS.R)
  phi = grid.to_modal(pe.vertical_matvec(g_w, tv - tref[:, None, None]))
  exp_div = (-grid.div_cos_lat(force, clip=False) - grid.laplacian(phi)
             - PS.g * grid.laplacian(oro))
  exp_vort = -grid.curl_cos_lat(force, clip=False)
  top = (Ellipsis, slice(0, -1))
  return (np.abs(tot.divergence[top] - exp_div[top]).max() / np.abs(exp_div[top]).max(),
          np.abs(tot.vorticity[top] - exp_vort[top]).max() / np.abs(exp_vort[top]).max())


@pytest.mark.parametrize('grid_name,lmax', [('T21', 7), ('T42', 14)])
def test_humidity_split_sums_to_unsplit_pressure_gradient(grid_name, lmax):
  """Pins the grad q . grad ln ps (divergence) and grad ln ps x grad q
  (vorticity) humidity terms, the correction q lap(ln ps) and their
  R_v - R coefficients: the humidity terms are ~2% of the divergence
  tendency and ~15% of the vorticity tendency here."""
  div_err, vort_err = _unsplit_mismatch(grid_name, lmax)
  assert div_err < 1e-11 and vort_err < 1e-9, (div_err, vort_err)
  bad_div, bad_vort = _unsplit_mismatch(grid_name, lmax, sign=-1.0)
  assert bad_div > 1e-3 and bad_vort > 1e-2, (bad_div, bad_vort)


def _adiabatic_profile_errors(layers=(8, 16, 32, 64, 128)):
  """Explicit T' tendency of a resting, flat, horizontally uniform T'(sigma)
  = sigma^2 driven by a divergence (1 - 2 sigma) Y (Y a single spherical
  harmonic, so ln ps stays uniform), against the continuous value
  (2 - kappa) sigma^2 (1 - sigma) Y: vertical advection sigma_dot dT/dsigma
  with sigma_dot = -sigma (1 - sigma) Y, plus kappa T' omega/p with
  omega/p = -(1/sigma) int_0^sigma delta = -(1 - sigma) Y (the T_ref part
  is implicit).  Returns the max error per resolution."""
  grid = sht.Grid.T21()
  n_l = grid.modal_shape[1]
  row, deg = 1, 3   # m = +1 cosine row, l = 3
  errs = []
  for k in layers:
    coords = _coords(grid, k)
    s = coords.vertical.centers.astype(np.float64)
    div = np.zeros(coords.modal_shape)
    div[:, row, deg] = 1 - 2 * s
    temp = np.zeros(coords.modal_shape)
    temp[:, 0, 0] = s ** 2 * np.sqrt(4 * np.pi)
    lnps = np.zeros(coords.surface_modal_shape)
    lnps[0, 0, 0] = np.log(units.pascal(1e5)) * np.sqrt(4 * np.pi)
    st = pe.State(np.zeros(coords.modal_shape), div, temp, lnps)
    eq = pe.PrimitiveEquationsSigma(np.full(k, 250.0), np.zeros(grid.modal_shape), coords, PS)
    tend = eq.explicit_terms(st).temperature_variation
    exact = (2 - PS.kappa) * s ** 2 * (1 - s)
    others = np.delete(tend.reshape(k, -1), row * n_l + deg, axis=1)
    assert np.abs(others).max() < 1e-11
    errs.append(np.abs(tend[:, row, deg] - exact).max())
  return np.array(errs)


def test_adiabatic_and_vertical_advection_converge_to_closed_form():
  """Pins _t_omega_over_sigma_sp (Durran 8.124) and the centred sigma-dot
  advection: second-order convergence (error ratio -> 4 per halving of the
  layer thickness) to the continuous tendency."""
  errs = _adiabatic_profile_errors()
  assert _second_order_to_closed_form(errs), errs


def _second_order_to_closed_form(errs):
  rates = errs[:-1] / errs[1:]
  return bool(errs[-1] < 1e-4 and np.all(rates > 3.7))


def test_adiabatic_closed_form_fails_for_planted_mistakes(monkeypatch):
  orig = pe.PrimitiveEquationsSigma._t_omega_over_sigma_sp

  def shifted(self, tfield, g_term, v_dot_grad_log_sp):
    # the alpha-weighted sum taken over the level's own integral twice
    # (a dropped shift)
    f = sigma.cumulative_sigma_integral(g_term, self.coords.vertical)
    alpha = pe.get_sigma_ratios(self.coords.vertical)[:, None, None]
    dsig = self.coords.vertical.layer_thickness[:, None, None]
    return tfield * (v_dot_grad_log_sp - 2 * alpha * f / dsig)
  monkeypatch.setattr(pe.PrimitiveEquationsSigma, '_t_omega_over_sigma_sp', shifted)
  assert not _second_order_to_closed_form(_adiabatic_profile_errors((32, 64, 128)))
  monkeypatch.setattr(pe.PrimitiveEquationsSigma, '_t_omega_over_sigma_sp', orig)
  monkeypatch.setattr(pe.PrimitiveEquationsSigma, '_vadv',
                      lambda self, w, x: -sigma.centered_vertical_advection(
                          w, x, self.coords.vertical))
  assert not _second_order_to_closed_form(_adiabatic_profile_errors((32, 64, 128)))


# ---------------------------------------------------------------- scale invariance


def test_explicit_tendencies_in_si_units_do_not_depend_on_scale():
  """primitive_equations_test.py:466-576: JW + baroclinic perturbation at
  wavenumber 64, 10 layers, in the default nondimensionalisation and in a
  custom one (length / 100, time x 55.3, mass x 16.4, temperature x 3.15):
  the tendencies in SI units agree (atol 5e-7, surface pressure 1e-7)."""
  custom = units.Scale(units.RADIUS_M / 100, 55.3 / 2 / units.OMEGA_PER_S, 16.4, 3.15)
  out = []
  for scale in (units.DEFAULT_SCALE, custom):
    ps = units.SimUnits.from_si(scale=scale)
    grid = sht.Grid.with_wavenumbers(64, radius=ps.radius)
    coords = _coords(grid, 10)
    st, oro, reft, _ = states.steady_state_jw(coords, ps)
    st = st + states.baroclinic_perturbation_jw(coords, ps)
    eq = pe.PrimitiveEquationsSigma(reft, pe.truncated_modal_orography(oro, coords), coords, ps)
    ten = eq.explicit_terms(st)
    out.append({
        'divergence': ps.dim(ten.divergence, time=-2),
        'vorticity': ps.dim(ten.vorticity, time=-2),
        'temperature': ps.dim(ten.temperature_variation, temperature=1, time=-1),
        'pressure': ps.dim(np.exp(ten.log_surface_pressure), mass=1, length=-1, time=-3),
    })
    if scale is custom:
      assert abs(ps.radius - 100.0) < 1e-9
  for key, atol in (('divergence', 5e-7), ('vorticity', 5e-7), ('temperature', 5e-7),
                    ('pressure', 1e-7)):
    np.testing.assert_allclose(out[0][key], out[1][key], atol=atol, err_msg=key)
  # the SI tendencies are small (divergence ~1e-9 s^-2), so the reference's
  # absolute tolerance alone is loose: they also agree normwise to 1e-9
  for key in ('divergence', 'vorticity', 'temperature'):
    rel = np.abs(out[0][key] - out[1][key]).max() / np.abs(out[0][key]).max()
    assert rel < 1e-9, (key, rel)
  # a constant given the wrong dimensions breaks the invariance
  bad = dataclasses.replace(units.SimUnits.from_si(scale=custom),
                            gravity_acceleration=units.GRAVITY_SI / custom.length)
  grid = sht.Grid.with_wavenumbers(64, radius=bad.radius)
  coords = _coords(grid, 10)
  st, oro, reft, _ = states.steady_state_jw(coords, bad)
  eq = pe.PrimitiveEquationsSigma(reft, pe.truncated_modal_orography(oro, coords), coords, bad)
  wrong = bad.dim(eq.explicit_terms(st).divergence, time=-2)
  rel = np.abs(wrong - out[0]['divergence']).max() / np.abs(out[0]['divergence']).max()
  assert rel > 1e-4, rel


# ---------------------------------------------------------------- rest state


def test_rest_state_surface_pressure_barometric_closed_form():
  """primitive_equations_states.py:107-124: ln ps of the rest atmosphere is
  ln p0 at sea level and falls with altitude at the sea-level rate
  -g M / (R0 T0) (1 / the 8.433 km scale height of air at 288.16 K), to first
  order in h; its second-order behaviour is that of the polytrope
  (1 - g h / (cp T0))^(cp M / R0)."""
  h = np.array([0.0, 1.0, 10.0, 100.0, 1000.0, 3000.0])
  rel = states.relative_pressure(h)
  assert rel[0] == 1.0
  g, cp, t0, mm, r0 = 9.80665, 1004.68506, 288.16, 0.02896968, 8.314462618
  scale_height = r0 * t0 / (g * mm)
  assert abs(scale_height - 8433.4) < 0.5
  # ln(1 - x) = -x (1 + x / 2 + ...): relative deviation g h / (2 cp T0)
  slope = np.log(rel[1:4]) / h[1:4]
  for sl, hh in zip(slope, h[1:4]):
    assert abs(sl * scale_height + 1) < 2e-5 * hh, (hh, sl)
  # exact polytrope: d ln p / dh = -(g M / R0) / (T0 - g h / cp)
  eps = 0.5
  dlog = (np.log(states.relative_pressure(h[1:] + eps)) -
          np.log(states.relative_pressure(h[1:] - eps))) / (2 * eps)
  np.testing.assert_allclose(dlog, -(g * mm / r0) / (t0 - g * h[1:] / cp), rtol=1e-6)
  # the oracle's rest state at those altitudes, nondimensional orography
  grid = sht.Grid.T21()
  coords = _coords(grid, 4)
  oro_m = 1500.0 * (1 + np.sin(3 * np.arcsin(grid.nodal_mesh[1])))
  st, _, _ = states.isothermal_rest_atmosphere(coords, PS,
                                               surface_height_nondim=oro_m / units.RADIUS_M)
  lnps = grid.to_nodal(st.log_surface_pressure)[0]
  want = np.log(units.pascal(1e5)) + np.log(states.relative_pressure(oro_m))
  # the modal state is the truncation of the nodal closed form
  assert np.abs(lnps - want).max() < 1e-6
