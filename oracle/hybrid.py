"""Hybrid sigma-pressure coordinates and the primitive equations on them
(oracle restatement; test infrastructure only).

Restates dinosaur/hybrid_coordinates.py:
  cumulative_integral_over_pressure / integral_over_pressure   :54-108
  cumulative_integral_over_sigma / integral_over_sigma         :111-156
  centered_difference / centered_vertical_advection            :159-238
  HybridCoordinates (from_coefficients, analytic_levels,
    from_sigma_levels, thicknesses, centres, sigma boundaries)  :241-596
and dinosaur/primitive_equations.py:1461-2518 (PrimitiveEquationsHybrid):
  _get_vertical_discretization_coeffs(_numpy)   :1483-1533  (Simmons & Burridge 1981)
  compute_diagnostic_state_hybrid               :1536-1602  (mass flux M_{k+1/2})
  get_geopotential_weights_hybrid / _diff       :1605-1674
  get_temperature_implicit_weights_hybrid       :1736-1783
  _get_pgf_lps_coefficient(_numpy)              :1826-1888
  _get_implicit_term_matrix_hybrid              :1890-1938
  hybrid_vertical_advection                     :1941-1957
  PrimitiveEquationsHybrid                      :1960-2518

Pinned in tests/test_oracle_hybrid.py by the reference's known answers for
the coordinate integrals (hybrid_coordinates_test.py), hybrid levels built
from sigma levels reproducing the sigma equations to the difference of the
two vertical discretisations (primitive_equations_test.py:785-854), the
implicit-inverse identity, closed forms of the Simmons-Burridge geopotential
for an isothermal column, pressure-coordinate consistency of the mass flux
and log-pressure tendency, and stationarity of the sigma-like
Jablonowski-Williamson state.
"""
from __future__ import annotations

import dataclasses

import numpy as np

from . import pe as pe_lib
from . import sht as sht_lib
from . import sigma as sigma_lib

State = pe_lib.State


# ---------------------------------------------------------------- coordinates


class HybridCoordinates:
  """Layer boundaries p = a + b ps, top to surface; a in hPa by convention."""

  def __init__(self, a_boundaries, b_boundaries):
    self.a_boundaries = np.asarray(a_boundaries, np.float64)
    self.b_boundaries = np.asarray(b_boundaries, np.float64)
    if len(self.a_boundaries) != len(self.b_boundaries):
      raise ValueError('Expected `a_boundaries` and `b_boundaries` to have the same length, '
                       f'got {len(self.a_boundaries)} and {len(self.b_boundaries)}.')

  @classmethod
  def from_coefficients(cls, a_coeffs, b_coeffs, p0=1000.0):
    return cls(np.asarray(a_coeffs, np.float64) * p0, np.asarray(b_coeffs, np.float64))

  @classmethod
  def analytic_levels(cls, n_levels, p_top=0.0, p0=1000.0, sigma_exponent=3.0,
                      stretch_exponent=2.0):
    eta = np.linspace(0, 1, n_levels + 1) ** stretch_exponent
    total = p_top + eta * (p0 - p_top)
    b = eta ** sigma_exponent
    a = total / p0 - b
    a[0], b[0], a[-1], b[-1] = p_top / p0, 0.0, 0.0, 1.0
    return cls.from_coefficients(a, b, p0)

  @classmethod
  def from_sigma_levels(cls, sigma_levels):
    b = np.asarray(sigma_levels.boundaries, np.float64)
    return cls(np.zeros_like(b), b)

  @property
  def layers(self):
    return len(self.a_boundaries) - 1

  @property
  def pressure_thickness(self):
    return np.diff(self.a_boundaries)

  @property
  def sigma_thickness(self):
    return np.diff(self.b_boundaries)

  @property
  def a_centers(self):
    return (self.a_boundaries[1:] + self.a_boundaries[:-1]) / 2

  @property
  def b_centers(self):
    return (self.b_boundaries[1:] + self.b_boundaries[:-1]) / 2

  @property
  def center_to_center(self):
    return np.diff(self.b_centers)

  @staticmethod
  def _column(v, ps):
    """v along a new leading axis with one trailing axis per non-singleton
    axis of ps (hybrid_coordinates.py:452-470)."""
    n = len([s for s in np.shape(ps) if s > 1])
    return np.asarray(v).reshape((-1,) + (1,) * n)

  def layer_thickness(self, ps):
    """Delta p of each layer for surface pressure ps."""
    return (self._column(self.pressure_thickness, ps)
            + self._column(self.sigma_thickness, ps) * ps)

  def pressure_boundaries(self, ps):
    return self._column(self.a_boundaries, ps) + self._column(self.b_boundaries, ps) * ps

  def pressure_centers(self, ps):
    bnd = self.pressure_boundaries(ps)
    return (bnd[1:] + bnd[:-1]) / 2

  def get_sigma_boundaries(self, ps):
    return self.a_boundaries / ps + self.b_boundaries

  def get_sigma_centers(self, ps):
    bnd = self.get_sigma_boundaries(ps)
    return (bnd[1:] + bnd[:-1]) / 2


def _check_layers(x, coords, axis):
  if coords.layers != x.shape[axis]:
    raise ValueError('`x.shape[axis]` must be equal to `coordinates.layers`; '
                     f'got {x.shape[axis]} and {coords.layers}.')


def _along(v, ndim, axis):
  shape = [1] * ndim
  shape[axis] = -1
  return np.asarray(v).reshape(shape)


def cumulative_integral_over_pressure(x, surface_pressure, coords, axis=-3, downward=True):
  """Running sum of x dp, dp = da + db ps (hybrid_coordinates.py:54-80)."""
  _check_layers(x, coords, axis)
  dp = (_along(coords.pressure_thickness, x.ndim, axis)
        + _along(coords.sigma_thickness, x.ndim, axis) * surface_pressure)
  xdp = x * dp
  if downward:
    return np.cumsum(xdp, axis=axis)
  return np.flip(np.cumsum(np.flip(xdp, axis), axis=axis), axis)


def integral_over_pressure(x, surface_pressure, coords, axis=-3, keepdims=True):
  """sum x da + ps sum x db (:83-108)."""
  _check_layers(x, coords, axis)
  ia = (x * _along(coords.pressure_thickness, x.ndim, axis)).sum(axis=axis, keepdims=keepdims)
  ib = (x * _along(coords.sigma_thickness, x.ndim, axis)).sum(axis=axis, keepdims=keepdims)
  return ia + surface_pressure * ib


def cumulative_integral_over_sigma(x, coords, axis=-3, downward=True):
  """Running sum of x db (the sigma part only, :111-136)."""
  _check_layers(x, coords, axis)
  xdb = x * _along(coords.sigma_thickness, x.ndim, axis)
  if downward:
    return np.cumsum(xdb, axis=axis)
  return np.flip(np.cumsum(np.flip(xdb, axis), axis=axis), axis)


def integral_over_sigma(x, coords, axis=-3, keepdims=True):
  _check_layers(x, coords, axis)
  return (x * _along(coords.sigma_thickness, x.ndim, axis)).sum(axis=axis, keepdims=keepdims)


def centered_difference(x, coords, axis=-3):
  """(x[n+1] - x[n]) / (b-centre spacing) at internal boundaries (:159-198)."""
  _check_layers(x, coords, axis)
  return np.diff(x, axis=axis) / _along(coords.center_to_center, x.ndim, axis)


def centered_vertical_advection(w, x, coords, axis=-3):
  """-(w dx/deta) averaged from the two faces of each layer, zero boundary
  values (:201-238)."""
  zw = list(w.shape)
  zw[axis] = 1
  wpad = np.concatenate([np.zeros(zw), w, np.zeros(zw)], axis=axis)
  xd = centered_difference(x, coords, axis)
  zx = list(xd.shape)
  zx[axis] = 1
  xdpad = np.concatenate([np.zeros(zx), xd, np.zeros(zx)], axis=axis)
  wx = wpad * xdpad
  n = wx.shape[axis]
  return -0.5 * (np.take(wx, np.arange(1, n), axis=axis) + np.take(wx, np.arange(n - 1), axis=axis))


# ---------------------------------------------------------------- Simmons-Burridge coefficients


def vertical_coefficients(coords: HybridCoordinates, ps):
  """(alpha_k, ln(p_{k+1/2} / p_{k-1/2})) at surface pressure ps, a scalar or
  a (1, ...) nodal field (primitive_equations.py:1483-1533): alpha_k =
  1 - p_{k-1/2} / dp_k ln(p_{k+1/2} / p_{k-1/2}), pressures floored at
  1e-6 ps (the model top has p = 0)."""
  a, b = coords.a_boundaries, coords.b_boundaries
  if np.ndim(ps) > 0:
    a, b = a[:, None, None], b[:, None, None]
  p_half = a + b * ps
  dp = np.diff(p_half, axis=0)
  if np.ndim(ps) == 0 and np.any(dp <= 0):
    raise ValueError('Pressure must decrease monotonically with height.')
  upper, lower = p_half[:-1], p_half[1:]
  floor = 1e-6 * ps
  log_ratio = np.log(np.maximum(lower, floor) / np.maximum(upper, floor))
  alpha = 1 - (upper / dp) * log_ratio
  return alpha, log_ratio


def get_geopotential_weights_hybrid(coords, ideal_gas_constant, p_s_ref):
  """R [alpha_k on the diagonal, ln(p ratio)_j for j > k] (:1605-1626)."""
  alpha, log_ratio = vertical_coefficients(coords, p_s_ref)
  k = coords.layers
  w = np.triu(np.ones((k, k)), 1) * log_ratio[None, :] + np.diag(alpha)
  return ideal_gas_constant * w


def get_geopotential_diff_hybrid(temperature, coords, ideal_gas_constant, ps):
  """R [sum_{j>k} T_j ln(p ratio)_j + alpha_k T_k] with the coefficients of
  ps (scalar or nodal) (:1629-1674, both methods)."""
  alpha, log_ratio = vertical_coefficients(coords, ps)
  if np.ndim(log_ratio) == 1:
    alpha, log_ratio = alpha[:, None, None], log_ratio[:, None, None]
  weighted = temperature * log_ratio
  below = np.flip(np.cumsum(np.flip(weighted, 0), axis=0), 0) - weighted
  return ideal_gas_constant * (below + alpha * temperature)


def get_temperature_implicit_weights_hybrid(coords, reference_temperature, kappa, p_s_ref):
  """H (:1736-1783): kappa T_ref (alpha-weighted lower-triangular sums) / dp
  minus the reference-lapse terms (with the sigma thicknesses), times dp."""
  tref = np.asarray(reference_temperature, np.float64)
  if tref.ndim != 1 or tref.shape[-1] != coords.layers:
    raise ValueError('`reference_temp` must be a vector of length `coordinates.layers`; '
                     f'got shape {tref.shape} and {coords.layers} layers.')
  k = coords.layers
  lower = np.tril(np.ones((k, k)))
  alpha, _ = vertical_coefficients(coords, p_s_ref)
  pa = lower * alpha[:, None]
  pa_prev = np.vstack([np.zeros((1, k)), pa[:-1]])
  dp = coords.pressure_thickness + coords.sigma_thickness * p_s_ref
  h0 = kappa * tref[:, None] * (pa + pa_prev) / dp[:, None]
  k0 = np.append(np.diff(tref) / (dp[:-1] + dp[1:]), 0.0)[:, None]
  kk = k0 * (lower - np.cumsum(coords.sigma_thickness)[:, None])
  kk_prev = np.vstack([np.zeros((1, k)), kk[:-1]])
  return (h0 - kk - kk_prev) * dp


def pgf_lps_coefficient(temperature, ps, coords, ideal_gas_constant):
  """C_k = R T / dp_k [ln(p ratio)_k B_{k-1/2} + alpha_k dB_k] ps, the
  coefficient of grad ln ps in the pressure-gradient force (:1826-1888);
  ps scalar (reference) or nodal."""
  alpha, log_ratio = vertical_coefficients(coords, ps)
  b_upper = coords.b_boundaries[:-1]
  db = np.diff(coords.b_boundaries)
  if np.ndim(ps) > 0:
    b_upper, db = b_upper[:, None, None], db[:, None, None]
  dp = np.maximum(coords.layer_thickness(ps), 1e-6 * ps)
  weight = log_ratio * b_upper + alpha * db
  return (ideal_gas_constant * temperature / dp) * weight * ps


def implicit_term_matrix_hybrid(eta, coords, reference_temperature, kappa, ideal_gas_constant,
                                p_s_ref):
  """(L, 2K+1, 2K+1) block matrix of (I - eta G) on hybrid levels (:1890-1938)."""
  v = coords.vertical
  k = v.layers
  lam = coords.horizontal.laplacian_eigenvalues
  g = get_geopotential_weights_hybrid(v, ideal_gas_constant, p_s_ref)
  h = get_temperature_implicit_weights_hybrid(v, reference_temperature, kappa, p_s_ref)
  c = pgf_lps_coefficient(np.asarray(reference_temperature, np.float64), p_s_ref, v,
                          ideal_gas_constant)
  eff = v.pressure_thickness / p_s_ref + v.sigma_thickness
  n_l = coords.horizontal.modal_shape[1]
  eye = np.broadcast_to(np.eye(k)[None], (n_l, k, k))
  top = np.concatenate([eye, eta * lam[:, None, None] * g[None],
                        eta * (lam[:, None] * c[None, :])[..., None]], axis=2)
  mid = np.concatenate([eta * np.broadcast_to(h[None], (n_l, k, k)), eye,
                        np.zeros((n_l, k, 1))], axis=2)
  bot = np.concatenate([np.broadcast_to(eta * eff[None, None, :], (n_l, 1, k)),
                        np.zeros((n_l, 1, k)), np.ones((n_l, 1, 1))], axis=2)
  return np.concatenate([top, mid, bot], axis=1)


def hybrid_vertical_advection(mass_flux, x, dp):
  """-(M_{k+1/2} (x_{k+1} - x_k) + M_{k-1/2} (x_k - x_{k-1})) / (2 dp_k)
  with M at all K+1 boundaries (zero at top and bottom) (:1941-1957)."""
  flux = mass_flux[1:-1] * np.diff(x, axis=0)
  z = np.zeros_like(flux[:1])
  padded = np.concatenate([z, flux, z], axis=0)
  return -0.5 * (padded[1:] + padded[:-1]) / dp


# ---------------------------------------------------------------- diagnostic state


@dataclasses.dataclass
class DiagnosticHybrid:
  vorticity: np.ndarray
  divergence: np.ndarray
  temperature_variation: np.ndarray
  cos_lat_u: tuple
  mass_flux_explicit: np.ndarray
  mass_flux_full: np.ndarray
  cos_lat_grad_log_sp: tuple
  u_dot_grad_log_sp: np.ndarray
  tracers: dict
  layer_pressure_thickness: np.ndarray


def compute_diagnostic_state_hybrid(state: State, coords) -> DiagnosticHybrid:
  """Nodal fields and the mass flux M_{k+1/2} = -sum_{r<=k} D_r +
  B_{k+1/2} sum_r D_r, D = dp div v + dB ps v.grad ln ps (:1536-1602)."""
  grid = coords.horizontal
  levels = coords.vertical
  nv = grid.to_nodal(state.vorticity)
  nd = grid.to_nodal(state.divergence)
  nt = grid.to_nodal(state.temperature_variation)
  tracers = {k: grid.to_nodal(v) for k, v in state.tracers.items()}
  cos_lat_u = tuple(grid.to_nodal(c) for c in sht_lib.get_cos_lat_vector(
      state.vorticity, state.divergence, grid, clip=False))
  nglsp = tuple(grid.to_nodal(c) for c in grid.cos_lat_grad(state.log_surface_pressure,
                                                            clip=False))
  udg = (cos_lat_u[0] * nglsp[0] + cos_lat_u[1] * nglsp[1]) * grid.sec2_lat
  ps = np.exp(grid.to_nodal(state.log_surface_pressure))
  dp = levels.layer_thickness(ps)
  adv = levels.sigma_thickness[:, None, None] * ps * udg
  full = dp * nd + adv
  b = levels.b_boundaries[:, None, None]

  def mass_flux(d):
    running = np.concatenate([np.zeros_like(d[:1]), np.cumsum(d, axis=0)], axis=0)
    return -running + b * d.sum(axis=0, keepdims=True)

  return DiagnosticHybrid(nv, nd, nt, cos_lat_u, mass_flux(adv), mass_flux(full), nglsp, udg,
                          tracers, dp)


# ---------------------------------------------------------------- equations


HPA_IN_PA = 100.0


@dataclasses.dataclass
class PrimitiveEquationsHybrid(pe_lib.PrimitiveEquationsSigma):
  """:1960-2518.  coords.vertical is a HybridCoordinates with a in hPa;
  reference_surface_pressure in Pa."""
  reference_surface_pressure: float = 101325.0

  def __post_init__(self):
    super().__post_init__()
    ps = self.physics_specs
    self.p_s_ref = ps.nondim(self.reference_surface_pressure, length=-1, time=-2, mass=1)
    if self.p_s_ref <= 0:
      raise ValueError('`reference_surface_pressure` must be positive.')
    a = ps.nondim(self.coords.vertical.a_boundaries * HPA_IN_PA, length=-1, time=-2, mass=1)
    self.levels = HybridCoordinates(a, self.coords.vertical.b_boundaries)
    self.nondim_coords = pe_lib.Coords(self.coords.horizontal, self.levels)

  def _vadv_dp(self, w, x, dp):
    return hybrid_vertical_advection(w, x, dp)

  def _t_omega_over_p(self, tfield, g_term, v_dot_grad_ln_p, ps):
    """T (v . grad ln p - (alpha F_k + alpha F_{k-1}) / dp) with F the running
    sum of g dp and alpha at the reference pressure (:2032-2056)."""
    dp = self.levels.layer_thickness(ps)
    f = np.cumsum(g_term * dp, axis=0)
    alpha, _ = vertical_coefficients(self.levels, self.p_s_ref)
    af = alpha[:, None, None] * f
    af_prev = np.concatenate([np.zeros_like(af[:1]), af[:-1]], axis=0)
    return tfield * (v_dot_grad_ln_p - (af + af_prev) / dp)

  def curl_and_div_tendencies(self, aux, ps):
    """:2058-2121, with the pressure-gradient coefficient of the full
    (virtual) temperature minus its linear reference part."""
    grid = self.coords.horizontal
    sec2 = grid.sec2_lat
    u, v = aux.cos_lat_u
    total_vort = aux.vorticity + self.coriolis_parameter
    nvu = -v * total_vort * sec2
    nvv = u * total_vort * sec2
    dp = aux.layer_pressure_thickness
    if self.include_vertical_advection:
      sdu = -self._vadv_dp(aux.mass_flux_full, u, dp)
      sdv = -self._vadv_dp(aux.mass_flux_full, v, dp)
    else:
      sdu = sdv = 0
    R = self.physics_specs.R
    coef = pgf_lps_coefficient(aux.temperature_variation + self.T_ref, ps, self.levels, R)
    coef = coef * self._virtual_temperature_adjustment(aux)
    coef = coef - pgf_lps_coefficient(np.asarray(self.reference_temperature, np.float64),
                                      self.p_s_ref, self.levels, R)[:, None, None]
    gu, gv = aux.cos_lat_grad_log_sp
    cu = grid.to_modal(nvu + (sdu + coef * gu) * sec2)
    cv = grid.to_modal(nvv + (sdv + coef * gv) * sec2)
    return (-grid.curl_cos_lat((cu, cv), clip=False), -grid.div_cos_lat((cu, cv), clip=False))

  def nodal_temperature_vertical_tendency(self, aux):
    """:2123-2144."""
    dp = aux.layer_pressure_thickness
    tend = 0
    if self.include_vertical_advection:
      tend = self._vadv_dp(aux.mass_flux_full, aux.temperature_variation, dp)
    if np.unique(self.T_ref.ravel()).size > 1:
      tend = tend + self._vadv_dp(aux.mass_flux_explicit, self.T_ref, dp)
    return tend

  def nodal_temperature_adiabatic_tendency(self, aux, ps):
    """:2146-2224."""
    levels = self.levels
    alpha, log_ratio = vertical_coefficients(levels, ps)
    b_upper = levels.b_boundaries[:-1, None, None]
    db = np.diff(levels.b_boundaries)[:, None, None]
    weight = log_ratio * b_upper + alpha * db
    dp = np.maximum(levels.layer_thickness(ps), 1e-6 * ps)
    v_grad_ln_p = aux.u_dot_grad_log_sp * ps * weight / dp
    g_full = aux.divergence + aux.u_dot_grad_log_sp * ps * db / levels.layer_thickness(ps)
    g_expl = g_full - aux.divergence
    ps_ = self.physics_specs
    mean_t = self._t_omega_over_p(self.T_ref, g_expl, v_grad_ln_p, ps)
    if self.humidity_key is None:
      var_t = self._t_omega_over_p(aux.temperature_variation, g_full, v_grad_ln_p, ps)
      return ps_.kappa * (mean_t + var_t)
    gcr = ps_.R_vapor / ps_.R
    hcr = ps_.Cp_vapor / ps_.Cp
    q = self._q(aux)
    vtc = aux.temperature_variation * ((1 + (gcr - 1) * q) / (1 + (hcr - 1) * q))
    hrc = self.T_ref * (((gcr - hcr) * q) / (1 + (hcr - 1) * q))
    part = self._t_omega_over_p(vtc + hrc, g_full, v_grad_ln_p, ps)
    return ps_.kappa * (mean_t + part)

  def nodal_log_pressure_tendency(self, aux, ps):
    """-(sum div da (1/ps - 1/ps_ref) + sum v.grad ln ps db) (:2226-2248)."""
    da = self.levels.pressure_thickness[:, None, None]
    db = self.levels.sigma_thickness[:, None, None]
    div_a = np.sum(aux.divergence * da * (1 / ps - 1 / self.p_s_ref), axis=0)
    udg_b = np.sum(aux.u_dot_grad_log_sp * db, axis=0)
    return -(div_a + udg_b)

  def explicit_terms(self, state: State) -> State:
    """:2250-2362."""
    grid = self.coords.horizontal
    aux = compute_diagnostic_state_hybrid(state, self.nondim_coords)
    ps = np.exp(grid.to_nodal(state.log_surface_pressure))
    vort_dot, div_dot = self.curl_and_div_tendencies(aux, ps)
    ke = self.kinetic_energy_tendency(aux)
    oro = self.orography_tendency()
    R = self.physics_specs.R
    temperature = self.T_ref + aux.temperature_variation
    geo_full = get_geopotential_diff_hybrid(
        temperature * self._virtual_temperature_adjustment(aux), self.levels, R, ps)
    geo_lin = pe_lib.vertical_matvec(
        get_geopotential_weights_hybrid(self.levels, R, self.p_s_ref), aux.temperature_variation)
    geo_resid = -grid.laplacian(grid.to_modal(geo_full - geo_lin))
    dT_hn, dT_hm = self.horizontal_scalar_advection(aux.temperature_variation, aux)
    tr_h = {k: self.horizontal_scalar_advection(v, aux) for k, v in aux.tracers.items()}
    dT_v = self.nodal_temperature_vertical_tendency(aux)
    dT_a = self.nodal_temperature_adiabatic_tendency(aux, ps)
    lsp = self.nodal_log_pressure_tendency(aux, ps)
    if self.include_vertical_advection:
      tr_v = {k: self._vadv_dp(aux.mass_flux_full, v, aux.layer_pressure_thickness)
              for k, v in aux.tracers.items()}
    else:
      tr_v = {k: 0 for k in aux.tracers}
    tend = State(
        vorticity=vort_dot,
        divergence=div_dot + ke + oro + geo_resid,
        temperature_variation=grid.to_modal(dT_hn + dT_v + dT_a) + dT_hm,
        log_surface_pressure=grid.to_modal(lsp)[None] if lsp.ndim == 2 else grid.to_modal(lsp),
        tracers={k: grid.to_modal(tr_v[k] + tr_h[k][0]) + tr_h[k][1] for k in aux.tracers},
        sim_time=None if state.sim_time is None else 1.0,
    )
    return tend.map(lambda x: grid.clip_wavenumbers(x) if isinstance(x, np.ndarray) else x)

  def implicit_terms(self, state: State) -> State:
    """:2364-2417 (dense vertical matvecs)."""
    grid = self.coords.horizontal
    ps_ = self.physics_specs
    g = get_geopotential_weights_hybrid(self.levels, ps_.R, self.p_s_ref)
    c = pgf_lps_coefficient(np.asarray(self.reference_temperature, np.float64), self.p_s_ref,
                            self.levels, ps_.R)
    div = -grid.laplacian(pe_lib.vertical_matvec(g, state.temperature_variation)
                          + c[:, None, None] * state.log_surface_pressure)
    h = -get_temperature_implicit_weights_hybrid(self.levels, self.reference_temperature,
                                                 ps_.kappa, self.p_s_ref)
    eff = self.levels.pressure_thickness / self.p_s_ref + self.levels.sigma_thickness
    return State(
        vorticity=np.zeros_like(state.vorticity),
        divergence=div,
        temperature_variation=pe_lib.vertical_matvec(h, state.divergence),
        log_surface_pressure=-pe_lib.vertical_matvec(eff[None], state.divergence),
        tracers={k: np.zeros_like(v) for k, v in state.tracers.items()},
        sim_time=None if state.sim_time is None else 0.0,
    )

  def implicit_inverse(self, state: State, step_size: float) -> State:
    """:2419-2489."""
    if not isinstance(step_size, (float, int, np.floating)):
      raise TypeError('`step_size` must be concrete')
    ps_ = self.physics_specs
    mat = implicit_term_matrix_hybrid(step_size, self.nondim_coords, self.reference_temperature,
                                      ps_.kappa, ps_.R, self.p_s_ref)
    inv = np.linalg.inv(mat)
    k = self.levels.layers
    sl = (slice(0, k), slice(k, 2 * k), slice(2 * k, 2 * k + 1))
    src = (state.divergence, state.temperature_variation, state.log_surface_pressure)
    vm = pe_lib.vertical_matvec_per_wavenumber
    out = [sum(vm(inv[:, r, c], x) for c, x in zip(sl, src)) for r in sl]
    return State(state.vorticity, out[0], out[1], out[2], dict(state.tracers), state.sim_time)

  def divergence_tendency_due_to_humidity(self, state, aux):
    """Zero: the virtual temperature enters the pressure-gradient coefficient
    and the geopotential directly (:2491-2504)."""
    return np.zeros_like(state.divergence)

  def vorticity_tendency_due_to_humidity(self, state, aux):
    """Zero (:2506-2518)."""
    return np.zeros_like(state.vorticity)
