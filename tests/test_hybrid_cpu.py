"""Host-side parts of PrimitiveEquationsHybrid against the oracle (no GPU).

* the hybrid coordinates' constructors and metadata against the oracle's
  (which the reference's known answers pin, tests/test_oracle_hybrid.py);
* the vertical matrices the implicit kernels receive (G, H, the
  pressure-gradient coefficient and the ln ps weights) against the oracle's
  block implicit matrix;
* the nodal terms written as tensor operations -- the backward of the fused
  hybrid kernel, and the whole nodal path for layouts the fused kernels do
  not address -- fed the oracle's nodal diagnostic state and assembled with
  the oracle's transforms, against the oracle's explicit_terms (FP64, 1e-11: the kinetic energy and the geopotential residual share one transform here),
  dry, moist with clouds, and with a level-varying reference temperature.
"""
import numpy as np
import pytest
import torch

from oracle import hybrid as ohyb
from oracle import pe as ope
from oracle import sht as osht
from oracle import units as ounits
from dinosaur_amd import coordinate_systems
from dinosaur_amd import hybrid_coordinates as hc
from dinosaur_amd import primitive_equations as pe
from dinosaur_amd import spherical_harmonic
from dinosaur_amd import units as units_lib


def _levels(k):
  return ohyb.HybridCoordinates.analytic_levels(k, p_top=5.0), hc.HybridCoordinates.analytic_levels(
      k, p_top=5.0)


def test_coordinates_match_oracle():
  for args in ((8,), (12, 2.0, 1000.0, 2.0, 1.5)):
    o = ohyb.HybridCoordinates.analytic_levels(*args)
    p = hc.HybridCoordinates.analytic_levels(*args)
    np.testing.assert_allclose(p.a_boundaries, o.a_boundaries, rtol=0, atol=1e-12)
    np.testing.assert_allclose(p.b_boundaries, o.b_boundaries, rtol=0, atol=1e-15)
    for ps in (1000.0, 870.0):
      np.testing.assert_allclose(p.layer_thickness(ps), o.layer_thickness(ps), rtol=1e-14)
      np.testing.assert_allclose(p.get_sigma_centers(ps), o.get_sigma_centers(ps), rtol=1e-14)
  e = hc.HybridCoordinates.analytic_ecmwf_like(40)
  assert e.layers == 40 and e.b_boundaries[0] == 0 and e.b_boundaries[-1] == 1.0
  assert np.all(np.diff(e.pressure_boundaries(1013.25)) > 0)
  assert np.all(e.a_boundaries >= 0)
  # pure pressure above the transition, hybrid below
  p = e.pressure_boundaries(1013.25)
  assert np.all(e.b_boundaries[p <= 80.0] == 0)
  with pytest.raises(ValueError, match='got 2 and 3'):
    hc.HybridCoordinates([1, 2], [1, 2, 3])
  s = hc.HybridCoordinates.from_sigma_levels(
      __import__('dinosaur_amd.sigma_coordinates', fromlist=['x']).SigmaCoordinates.equidistant(6))
  assert s == hc.HybridCoordinates(np.zeros(7), np.linspace(0, 1, 7, dtype=np.float32))
  sig = e.to_approx_sigma_coords(10)
  assert sig.layers == 10 and sig.boundaries[0] == 0 and sig.boundaries[-1] == 1


def test_operators_match_oracle():
  o, p = _levels(10)
  rs = np.random.RandomState(0)
  x = rs.normal(size=(10, 4, 3))
  ps = 900 + 50 * rs.uniform(size=(1, 4, 3))
  xt, pst = torch.as_tensor(x), torch.as_tensor(ps)
  np.testing.assert_allclose(hc.cumulative_integral_over_pressure(xt, pst, p).numpy(),
                             ohyb.cumulative_integral_over_pressure(x, ps, o), rtol=1e-13)
  np.testing.assert_allclose(
      hc.cumulative_integral_over_pressure(xt, pst, p, downward=False).numpy(),
      ohyb.cumulative_integral_over_pressure(x, ps, o, downward=False), rtol=1e-13)
  np.testing.assert_allclose(hc.integral_over_pressure(xt, pst, p).numpy(),
                             ohyb.integral_over_pressure(x, ps, o), rtol=1e-13)
  np.testing.assert_allclose(hc.cumulative_integral_over_sigma(xt, p).numpy(),
                             ohyb.cumulative_integral_over_sigma(x, o), rtol=1e-13, atol=1e-15)
  np.testing.assert_allclose(hc.integral_over_sigma(xt, p).numpy(),
                             ohyb.integral_over_sigma(x, o), rtol=1e-13)
  w = torch.as_tensor(rs.normal(size=(9, 4, 3)))
  np.testing.assert_allclose(hc.centered_vertical_advection(w, xt, p).numpy(),
                             ohyb.centered_vertical_advection(w.numpy(), x, o), rtol=1e-13)
  with pytest.raises(ValueError):
    hc.integral_over_sigma(xt[:3], p)


GRID = 'T21'


def _pair(k, moist=False, tref=None):
  og = osht.Grid.T21()
  grid = spherical_harmonic.Grid.T21()
  olev, plev = _levels(k)
  ops_ = ounits.SimUnits.from_si()
  pps = units_lib.SimUnits.from_si()
  rs = np.random.RandomState(3)
  _, l = og.modal_mesh
  oro = rs.normal(size=og.modal_shape) * og.mask * 1e-4 / (l + 1) ** 2
  tref = np.full(k, 260.0) if tref is None else tref
  kw = {}
  if moist:
    kw = dict(humidity_key='specific_humidity',
              cloud_keys=('specific_cloud_liquid_water_content', 'specific_cloud_ice_water_content'))
  oeq = ohyb.PrimitiveEquationsHybrid(tref, oro, ope.Coords(og, olev), ops_, **kw)
  peq = pe.PrimitiveEquationsHybrid(tref, torch.as_tensor(oro),
                                    coordinate_systems.CoordinateSystem(grid, plev), pps, **kw)

  def field(n, amp):
    return rs.normal(size=(n,) + og.modal_shape) * og.mask * amp / (l + 1) ** 2

  lnps = field(1, 0.05)
  lnps[0, 0, 0] += np.sqrt(4 * np.pi) * np.log(ops_.nondim(1e5, length=-1, time=-2, mass=1))
  tr = {}
  if moist:
    tr = {'specific_humidity': np.abs(field(k, 1e-3)) + 0.0,
          'specific_cloud_liquid_water_content': np.abs(field(k, 1e-5)),
          'specific_cloud_ice_water_content': np.abs(field(k, 1e-5))}
    tr['specific_humidity'][:, 0, 0] += 0.01 * np.sqrt(4 * np.pi)
  st = ope.State(field(k, 1e-2), field(k, 1e-2), field(k, 1.0), lnps, tracers=tr)
  return og, oeq, peq, st


def _assemble(og, outs, names):
  """Modal tendencies from the nodal products (the sigma epilogue's algebra,
  written with the oracle's operators)."""
  cu, cv, ke, temp, ut, vt, lsp = (og.to_modal(o) for o in outs[:7])
  div = -og.div_cos_lat((cu, cv), clip=False) - og.laplacian(ke)
  vort = -og.curl_cos_lat((cu, cv), clip=False)
  t = temp - og.div_cos_lat((ut, vt), clip=False)
  trs = {}
  for i, n in enumerate(names):
    a, b, c = (og.to_modal(o) for o in outs[7 + 3 * i:10 + 3 * i])
    trs[n] = a - og.div_cos_lat((b, c), clip=False)
  return vort, div, t, lsp, trs


@pytest.mark.parametrize('moist,tref', [(False, None), (True, None),
                                        (False, np.linspace(220.0, 290.0, 9))])
def test_nodal_terms_match_oracle_explicit_terms(moist, tref):
  k = 9
  og, oeq, peq, st = _pair(k, moist, tref)
  ref = oeq.explicit_terms(st)
  aux = ohyb.compute_diagnostic_state_hybrid(st, oeq.nondim_coords)
  lnps_n = og.to_nodal(st.log_surface_pressure)
  names = list(st.tracers)
  T = lambda a: torch.as_tensor(np.ascontiguousarray(a))
  outs = peq._nodal_terms(T(aux.vorticity), T(aux.divergence), T(aux.temperature_variation),
                          {n: T(aux.tracers[n]) for n in names}, T(aux.cos_lat_u[0]),
                          T(aux.cos_lat_u[1]), T(aux.cos_lat_grad_log_sp[0]),
                          T(aux.cos_lat_grad_log_sp[1]), T(lnps_n))
  assert len(outs) == 7 + 3 * len(names)
  vort, div, t, lsp, trs = _assemble(og, [o.numpy() for o in outs], names)
  div = div - oeq.physics_specs.g * og.laplacian(oeq.orography)
  got = {'vorticity': vort, 'divergence': div, 'temperature_variation': t,
         'log_surface_pressure': lsp}
  for f, g in got.items():
    g = og.clip_wavenumbers(g)
    r = getattr(ref, f)
    err = np.abs(g - r).max() / np.abs(r).max()
    assert err < 1e-11, (f, err)
  for n in names:
    g = og.clip_wavenumbers(trs[n])
    r = ref.tracers[n]
    assert np.abs(g - r).max() / np.abs(r).max() < 1e-11, n


def test_nodal_buffer_terms_group_order():
  """The (fields, I, J) buffer form splits and concatenates in the fused
  kernel's group order, members outermost within each group."""
  og, oeq, peq, st = _pair(5, moist=True)
  names = list(st.tracers)
  E, K = 2, 5
  I, J = og.nodal_shape
  rs = np.random.RandomState(1)
  groups = [K] * (3 + len(names)) + [K, K, 1, 1, 1]
  parts = [rs.normal(size=(E, n, I, J)) for n in groups]
  parts[3 + len(names) + 4] = np.log(np.full((E, 1, I, J), 1.0)) + rs.normal(size=(E, 1, I, J)) * 0.01
  parts[3] = np.abs(parts[3]) * 1e-3  # humidity
  buf = torch.as_tensor(np.concatenate([p.reshape(-1, I, J) for p in parts]))
  out = peq._nodal_buffer_terms(buf, E, names)
  for e in range(E):
    tp = [torch.as_tensor(p[e]) for p in parts]
    nt = len(names)
    ref = peq._nodal_terms(tp[0], tp[1], tp[2], dict(zip(names, tp[3:3 + nt])), *tp[3 + nt:])
    off = 0
    for r in ref:
      n = r.shape[0]
      got = out[off + e * n: off + (e + 1) * n]
      np.testing.assert_allclose(got.numpy(), r.numpy(), rtol=1e-14, atol=1e-14)
      off += E * n


def test_implicit_matrices_match_oracle_block_matrix():
  k = 9
  og, oeq, peq, st = _pair(k, tref=np.linspace(230.0, 290.0, k))
  op = peq._implicit
  eta = 0.07
  mat = ohyb.implicit_term_matrix_hybrid(eta, oeq.nondim_coords, oeq.reference_temperature,
                                         oeq.physics_specs.kappa, oeq.physics_specs.R,
                                         oeq.p_s_ref)
  lam = og.laplacian_eigenvalues
  for li in (1, 7, 20):
    np.testing.assert_allclose(mat[li, :k, k:2 * k], eta * lam[li] * op.G, rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(mat[li, :k, 2 * k], eta * lam[li] * op.r_gas * op.tref,
                               rtol=1e-13)
    np.testing.assert_allclose(mat[li, k:2 * k, :k], eta * op.H, rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(mat[li, 2 * k, :k], eta * op.dsig, rtol=1e-13)
  # the vertical modes diagonalise A = G H + c w^T
  a = op.G @ op.H + op.r_gas * np.outer(op.tref, op.dsig)
  np.testing.assert_allclose(op.V @ np.diag(op.d) @ op.Vinv, a, rtol=1e-9, atol=1e-12 * abs(a).max())
