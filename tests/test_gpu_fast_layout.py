"""The reference's fast modal layout (FastSphericalHarmonics, rows m = [0, 0',
+1, -1, ...], spherical_harmonic.py:411-621) on the GPU: transforms, grid
operators and primitive-equation tendencies equal the Real layout's with the
row mapping (bitwise for the transforms, which share the kernels), the padded
variant keeps its padding zero, and the basis matches the Real one."""
import numpy as np
import pytest
import torch

from dinosaur_amd import coordinate_systems, primitive_equations as pe
from dinosaur_amd import sigma_coordinates, spherical_harmonic as sh, time_integration as ti
from dinosaur_amd import units

pytestmark = pytest.mark.gpu


def to_fast(x, pad_rows=0, pad_l=0):
  z = torch.zeros_like(x[..., :1, :])
  out = torch.cat([x[..., :1, :], z, x[..., 1:, :]], -2)
  return torch.nn.functional.pad(out, (0, pad_l, 0, pad_rows))


def to_real(x, M, L):
  idx = torch.tensor([0] + list(range(2, 2 * M)), device=x.device)
  return torch.index_select(x[..., :, :L], -2, idx)


def _modal(grid, n, seed, device, dtype):
  rs = np.random.RandomState(seed)
  _, l = grid.modal_mesh
  return torch.as_tensor(rs.normal(size=(n,) + grid.modal_shape) * grid.mask / (l + 1) ** 1.5,
                         device=device, dtype=dtype)


@pytest.mark.parametrize('dtype', [torch.float64, torch.float32])
@pytest.mark.parametrize('multiple', [1, 8])
def test_fast_transforms_match_real(dtype, multiple, gpu_device):
  print(f'[test] fast layout transforms {dtype} base_shape_multiple={multiple}', flush=True)
  real = sh.Grid.T21()
  fast = sh.Grid.T21(spherical_harmonics_impl=lambda **kw: sh.FastSphericalHarmonics(
      **kw, base_shape_multiple=multiple))
  M, L = real.longitude_wavenumbers, real.total_wavenumbers
  pr, pl = fast.spherical_harmonics.modal_padding
  ni, nj = fast.spherical_harmonics.nodal_padding
  x = _modal(real, 5, 0, gpu_device, dtype)
  xf = to_fast(x, pr, pl)
  assert tuple(xf.shape[-2:]) == fast.modal_shape
  assert not (xf * torch.as_tensor(~fast.mask, device=gpu_device)).any()
  y = real.to_nodal(x)
  yf = fast.to_nodal(xf)
  assert torch.equal(yf[..., :real.nodal_shape[0], :real.nodal_shape[1]], y)
  if ni or nj:
    assert not yf[..., real.nodal_shape[0]:, :].any() and not yf[..., :, real.nodal_shape[1]:].any()
  zf = fast.to_modal(yf)
  z = real.to_modal(y)
  assert torch.equal(to_real(zf, M, L), z)
  assert not zf[..., 1, :].any() and not zf[..., 2 * M:, :].any() and not zf[..., L:].any()


def test_fast_grid_operators_match_real(gpu_device):
  print('[test] fast layout grid operators', flush=True)
  real = sh.Grid.T21()
  fast = sh.Grid.T21(spherical_harmonics_impl=sh.FastSphericalHarmonics)
  M, L = real.longitude_wavenumbers, real.total_wavenumbers
  x = _modal(real, 3, 1, gpu_device, torch.float64)
  y = _modal(real, 3, 2, gpu_device, torch.float64)
  xf, yf = to_fast(x), to_fast(y)
  for name in ('d_dlon', 'cos_lat_d_dlat', 'sec_lat_d_dlat_cos2', 'laplacian',
               'inverse_laplacian', 'clip_wavenumbers'):
    a = getattr(fast, name)(xf)
    b = getattr(real, name)(x)
    err = (to_real(a, M, L) - b).abs().max().item()
    print(f'  {name}: {err:.1e}', flush=True)
    assert err <= 1e-14 * max(b.abs().max().item(), 1.0), name
    assert not a[..., 1, :].any(), name
  uf, vf = sh.vor_div_to_uv_nodal(fast, xf, yf)
  u, v = sh.vor_div_to_uv_nodal(real, x, y)
  assert (uf - u).abs().max().item() <= 1e-12 * u.abs().max().item()
  assert (vf - v).abs().max().item() <= 1e-12 * v.abs().max().item()
  b_fast, b_real = fast.spherical_harmonics.basis, real.spherical_harmonics.basis
  assert np.all(b_fast.f[:, 1] == 0)
  np.testing.assert_array_equal(b_fast.f[:, 2:], b_real.f[:, 1:])
  np.testing.assert_array_equal(b_fast.p, b_real.p[0::2])
  np.testing.assert_array_equal(b_fast.w, b_real.w)


def test_fast_layout_primitive_equations_match_real(gpu_device):
  """Moist T21 L6 explicit and implicit terms, and one SIL3 step, in the fast
  layout (tensor-op explicit terms) against the Real layout (fused kernels),
  FP64."""
  print('[test] fast layout primitive equations', flush=True)
  K = 6
  out = {}
  for name, impl in (('real', sh.RealSphericalHarmonics), ('fast', sh.FastSphericalHarmonics)):
    grid = sh.Grid.T21(spherical_harmonics_impl=impl)
    coords = coordinate_systems.CoordinateSystem(
        grid, sigma_coordinates.SigmaCoordinates.equidistant(K))
    rs = np.random.RandomState(21)
    real_grid = sh.Grid.T21()
    _, l = real_grid.modal_mesh
    def f(shape, amp):
      return torch.as_tensor(rs.normal(size=shape) * real_grid.mask * amp / (l + 1) ** 1.5,
                             device=gpu_device)
    ms, ss = (K,) + real_grid.modal_shape, (1,) + real_grid.modal_shape
    lnps = f(ss, 0.05)
    lnps[..., 0, 0] += 3.5449077 * 4.0
    fields = [f(ms, 1e-1), f(ms, 1e-1), f(ms, 5.0), lnps]
    q = f(ms, 1e-3)
    oro = rs.normal(size=real_grid.modal_shape) * real_grid.mask * 1e-4
    if name == 'fast':
      fields = [to_fast(t) for t in fields]
      q = to_fast(q)
      oro = to_fast(torch.as_tensor(oro)).numpy()
    st = pe.State(*fields, tracers={'q': q}, sim_time=0.0)
    eq = pe.PrimitiveEquationsSigma(np.linspace(230., 280., K), oro, coords,
                                    units.SimUnits.from_si(), humidity_key='q')
    step = ti.imex_rk_sil3(eq, 0.01)
    with torch.no_grad():
      out[name] = (eq.explicit_terms(st), eq.implicit_terms(st), step(st))
  M, L = 22, 23
  for i, what in enumerate(('explicit', 'implicit', 'sil3 step')):
    a, b = out['fast'][i], out['real'][i]
    for fld in pe.State._ARRAYS:
      x, y = to_real(getattr(a, fld), M, L), getattr(b, fld)
      err = ((x - y).abs().max() / max(y.abs().max().item(), 1e-300)).item()
      print(f'  {what} {fld}: {err:.1e}', flush=True)
      assert err < 1e-11, (what, fld, err)


@pytest.mark.parametrize('multiple', [1, 8])
def test_fast_layout_matches_oracle_fast_restatement(multiple, gpu_device):
  """Against the oracle's own einsum restatement of the Fast layout (not the
  Real layout): to_nodal, to_modal, d_dlon and basis metadata (modal axes,
  mask, shapes), FP64 1e-12 with and without padding; then one SIL3 step of
  T21 L6 dry primitive equations in the Fast layout against the oracle's
  Real-layout step relabelled into Fast rows."""
  from oracle import pe as ope, sht as osht, sigma as osig, time_integration as oti
  from oracle import units as ounits
  from dinosaur_amd import workloads
  print(f'[test] fast layout vs oracle fast restatement, multiple={multiple}', flush=True)
  og = osht.Grid.T21()
  ofg = osht.FastGrid(og, base_shape_multiple=multiple)
  fast = sh.Grid.T21(spherical_harmonics_impl=lambda **kw: sh.FastSphericalHarmonics(
      **kw, base_shape_multiple=multiple))
  assert tuple(fast.modal_shape) == ofg.modal_shape
  assert tuple(fast.nodal_shape) == ofg.nodal_shape
  np.testing.assert_array_equal(fast.modal_axes[0], ofg.modal_axes[0])
  np.testing.assert_array_equal(np.asarray(fast.mask), ofg.mask)
  rs = np.random.RandomState(5)
  _, l = og.modal_mesh
  xf = ofg.from_real(rs.normal(size=(4,) + og.modal_shape) * og.mask / (l + 1) ** 1.5)
  t = torch.as_tensor(xf, device=gpu_device)
  rel = lambda a, b: np.abs(a - b).max() / np.abs(b).max()
  zn = fast.to_nodal(t).cpu().numpy()
  assert rel(zn, ofg.to_nodal(xf)) < 1e-12
  assert rel(fast.to_modal(torch.as_tensor(zn, device=gpu_device)).cpu().numpy(),
             ofg.to_modal(zn)) < 1e-12
  assert rel(fast.d_dlon(t).cpu().numpy(), ofg.d_dlon(xf)) < 1e-14
  if multiple != 1:
    return
  eq, st, coords, ps = workloads.build_primitive('T21', 6, dtype=torch.float64, device=gpu_device)
  # The Fast-layout equation takes the Real one's parameters relabelled into
  # Fast rows (the seeded recipe would draw a different orography in the
  # (2M, L) shape).
  fcoords = coordinate_systems.CoordinateSystem(
      sh.Grid.T21(spherical_harmonics_impl=sh.FastSphericalHarmonics), coords.vertical)
  foro = torch.as_tensor(ofg.from_real(eq.orography.double().cpu().numpy()), device=gpu_device)
  feq = pe.PrimitiveEquationsSigma(np.asarray(eq.reference_temperature, np.float64), foro,
                                   fcoords, ps)
  ocoords = ope.Coords(og, osig.SigmaCoordinates.equidistant(6))
  oeq = ope.PrimitiveEquationsSigma(np.asarray(eq.reference_temperature, np.float64),
                                    eq.orography.double().cpu().numpy(), ocoords,
                                    ounits.SimUnits.from_si())
  fields = ('vorticity', 'divergence', 'temperature_variation', 'log_surface_pressure')
  ost = ope.State(*(getattr(st, f).cpu().numpy() for f in fields))
  dt = ps.nondimensionalize(1200 * units.scales.units.s)
  ref = oti.imex_rk_sil3(oeq, dt)(ost)
  fin = pe.State(*(torch.as_tensor(ofg.from_real(getattr(ost, f)), device=gpu_device)
                   for f in fields))
  out = ti.imex_rk_sil3(feq, dt)(fin)
  for f in fields:
    got = ofg.to_real(getattr(out, f).cpu().numpy())
    err = rel(got, getattr(ref, f))
    print(f'  fast SIL3 step {f}: rel err {err:.2e}', flush=True)
    assert err < 1e-10, (f, err)
    assert np.abs(getattr(out, f).cpu().numpy()[..., 1, :]).max() == 0
