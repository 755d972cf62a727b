"""GPU parity of PrimitiveEquationsHybrid (fused hybrid nodal kernel, hybrid
modal prep, FP64 vertical-mode implicit kernels with the hybrid matrices)
against the FP64 oracle on the same seeded inputs:

* explicit terms, FP64, dry / moist with clouds / level-varying T_ref: 1e-10
  (divergence 2e-9, its rounding floor being ~4e-9);
* explicit terms in FP32 against the FP64 oracle: 2e-5 normwise (divergence
  2e-2 and ln ps 5e-4: the rounding floors of ps = exp(ln ps) at FP32);
* implicit terms and implicit inverse, FP64: 1e-11;
* a 2-member batch equals the members evaluated one by one;
* 12 SIL3 steps (fused IMEX step on flat buffers), FP64: 1e-9;
* gradients through the fused kernel equal those of the tensor-op path
  (fused_nodal=False) and central finite differences.
"""
import numpy as np
import pytest
import torch

from oracle import hybrid as ohyb
from oracle import pe as ope
from oracle import sht as osht
from oracle import time_integration as oti
from oracle import units as ounits
from dinosaur_amd import coordinate_systems
from dinosaur_amd import hybrid_coordinates as hc
from dinosaur_amd import primitive_equations as pe
from dinosaur_amd import spherical_harmonic as sh
from dinosaur_amd import time_integration as ti
from dinosaur_amd import units as units_lib

pytestmark = pytest.mark.gpu

FIELDS = ('vorticity', 'divergence', 'temperature_variation', 'log_surface_pressure')
CLOUDS = ('specific_cloud_liquid_water_content', 'specific_cloud_ice_water_content')


def _rel(a, b):
  a = np.asarray(a, np.float64)
  b = np.asarray(b, np.float64)
  return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _setup(k, moist, tref, dtype, dev, members=0, fused=True, seed=3):
  og = osht.Grid.T21()
  olev = ohyb.HybridCoordinates.analytic_levels(k, p_top=5.0)
  plev = hc.HybridCoordinates.analytic_levels(k, p_top=5.0)
  ops_ = ounits.SimUnits.from_si()
  rs = np.random.RandomState(seed)
  _, l = og.modal_mesh
  oro = rs.normal(size=og.modal_shape) * og.mask * 1e-4 / (l + 1) ** 2
  tref = np.full(k, 260.0) if tref is None else np.asarray(tref, np.float64)
  kw = dict(humidity_key='specific_humidity', cloud_keys=CLOUDS) if moist else {}
  oeq = ohyb.PrimitiveEquationsHybrid(tref, oro, ope.Coords(og, olev), ops_, **kw)
  peq = pe.PrimitiveEquationsHybrid(tref, torch.as_tensor(oro, device=dev),
                                    coordinate_systems.CoordinateSystem(sh.Grid.T21(), plev),
                                    units_lib.SimUnits.from_si(), fused_nodal=fused, **kw)

  def field(n, amp):
    lead = (members,) if members else ()
    return rs.normal(size=lead + (n,) + og.modal_shape) * og.mask * amp / (l + 1) ** 2

  lnps = field(1, 0.05)
  lnps[..., 0, 0, 0] += np.sqrt(4 * np.pi) * np.log(ops_.nondim(1e5, length=-1, time=-2, mass=1))
  tr = {}
  if moist:
    tr = {'specific_humidity': np.abs(field(k, 1e-3)), CLOUDS[0]: np.abs(field(k, 1e-5)),
          CLOUDS[1]: np.abs(field(k, 1e-5))}
    tr['specific_humidity'][..., 0, 0] += 0.01 * np.sqrt(4 * np.pi)
  arrs = [field(k, 1e-2), field(k, 1e-2), field(k, 1.0), lnps]
  ost = ope.State(*arrs, tracers=tr)
  T = lambda a: torch.as_tensor(a, device=dev).to(dtype)
  gst = pe.State(*(T(a) for a in arrs), tracers={n: T(v) for n, v in tr.items()})
  return oeq, peq, ost, gst


def _cmp(got, ref, tol, what, member=None, div_tol=None, lnps_tol=None):
  """Per-field normwise comparison.  div_tol / lnps_tol (default tol) for the
  divergence and ln ps tendencies, whose rounding floors are higher: ps =
  exp(ln ps) with nondimensional ln ps ~ 45 and the Laplacian of the
  geopotential amplify round-off (tests/test_oracle_hybrid.py::
  test_hybrid_divergence_rounding_floor: one FP64 round trip of ln ps moves
  the divergence by ~4e-9), and the ln ps tendency carries 1/ps - 1/ps_ref."""
  sel = (lambda a: a[member]) if member is not None else (lambda a: a)
  special = {'divergence': div_tol, 'log_surface_pressure': lnps_tol}
  for f in FIELDS:
    err = _rel(getattr(got, f).cpu().numpy(), sel(getattr(ref, f)))
    print(f'  {what} {f}: rel err {err:.2e}', flush=True)
    t = special[f] if special.get(f) is not None else tol
    assert err < t, (what, f, err)
  for n in ref.tracers:
    err = _rel(got.tracers[n].cpu().numpy(), sel(ref.tracers[n]))
    print(f'  {what} tracer {n}: rel err {err:.2e}', flush=True)
    assert err < tol, (what, n, err)


@pytest.mark.parametrize('moist,tref', [(False, None), (True, None),
                                        (False, np.linspace(220.0, 290.0, 12))])
def test_hybrid_explicit_terms_fp64(moist, tref, gpu_device):
  print(f'[test] hybrid explicit fp64 moist={moist} tref_varies={tref is not None}', flush=True)
  oeq, peq, ost, gst = _setup(12, moist, tref, torch.float64, gpu_device)
  _cmp(peq.explicit_terms(gst), oeq.explicit_terms(ost), 1e-10, 'explicit', div_tol=2e-9)


def test_hybrid_explicit_terms_fp32(gpu_device):
  print('[test] hybrid explicit fp32 vs fp64 oracle (moist)', flush=True)
  oeq, peq, ost, gst = _setup(12, True, None, torch.float32, gpu_device)
  # The same conditioning at FP32 round-off: the oracle itself, given FP32
  # state arrays, is off its FP64 result by 3.1e-3 (divergence) and 6.1e-5
  # (ln ps tendency); the other fields by <= 4.3e-6.
  _cmp(peq.explicit_terms(gst), oeq.explicit_terms(ost), 2e-5, 'explicit fp32', div_tol=2e-2,
       lnps_tol=5e-4)


def test_hybrid_implicit_terms_and_inverse_fp64(gpu_device):
  print('[test] hybrid implicit terms / inverse fp64', flush=True)
  oeq, peq, ost, gst = _setup(12, False, np.linspace(230.0, 290.0, 12), torch.float64,
                              gpu_device)
  _cmp(peq.implicit_terms(gst), oeq.implicit_terms(ost), 1e-11, 'implicit terms')
  for eta in (0.01, 0.2):
    _cmp(peq.implicit_inverse(gst, eta), oeq.implicit_inverse(ost, eta), 1e-11,
         f'inverse eta={eta}')


def test_hybrid_members_match_single(gpu_device):
  print('[test] hybrid 2-member batch vs members one by one (moist)', flush=True)
  oeq, peq, ost, gst = _setup(8, True, None, torch.float64, gpu_device, members=2)
  got = peq.explicit_terms(gst)
  for e in range(2):
    one = pe.State(*(getattr(gst, f)[e] for f in FIELDS),
                   tracers={n: v[e] for n, v in gst.tracers.items()})
    ref = peq.explicit_terms(one)
    for f in FIELDS:
      assert torch.allclose(getattr(got, f)[e], getattr(ref, f), rtol=0, atol=1e-14), f
    ost_e = ope.State(*(getattr(ost, f)[e] for f in FIELDS),
                      tracers={n: v[e] for n, v in ost.tracers.items()})
    _cmp(ref, oeq.explicit_terms(ost_e), 1e-10, f'member {e}', div_tol=2e-9)


def test_hybrid_sil3_rollout_fp64(gpu_device):
  print('[test] hybrid SIL3 12 steps fp64 vs oracle (moist, fused IMEX step)', flush=True)
  oeq, peq, ost, gst = _setup(10, True, None, torch.float64, gpu_device)
  dt = units_lib.SimUnits.from_si().nondimensionalize(600 * units_lib.scales.units.s)
  step = ti.imex_rk_sil3(peq, dt)
  ostep = oti.imex_rk_sil3(oeq, dt)
  y = pe.to_flat(gst)
  oy = ost
  for _ in range(12):
    y = step(y)
    oy = ostep(oy)
  _cmp(y, oy, 1e-9, 'rollout')


def test_hybrid_gradients_fused_vs_tensor_ops_and_fd(gpu_device):
  print('[test] hybrid gradient: fused kernel vs tensor-op path vs finite differences',
        flush=True)
  _, peq, _, gst = _setup(8, True, None, torch.float64, gpu_device)
  _, peq_u, _, _ = _setup(8, True, None, torch.float64, gpu_device, fused=False)
  rs = np.random.RandomState(11)
  w = {f: torch.as_tensor(rs.normal(size=getattr(gst, f).shape), device=gpu_device)
       for f in FIELDS}

  def loss(eq, leaves):
    st = pe.State(*leaves, tracers=dict(gst.tracers))
    out = eq.explicit_terms(st)
    return sum((getattr(out, f) * w[f]).sum() for f in FIELDS)

  def grads(eq):
    leaves = [getattr(gst, f).detach().clone().requires_grad_(True) for f in FIELDS]
    return torch.autograd.grad(loss(eq, leaves), leaves)

  g_f, g_u = grads(peq), grads(peq_u)
  for f, a, b in zip(FIELDS, g_f, g_u):
    err = _rel(a.cpu().numpy(), b.cpu().numpy())
    print(f'  fused vs tensor-op grad {f}: {err:.2e}', flush=True)
    assert err < 1e-11, (f, err)
  # central differences along one seeded direction
  d = [torch.as_tensor(rs.normal(size=getattr(gst, f).shape), device=gpu_device)
       * (1e-2 if f != 'temperature_variation' else 1.0) for f in FIELDS]
  base = [getattr(gst, f).detach() for f in FIELDS]
  eps = 1e-5
  with torch.no_grad():
    lp = loss(peq, [b + eps * v for b, v in zip(base, d)]).item()
    lm = loss(peq, [b - eps * v for b, v in zip(base, d)]).item()
  fd = (lp - lm) / (2 * eps)
  an = sum((g * v).sum().item() for g, v in zip(g_f, d))
  print(f'  directional derivative: adjoint {an:.10e} fd {fd:.10e}', flush=True)
  assert abs(an - fd) <= 1e-6 * abs(fd)
