"""GPU parity of the spherical harmonic transforms against the FP64 oracle.

FP64 mode must agree with the oracle's dense einsums to ~1e-12 (normwise);
FP32 to the FP32 rounding level.  Covers the FFT path (power-of-two I), the
direct-DFT path (odd I, odd J with an equator node), equiangular nodes, the
GPU-generated FP64 P_l^m table, and the adjoint kernels (dot-product test).
"""
import numpy as np
import pytest
import torch

from oracle import sht as osht
from dinosaur_amd import spherical_harmonic as sh

pytestmark = pytest.mark.gpu


def _rel(a, b):
  a = np.asarray(a, np.float64)
  b = np.asarray(b, np.float64)
  return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _grids():
  return {
      'T21': (sh.Grid.T21(), osht.Grid.T21()),
      'wn16_direct': (sh.Grid.with_wavenumbers(16), osht.Grid.with_wavenumbers(16)),
      'TL127': (sh.Grid.TL127(), osht.Grid.TL127()),
      'equiangular': (sh.Grid(24, 24, 96, 48, latitude_spacing='equiangular'),
                      osht.Grid(24, 24, 96, 48, latitude_spacing='equiangular')),
      # odd J (equator node), and J % 8 != 0 with L % 4 == 0: the scalar-load
      # and vector-store paths of the FP32 / split Legendre kernels
      'odd_j25': (sh.Grid(12, 13, 32, 25), osht.Grid(12, 13, 32, 25)),
      'j36_l16': (sh.Grid(16, 16, 48, 36), osht.Grid(16, 16, 48, 36)),
  }


def _modal(og, n, seed):
  rs = np.random.RandomState(seed)
  _, l = og.modal_mesh
  return rs.normal(size=(n,) + og.modal_shape) * og.mask / (l + 1) ** 1.5


@pytest.mark.parametrize('name', ['T21', 'wn16_direct', 'TL127', 'equiangular'])
def test_transforms_fp64_match_oracle(name, gpu_device):
  g, og = _grids()[name]
  print(f'[test] fp64 transforms {name}', flush=True)
  x = _modal(og, 3, 0)
  xt = torch.tensor(x, device=gpu_device)
  y = g.to_nodal(xt).cpu().numpy()
  y_ref = og.to_nodal(x)
  assert _rel(y, y_ref) < 1e-12
  z = g.to_modal(torch.tensor(y_ref, device=gpu_device)).cpu().numpy()
  z_ref = og.to_modal(y_ref)
  assert _rel(z, z_ref) < 1e-12
  # round trip recovers the modal coefficients below the top wavenumber
  # (construct() grids carry one extra, unresolved l that the model clips)
  assert _rel(z[..., :-1], x[..., :-1]) < 1e-11


@pytest.mark.parametrize('name', ['T21', 'wn16_direct', 'TL127', 'equiangular', 'odd_j25',
                                  'j36_l16'])
def test_transforms_fp32(name, gpu_device):
  g, og = _grids()[name]
  print(f'[test] fp32 transforms {name}', flush=True)
  x = _modal(og, 4, 1)
  y = g.to_nodal(torch.tensor(x, dtype=torch.float32, device=gpu_device)).cpu().numpy()
  assert _rel(y, og.to_nodal(x)) < 2e-6
  yr = og.to_nodal(x)
  z = g.to_modal(torch.tensor(yr, dtype=torch.float32, device=gpu_device)).cpu().numpy()
  assert _rel(z, og.to_modal(yr)) < 2e-6


def test_gpu_table_matches_oracle_recurrence(gpu_device):
  g = sh.Grid.TL127()
  plan = g.spherical_harmonics.plan(gpu_device)
  x, w = plan.nodes()
  folded = plan.folded_table()
  M, L, J = g.longitude_wavenumbers, g.total_wavenumbers, g.latitude_nodes
  jh = (J + 1) // 2
  ref = osht.legendre_table(M, L, x)  # same nodes -> same recurrence
  pos = 0
  err = 0.0
  for m in range(M):
    rows = folded[pos:pos + (L - m) * jh].reshape(L - m, jh)
    pos += (L - m) * jh
    err = max(err, np.abs(rows.T - ref[m, J - jh:, m:]).max())
  assert err < 1e-15
  xs, ws = osht.gauss_nodes(J)
  np.testing.assert_allclose(x, xs, atol=2e-15)
  np.testing.assert_allclose(w, ws, atol=1e-13)


def test_adjoints_dot_product(gpu_device):
  g, og = _grids()['T21']
  rs = np.random.RandomState(3)
  x = torch.tensor(_modal(og, 2, 3), device=gpu_device, requires_grad=True)
  ybar = torch.tensor(rs.normal(size=(2,) + og.nodal_shape), device=gpu_device)
  y = g.to_nodal(x)
  (gx,) = torch.autograd.grad((y * ybar).sum(), x)
  # <F P x, ybar> == <x, P^T F^T ybar>; check against the oracle's explicit matrices
  f, p, w = og.basis
  gref = np.einsum('mjl,...mj->...ml', p, np.einsum('im,...ij->...mj', f, ybar.cpu().numpy()))
  assert _rel(gx.cpu().numpy(), gref) < 1e-12
  z = torch.tensor(rs.normal(size=(2,) + og.nodal_shape), device=gpu_device, requires_grad=True)
  xbar = torch.tensor(_modal(og, 2, 4), device=gpu_device)
  (gz,) = torch.autograd.grad((g.to_modal(z) * xbar).sum(), z)
  gzref = w * np.einsum('im,...mj->...ij', f, np.einsum('mjl,...ml->...mj', p, xbar.cpu().numpy()))
  assert _rel(gz.cpu().numpy(), gzref) < 1e-12


def test_batched_and_pytree_api(gpu_device):
  g, og = _grids()['T21']
  x = _modal(og, 5, 5)
  tree = {'a': torch.tensor(x[:2], device=gpu_device), 'b': (torch.tensor(x[2], device=gpu_device), 3.0)}
  out = g.to_nodal(tree)
  assert out['b'][1] == 3.0
  assert _rel(out['a'].cpu().numpy(), og.to_nodal(x[:2])) < 1e-12
  parts = g.to_nodal_batched([tree['a'], tree['b'][0]])
  assert _rel(parts[1].cpu().numpy(), og.to_nodal(x[2])) < 1e-12
  # numpy in, numpy out
  y = g.to_nodal(x[0])
  assert isinstance(y, np.ndarray)
  assert _rel(y, og.to_nodal(x[0])) < 1e-12


def test_grid_operators_match_oracle(gpu_device):
  g, og = _grids()['T21']
  x = _modal(og, 2, 6)
  xt = torch.tensor(x, device=gpu_device)
  for name in ('d_dlon', 'cos_lat_d_dlat', 'sec_lat_d_dlat_cos2', 'laplacian',
               'inverse_laplacian', 'clip_wavenumbers'):
    got = getattr(g, name)(xt).cpu().numpy()
    assert _rel(got, getattr(og, name)(x)) < 1e-14, name
  u, v = sh.get_cos_lat_vector(xt[0], xt[1], g, clip=False)
  ou, ov = osht.get_cos_lat_vector(x[0], x[1], og, clip=False)
  assert _rel(u.cpu().numpy(), ou) < 1e-14 and _rel(v.cpu().numpy(), ov) < 1e-14
  d = g.div_cos_lat((xt[0], xt[1])).cpu().numpy()
  assert _rel(d, og.div_cos_lat((x[0], x[1]))) < 1e-14


@pytest.mark.parametrize('name', ['T21', 'TL127', 'odd_j25', 'j36_l16', 'equiangular'])
def test_split_bf16x3_transforms_within_1e5(name, gpu_device):
  """Precision pattern: split-BF16 x3 MFMA Legendre with FP32 accumulation
  stays within 1e-5 normwise (max-abs and RMS) of the FP64 oracle."""
  import functools
  impl = functools.partial(sh.RealSphericalHarmonics, transform_precision='bf16x3')
  shapes = {'odd_j25': ((12, 13, 32, 25), {}), 'j36_l16': ((16, 16, 48, 36), {}),
            'equiangular': ((24, 24, 96, 48), {'latitude_spacing': 'equiangular'})}
  if name in shapes:
    args, kw = shapes[name]
    g = sh.Grid(*args, spherical_harmonics_impl=impl, **kw)
    og = osht.Grid(*args, **kw)
  else:
    g = getattr(sh.Grid, name)(spherical_harmonics_impl=impl)
    og = getattr(osht.Grid, name)()
  x = _modal(og, 6, 9)
  for i, s in enumerate([1.0, 1e-2, 300.0, 1e-5, 1.0, 1.0]):
    x[i] *= s
  y = g.to_nodal(torch.tensor(x, dtype=torch.float32, device=gpu_device)).cpu().numpy()
  yr = og.to_nodal(x)
  z = g.to_modal(torch.tensor(yr, dtype=torch.float32, device=gpu_device)).cpu().numpy()
  zr = og.to_modal(yr)
  for i in range(6):
    for got, ref, what in ((y[i], yr[i], 'to_nodal'), (z[i], zr[i], 'to_modal')):
      mx = _rel(got, ref)
      rms = np.sqrt(((got - ref) ** 2).mean() / (ref ** 2).mean())
      print(f'bf16x3 {name} field {i} {what}: max-rel {mx:.2e} rms-rel {rms:.2e}', flush=True)
      assert mx < 1e-5 and rms < 1e-5


def test_tl255_round_trip_config1_all_modes(gpu_device):
  """Isolated SHT round trip at TL255 (37 levels x 6 fields = 222 fields,
  i.i.d. N(0,1) (l+1)^-1.5 coefficients, m = 0 sin slot zero): to_nodal then
  to_modal in FP64, FP32 and split-BF16x3.  FP64 matches the oracle on a
  subset; FP32 and BF16x3 are compared with the FP64 round trip on all 222
  fields (max-abs and RMS ratios, reported)."""
  import functools
  print('[test] TL255 round trip, 222 fields', flush=True)
  og = osht.Grid.TL255()
  x = _modal(og, 222, 21)
  xt = torch.tensor(x, device=gpu_device)
  g64 = sh.Grid.TL255()
  y64 = g64.to_nodal(xt)
  z64 = g64.to_modal(y64)
  sub = [0, 7, 55, 110, 111, 150, 200, 221]
  y_ref = og.to_nodal(x[sub])
  assert _rel(y64[sub].cpu().numpy(), y_ref) < 1e-12
  assert _rel(z64[sub].cpu().numpy(), og.to_modal(y_ref)) < 1e-12
  ref = z64.cpu().numpy()
  for prec, tol in (('float32', 2e-6), ('bf16x3', 1e-5)):
    impl = functools.partial(sh.RealSphericalHarmonics, transform_precision=prec)
    g = sh.Grid.TL255(spherical_harmonics_impl=impl)
    z = g.to_modal(g.to_nodal(xt.float())).double().cpu().numpy()
    mx = max(_rel(z[i], ref[i]) for i in range(222))
    rms = max(np.sqrt(((z[i] - ref[i]) ** 2).mean() / (ref[i] ** 2).mean()) for i in range(222))
    print(f'TL255 round trip {prec}: worst max-rel {mx:.2e}, worst rms-rel {rms:.2e}', flush=True)
    assert mx < tol and rms < tol, (prec, mx, rms)


@pytest.mark.parametrize('fn', ['dino_sht_forward', 'dino_sht_inverse_adjoint'])
def test_triangle_only_flag_leaves_zero_block_untouched(fn, gpu_device):
  """DINO_SHT_TRIANGLE_ONLY: each modal row of wavenumber m is written from
  degree s(m) = max(0, L - 32 ceil((L - m) / 32)) on, bitwise equal to the full
  call there, and each degree below keeps what the buffer held (a NaN
  sentinel here) or holds its true value, zero (whole-line stores may zero at
  most 31 of them on each side of a row); most stay untouched.  TL127, FP32,
  6 fields."""
  import ctypes
  from dinosaur_amd import _lib
  print(f'[test] triangle-only flag {fn}', flush=True)
  g = sh.Grid.TL127()
  plan = g.spherical_harmonics.plan(gpu_device)
  rs = np.random.RandomState(9)
  z = torch.tensor(rs.normal(size=(6,) + g.nodal_shape), dtype=torch.float32, device=gpu_device)
  full = sh.sht_call(fn, plan, z, (plan.R, plan.L), _lib.PREC_NATIVE)
  out = torch.full((6, plan.R, plan.L), float('nan'), device=gpu_device)
  work = torch.empty((6, plan.R, plan.J), device=gpu_device)
  p = lambda t: ctypes.c_void_p(t.data_ptr())
  stream = ctypes.c_void_p(torch.cuda.current_stream(gpu_device).cuda_stream)
  rc = getattr(_lib.lib(), fn)(plan.handle, p(z), p(out), p(work), 6, _lib.DINO_F32,
                               _lib.PREC_NATIVE | _lib.SHT_TRIANGLE_ONLY, stream)
  _lib.check(rc, fn)
  torch.cuda.synchronize()
  kept = below = 0
  for r in range(plan.R):
    m = (r + 1) // 2
    lt0 = max(0, plan.L - 32 * (-(-(plan.L - m) // 32)))
    assert lt0 <= m
    low = out[:, r, :lt0]
    assert (torch.isnan(low) | (low == 0)).all(), r
    assert (torch.isnan(low).sum(dim=1) >= lt0 - 62).all(), r
    kept += int(torch.isnan(low).sum())
    below += low.numel()
    assert torch.equal(out[:, r, lt0:], full[:, r, lt0:]), r
  assert kept >= 0.5 * below, (kept, below)


@pytest.mark.parametrize('name', ['T21', 'TL127', 'TL255', 'j36_l16'])
def test_wide_forward_rows_bitwise_equal(name, gpu_device):
  """The forward row kernel with 128-row workgroups (8 waves) does the same
  arithmetic per row as the 64-row one: outputs are bitwise equal, with
  batches that leave the last workgroup partly empty (and half the rows at
  m = 0), full and triangle-only writes, FP32 and split-BF16x3, and the
  scalar-load path (J % 8 != 0, j36_l16)."""
  import ctypes
  from dinosaur_amd import _lib
  grids = {'T21': sh.Grid.T21(), 'TL127': sh.Grid.TL127(), 'TL255': sh.Grid.TL255(),
           'j36_l16': sh.Grid(16, 16, 48, 36)}
  plan = grids[name].spherical_harmonics.plan(gpu_device)
  lib = _lib.lib()
  stream = ctypes.c_void_p(torch.cuda.current_stream(gpu_device).cuda_stream)
  p = lambda t: ctypes.c_void_p(t.data_ptr())
  st = _lib.STAGE_LEGENDRE_FORWARD
  old = lib.dino_plan_get_schedule(plan.handle, st)
  batches = (3, 70) if name == 'TL255' else (1, 33, 70, 130)
  try:
    for b in batches:
      rs = np.random.RandomState(b)
      four = torch.tensor(rs.normal(size=(b, plan.R, plan.J)), dtype=torch.float32,
                          device=gpu_device)
      for prec in (_lib.PREC_NATIVE, _lib.PREC_BF16X3):
        for flag in (0, _lib.SHT_TRIANGLE_ONLY):
          outs = []
          for sched in (_lib.SCHED_ROWS, _lib.SCHED_ROWS_WIDE,
                        _lib.SCHED_ROWS_WIDE | _lib.SCHED_MPAIRED):
            _lib.check(lib.dino_plan_set_schedule(plan.handle, st, sched), 'schedule')
            out = torch.full((b, plan.R, plan.L), float('nan'), device=gpu_device)
            _lib.check(lib.dino_legendre_forward(plan.handle, p(four), p(out), b, _lib.DINO_F32,
                                                 prec | flag, stream), 'forward')
            torch.cuda.synchronize()
            outs.append(out.nan_to_num(nan=7.0))
          print(f'[test] wide rows {name} B={b} prec={prec} flag={flag}', flush=True)
          assert torch.equal(outs[0], outs[1]) and torch.equal(outs[0], outs[2])
          if not flag:
            assert not torch.isnan(outs[0]).any()
  finally:
    lib.dino_plan_set_schedule(plan.handle, st, old)
