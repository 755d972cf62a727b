"""Vertical regridding on the GPU (csrc/regrid.hip) against the oracle: the
reference's known answers (vertical_interpolation_test.py:168-228), then
ERA5-like 137 analytic hybrid levels onto 37 sigma levels over a T42 grid of
surface pressures, conservative (sigma and pressure units) and linear, FP64
1e-12 and FP32 2e-6, with member-varying surface pressure."""
import numpy as np
import pytest
import torch

from oracle import hybrid as ohyb
from oracle import sigma as osig
from oracle import vertical_interpolation as ovi
from dinosaur_amd import hybrid_coordinates as hc
from dinosaur_amd import sigma_coordinates as sc
from dinosaur_amd import vertical_interpolation as vi

pytestmark = pytest.mark.gpu


def test_known_answers(gpu_device):
  print('[test] regrid known answers', flush=True)
  sig = sc.SigmaCoordinates.equidistant(5)
  hyb = hc.HybridCoordinates(np.array([0, 30, 75, 200, 300, 300, 150, 0]),
                             np.array([0, 0, 0, 0, 0.15, 0.4, 0.7, 1]))
  ps = torch.tensor([[1000.0]], dtype=torch.float64, device=gpu_device)
  x = torch.arange(1.0, 8.0, dtype=torch.float64, device=gpu_device)[:, None, None]
  got = vi.regrid_hybrid_to_sigma(x, hyb, sig, ps).cpu().numpy().ravel()
  want = [1 * 30 / 200 + 2 * 45 / 200 + 3 * 125 / 200, 4.0, 4 * 50 / 200 + 5 * 150 / 200,
          5 * 100 / 200 + 6 * 100 / 200, 6 * 50 / 200 + 7 * 150 / 200]
  np.testing.assert_allclose(got, want, atol=1e-12)
  src = hc.HybridCoordinates(np.array([0, 400, 300, 0]), np.array([0, 0, 0.5, 1.0]))
  tgt = hc.HybridCoordinates(np.array([0, 250, 0]), np.array([0, 0.25, 1.0]))
  x3 = torch.tensor([1.0, 2.0, 3.0], dtype=torch.float64, device=gpu_device)[:, None, None]
  got = vi.regrid_hybrid_to_hybrid(x3, src, tgt, ps).cpu().numpy().ravel()
  np.testing.assert_allclose(got, [1.2, 2.4], atol=1e-12)
  got = vi.interp_hybrid_to_sigma(x3, src, sig, ps).cpu().numpy().ravel()
  np.testing.assert_allclose(got, [0.75, 1.25, 1.75, 2 + 1 / 3, 3.0], atol=1e-12)


@pytest.mark.parametrize('dtype,tol', [(torch.float64, 1e-12), (torch.float32, 2e-6)])
def test_era5_like_levels_match_oracle(dtype, tol, gpu_device):
  print(f'[test] regrid 137 hybrid -> 37 sigma levels vs oracle, {dtype}', flush=True)
  rs = np.random.RandomState(4)
  ohyb137 = ohyb.HybridCoordinates(*_ecmwf_like(137))
  hyb137 = hc.HybridCoordinates(*_ecmwf_like(137))
  osig37 = osig.SigmaCoordinates.equidistant(37)
  sig37 = sc.SigmaCoordinates.equidistant(37)
  X, Y, E = 12, 10, 2
  ps = 700 + 330 * rs.uniform(size=(E, 1, X, Y))
  f = rs.normal(size=(E, 3, 137, X, Y)).cumsum(axis=2) / 10
  T = lambda a: torch.as_tensor(a, device=gpu_device).to(dtype)
  for fn, ofn, tgt, otgt in ((vi.regrid_hybrid_to_sigma, ovi.regrid_hybrid_to_sigma, sig37,
                              osig37),
                             (vi.interp_hybrid_to_sigma, ovi.interp_hybrid_to_sigma, sig37,
                              osig37)):
    got = fn(T(f), hyb137, tgt, T(ps)).cpu().numpy()
    ref = np.stack([ofn(f[e], ohyb137, otgt, ps[e, 0]) for e in range(E)])
    finite = np.isfinite(ref)
    assert np.array_equal(finite, np.isfinite(got)), fn.__name__
    err = np.abs(got[finite] - ref[finite]).max() / np.abs(ref[finite]).max()
    print(f'  {fn.__name__}: rel err {err:.2e}', flush=True)
    assert err < tol, (fn.__name__, err)
  o40 = ohyb.HybridCoordinates.analytic_levels(40)
  p40 = hc.HybridCoordinates.analytic_levels(40)
  got = vi.regrid_hybrid_to_hybrid(T(f), hyb137, p40, T(ps)).cpu().numpy()
  ref = np.stack([ovi.regrid_hybrid_to_hybrid(f[e], ohyb137, o40, ps[e, 0]) for e in range(E)])
  err = np.abs(got - ref).max() / np.abs(ref).max()
  print(f'  regrid_hybrid_to_hybrid: rel err {err:.2e}', flush=True)
  assert err < tol


def _ecmwf_like(n):
  lev = hc.HybridCoordinates.analytic_ecmwf_like(n)
  return lev.a_boundaries, lev.b_boundaries
