"""Pins the oracle's vertical regridding (oracle/vertical_interpolation.py):
the reference's known answers (vertical_interpolation_test.py:168-238), the
row sums of the conservative weights, conservation of the column integral
over the common range, identity regridding onto the same levels, and planted
mistakes."""
import numpy as np
import pytest

from oracle import hybrid as ohyb
from oracle import sigma as osig
from oracle import vertical_interpolation as vi


def test_known_answers():
  # interval overlap (:230-238)
  np.testing.assert_array_equal(
      vi.interval_overlap(np.array([1, 4, 6, 9, 10]), np.array([0, 3, 7, 10])),
      np.array([[2, 0, 0, 0], [1, 2, 1, 0], [0, 0, 2, 1]]))
  # conservative hybrid -> sigma (:185-205)
  sig = osig.SigmaCoordinates.equidistant(5)
  hyb = ohyb.HybridCoordinates(np.array([0, 30, 75, 200, 300, 300, 150, 0]),
                               np.array([0, 0, 0, 0, 0.15, 0.4, 0.7, 1]))
  got = vi.regrid_hybrid_to_sigma(np.arange(1.0, 8.0)[:, None, None], hyb, sig,
                                  np.array([[1000]])).ravel()
  want = [1 * 30 / 200 + 2 * 45 / 200 + 3 * 125 / 200, 4.0, 4 * 50 / 200 + 5 * 150 / 200,
          5 * 100 / 200 + 6 * 100 / 200, 6 * 50 / 200 + 7 * 150 / 200]
  np.testing.assert_allclose(got, want, atol=1e-6)
  # conservative hybrid -> hybrid (:207-228)
  src = ohyb.HybridCoordinates(np.array([0, 400, 300, 0]), np.array([0, 0, 0.5, 1.0]))
  tgt = ohyb.HybridCoordinates(np.array([0, 250, 0]), np.array([0, 0.25, 1.0]))
  got = vi.regrid_hybrid_to_hybrid(np.array([1.0, 2.0, 3.0])[:, None, None], src, tgt,
                                   np.array([[1000]]))
  np.testing.assert_allclose(got.ravel(), [1.2, 2.4], atol=1e-6)
  # linear hybrid -> sigma with one extrapolated cell (:168-183)
  got = vi.interp_hybrid_to_sigma(np.array([1.0, 2.0, 3.0])[:, None, None], src, sig,
                                  np.array([[1000]])).ravel()
  np.testing.assert_allclose(got, [0.75, 1.25, 1.75, 2 + 1 / 3, 3.0], atol=1e-6)


def test_weights_rows_sum_to_one_and_conserve():
  rs = np.random.RandomState(0)
  hyb = ohyb.HybridCoordinates.analytic_levels(20, p_top=0.0)
  sig = osig.SigmaCoordinates.equidistant(7)
  ps = 800 + 300 * rs.uniform(size=(3, 4))
  f = rs.normal(size=(2, 20, 3, 4))
  out = vi.regrid_hybrid_to_sigma(f, hyb, sig, ps)
  for x in range(3):
    for y in range(4):
      s = hyb.a_boundaries / ps[x, y] + hyb.b_boundaries
      w = vi.conservative_regrid_weights(s, sig.boundaries)
      np.testing.assert_allclose(w.sum(axis=1), 1.0, rtol=1e-12)
      # both level sets span [0, 1] in sigma: the column integral is kept
      np.testing.assert_allclose((out[:, :, x, y] * np.diff(sig.boundaries)).sum(-1),
                                 (f[:, :, x, y] * np.diff(s)).sum(-1), rtol=1e-12)
  # onto the same levels: identity
  same = ohyb.HybridCoordinates.from_sigma_levels(sig)
  g = rs.normal(size=(7, 3, 4))
  np.testing.assert_allclose(vi.regrid_hybrid_to_sigma(g, same, sig, ps), g, rtol=1e-12)


def test_planted_mistake_fails(monkeypatch):
  sig = osig.SigmaCoordinates.equidistant(5)
  hyb = ohyb.HybridCoordinates(np.array([0, 30, 75, 200, 300, 300, 150, 0]),
                               np.array([0, 0, 0, 0, 0.15, 0.4, 0.7, 1]))
  orig = vi.interval_overlap
  monkeypatch.setattr(vi, 'interval_overlap',
                      lambda s, t: orig(np.asarray(s)[::-1][::-1] * 1.0 + 0.01, t))
  got = vi.regrid_hybrid_to_sigma(np.arange(1.0, 8.0)[:, None, None], hyb, sig,
                                  np.array([[1000]])).ravel()
  assert abs(got[0] - (30 / 200 + 2 * 45 / 200 + 3 * 125 / 200)) > 1e-3
  with pytest.raises(ValueError, match='layers'):
    vi.regrid_hybrid_to_sigma(np.zeros((3, 1, 1)), hyb, sig, np.array([[1000]]))
