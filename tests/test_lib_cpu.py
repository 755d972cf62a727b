"""CPU checks of the C-ABI library: it loads, exports every symbol the header
declares, and its host-side FP64 Gauss nodes match library routines."""
import os
import re

import numpy as np
import pytest
import scipy.special

from dinosaur_amd import _lib

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                      'include', 'dinosaur_amd.h')


def _declared():
  text = open(HEADER).read()
  return sorted(set(re.findall(r'\b(dino_[a-z0-9_]+)\s*\(', text)))


def test_library_exports_every_declared_symbol():
  lib = _lib.lib()
  names = _declared()
  assert len(names) >= 20
  for name in names:
    assert hasattr(lib, name), name
  assert set(names) == set(_lib.SIGNATURES)


@pytest.mark.parametrize('n', [1, 2, 3, 16, 33, 128, 256, 512])
def test_host_gauss_nodes(n):
  x, w = _lib.gauss_nodes_host(n)
  xs, ws = scipy.special.roots_legendre(n)
  np.testing.assert_allclose(x, xs, atol=2e-15)
  # scipy's smallest (polar) weights carry ~1e-9 relative error at n = 512;
  # absolute agreement plus exact integration of x^2k (k < n) pins the nodes.
  np.testing.assert_allclose(w, ws, rtol=0, atol=1e-13)
  for k in range(n):
    assert abs((w * x ** (2 * k)).sum() - 2 / (2 * k + 1)) < 5e-15
  np.testing.assert_array_equal(x, -x[::-1])


def test_error_mapping():
  with pytest.raises(ValueError):
    _lib.check(_lib.lib().dino_gauss_nodes_host(0, None, None))


def test_schedule_entry_points_reject_invalid_arguments():
  lib = _lib.lib()
  assert lib.dino_plan_set_schedule(None, _lib.STAGE_LEGENDRE_INVERSE, 0) == _lib.DINO_EINVAL
  assert b'NULL plan' in lib.dino_last_error()
  assert lib.dino_plan_get_schedule(None, 0) == -1


def test_autotune_batch_buckets_and_lookup():
  from dinosaur_amd import autotune
  assert [autotune.bucket(b) for b in (1, 2, 3, 64, 65, 5984, 10672)] == \
      [1, 2, 4, 64, 128, 8192, 16384]
  tuned = {(8192, 0, 0): {0: 1}, (16384, 0, 0): {0: 2}, (64, 0, 2): {0: 3}}
  assert autotune.lookup(tuned, 5000, 0, 0) == {0: 1}
  assert autotune.lookup(tuned, 12000, 0, 0) == {0: 2}
  assert autotune.lookup(tuned, 10 ** 6, 0, 0) == {0: 2}   # nearest bucket
  assert autotune.lookup(tuned, 100, 0, 2) == {0: 3}       # other precision
  assert autotune.lookup(tuned, 100, 1, 0) is None         # nothing tuned for FP64


@pytest.mark.parametrize('rows,degrees', [(43, 23), (511, 257), (85, 44), (1, 1)])
def test_triangle_entries_counts_mask(rows, degrees):
  """Entries with |m| <= l of a Real-layout (2M-1, L) array, m order
  0, +1, -1, +2, -2, ... (TL255, (511, 257): 66,047 entries)."""
  m = np.array([0] + [k for j in range(1, (rows + 1) // 2) for k in (j, j)])
  l = np.arange(degrees)
  assert _lib.triangle_entries(rows, degrees) == int((m[:, None] <= l[None, :]).sum())


def test_work_tally_roofs():
  """Time at the binding roof per family: bytes at 8 TB/s or flops at the
  unit's peak, whichever is longer; nothing is recorded outside the context."""
  _lib.tally('x', 1e9)
  with _lib.work_tally() as t:
    _lib.tally('leg', 8e9, 157.3e12 * 2e-3, 'f32_mfma')
    _lib.tally('leg', 8e9, 0.0, 'f32_mfma')
    _lib.tally('copy', 16e9)
    _lib.tally('solve', 0.8e9, 78.6e12 * 1e-4, 'f64')
  _lib.tally('x', 1e9)
  r = t.roof_seconds()
  assert set(r) == {'leg', 'copy', 'solve'}
  assert t.families['leg']['launches'] == 2
  assert abs(r['leg'] - 2e-3) < 1e-12          # 16 GB = 2 ms, 2 ms of MFMA
  assert abs(r['copy'] - 2e-3) < 1e-12
  assert abs(r['solve'] - 1e-4) < 1e-12        # 0.1 ms of memory, 0.1 ms of FP64


def test_release_defers_destroy_during_graph_capture(monkeypatch):
  """Library objects dropped while a stream is being captured (the garbage
  collector may run there) are destroyed after the capture, not inside it:
  their destroy calls free device memory, which would invalidate the
  capture.  A fake library records the calls."""
  from dinosaur_amd import _lib

  calls = []

  class Fake:
    def __getattr__(self, name):
      return lambda h: calls.append((name, h))

  monkeypatch.setattr(_lib, '_lib', Fake())
  monkeypatch.setattr(_lib, '_released', [])
  capturing = [True]
  monkeypatch.setattr(_lib, '_capturing', lambda: capturing[0])
  _lib.release('dino_plan_destroy', 1)
  _lib.release('dino_modal_destroy', 2)
  assert calls == [] and len(_lib._released) == 2
  capturing[0] = False
  _lib.release('dino_nodal_destroy', 3)
  assert sorted(calls[:2]) == [('dino_modal_destroy', 2), ('dino_plan_destroy', 1)]
  assert calls[2] == ('dino_nodal_destroy', 3) and _lib._released == []
  monkeypatch.setattr(_lib, '_lib', None)
  _lib.release('dino_plan_destroy', 4)  # no library loaded: nothing to do
  assert len(calls) == 3
