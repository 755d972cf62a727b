"""The allocation-free inverse transform (dino_sht_inverse_scratch) against
dino_sht_inverse with a caller workspace, called through the C-ABI.

Both entry points run the same Legendre and DFT kernels, so their outputs
must be bitwise equal.  Each case makes two calls in a row, the second with
a smaller batch, so the second reads a scratch array that still holds the
first call's rows.
"""
import ctypes

import numpy as np
import pytest
import torch

from dinosaur_amd import _lib, spherical_harmonic as sh

pytestmark = pytest.mark.gpu


def _p(t):
  return ctypes.c_void_p(t.data_ptr())


def _stream(dev):
  return ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)


def _random_modal(g, batch, dtype, dev, seed):
  rs = np.random.RandomState(seed)
  R, L = g.modal_shape
  x = rs.normal(size=(batch, R, L)) / (np.arange(L) + 1.0) ** 1.5
  x *= np.asarray(g.mask, dtype=np.float64)
  return torch.tensor(x, dtype=dtype, device=dev)


def _both(g, modal, precision, dev):
  """(workspace result, scratch result) of one inverse transform."""
  lib = _lib.lib()
  plan = g.spherical_harmonics.plan(dev)
  batch = modal.shape[0]
  code = _lib.DINO_F64 if modal.dtype == torch.float64 else _lib.DINO_F32
  nan = float('nan')
  ref = torch.full((batch,) + g.nodal_shape, nan, dtype=modal.dtype, device=dev)
  out = torch.full((batch,) + g.nodal_shape, nan, dtype=modal.dtype, device=dev)
  work = torch.empty((batch, plan.R, plan.J), dtype=modal.dtype, device=dev)
  with torch.cuda.device(dev):
    _lib.check(lib.dino_sht_inverse(plan.handle, _p(modal), _p(ref), _p(work), batch, code,
                                    precision, _stream(dev)), 'dino_sht_inverse')
    torch.cuda.synchronize(dev)
    before = torch.cuda.memory_allocated(dev)
    _lib.check(lib.dino_sht_inverse_scratch(plan.handle, _p(modal), _p(out), batch, code,
                                            precision, _stream(dev)),
               'dino_sht_inverse_scratch')
    torch.cuda.synchronize(dev)
    assert torch.cuda.memory_allocated(dev) == before
  return ref, out


@pytest.mark.parametrize('grid,dtype,precision,batches', [
    ('TL255', torch.float32, _lib.PREC_NATIVE, (120, 7)),
    ('TL255', torch.float32, _lib.PREC_BF16X3, (37, 5)),
    ('T21', torch.float64, _lib.PREC_NATIVE, (64, 3)),
    ('TL127', torch.float32, _lib.PREC_NATIVE, (222, 1)),
])
def test_scratch_inverse_matches_workspace_inverse(gpu_device, grid, dtype, precision, batches):
  g = getattr(sh.Grid, grid)()
  for i, batch in enumerate(batches):
    print(f'[test] scratch inverse {grid} {dtype} prec={precision} batch={batch}', flush=True)
    modal = _random_modal(g, batch, dtype, gpu_device, seed=100 + i)
    ref, out = _both(g, modal, precision, gpu_device)
    assert not torch.isnan(ref).any()
    assert torch.equal(out, ref), (grid, batch, (out - ref).abs().max().item())


def test_scratch_inverse_rejects_batches_beyond_the_array(gpu_device):
  g = sh.Grid.TL255()
  plan = g.spherical_harmonics.plan(gpu_device)
  per_field = 4 * plan.R * plan.J
  batch = (64 << 20) // per_field + 1
  print(f'[test] scratch inverse rejects batch={batch}', flush=True)
  modal = torch.zeros((1,) + g.modal_shape, dtype=torch.float32, device=gpu_device)
  nodal = torch.zeros((1,) + g.nodal_shape, dtype=torch.float32, device=gpu_device)
  # The size check runs on the host before any launch, so the undersized
  # buffers are never touched.
  rc = _lib.lib().dino_sht_inverse_scratch(plan.handle, _p(modal), _p(nodal), batch,
                                           _lib.DINO_F32, _lib.PREC_NATIVE, _stream(gpu_device))
  assert rc == _lib.DINO_EINVAL
  assert b'scratch array' in _lib.lib().dino_last_error()
