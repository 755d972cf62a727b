"""Mixed-truncation ensembles on the GPU: graph-captured ticks (concurrent
streams or one stream) equal eager ticks bitwise, and a member advanced inside
its batched group equals the same member advanced alone."""
import time

import numpy as np
import pytest
import torch

from dinosaur_amd import ensemble, time_integration as ti, units as units_lib, workloads

pytestmark = pytest.mark.gpu


def _group(name, layers, members, first, dt_seconds, every, gpu_device):
  eq, st, coords, ps = workloads.build_primitive(name, layers, moist=True, members=members,
                                                 device=gpu_device, first_member=first)
  dt = ps.nondimensionalize(dt_seconds * units_lib.scales.units.s)
  step = ti.step_with_filters(ti.imex_rk_sil3(eq, dt),
                              [ti.exponential_step_filter(coords.horizontal, dt)])
  flat = st.as_flat()
  flat = type(flat).from_flat(flat.flat_buffer(), *flat._flat[1:], sim_time=0.0)
  return ensemble.Group(name, step, flat, every)


def _buffers(ens):
  return [s.flat_buffer().clone() for s in ens.states()]


def test_mixed_ensemble_graph_ticks_match_eager(gpu_device):
  """T21 (3 members, 20 min, every tick) + T42 (2 members, 40 min, every other
  tick), moist L6: 4 ticks eager, graphed on concurrent streams and graphed on
  one stream give bitwise equal states and equal simulated times."""
  print('[test] mixed ensemble T21 + T42: graph vs eager', flush=True)

  def groups():
    return [_group('T21', 6, 3, 0, 1200.0, 1, gpu_device),
            _group('T42', 6, 2, 3, 2400.0, 2, gpu_device)]

  runs = [ensemble.MixedTruncationEnsemble(groups(), graph=False),
          ensemble.MixedTruncationEnsemble(groups(), graph=True, concurrent=True),
          ensemble.MixedTruncationEnsemble(groups(), graph=True, concurrent=False)]
  for tick in range(4):
    for e in runs:
      e.tick()
    ref = _buffers(runs[0])
    for e in runs[1:]:
      for a, b in zip(_buffers(e), ref):
        assert torch.equal(a, b), tick
  times = [[s.sim_time for s in e.states()] for e in runs]
  assert times[0] == times[1] == times[2]
  assert times[0][0] > 0 and np.isclose(times[0][0], times[0][1])  # both 80 min


def test_member_in_group_matches_member_alone(gpu_device):
  """Member 1 of a 3-member T42 group against the same seeded member stepped
  alone (FP32, 3 graphed ticks)."""
  print('[test] batched member vs member alone (T42 L6 moist)', flush=True)
  grp = ensemble.MixedTruncationEnsemble([_group('T42', 6, 3, 0, 1200.0, 1, gpu_device)])
  one = ensemble.MixedTruncationEnsemble([_group('T42', 6, 1, 1, 1200.0, 1, gpu_device)])
  grp.run(3)
  one.run(3)
  a, b = grp.states()[0], one.states()[0]
  for f in ('vorticity', 'divergence', 'temperature_variation', 'log_surface_pressure'):
    x, y = getattr(a, f)[1], getattr(b, f)
    err = ((x - y).abs().max() / y.abs().max()).item()
    print(f'  {f}: bitwise {torch.equal(x, y)} max-abs ratio {err:.2e}', flush=True)
    assert err <= 1e-6, (f, err)
  for k in b.tracers:
    x, y = a.tracers[k][1], b.tracers[k]
    scale = y.abs().max().item()
    if scale:
      assert ((x - y).abs().max() / scale).item() <= 1e-6, k


def test_config3_mixed_ensemble_graph_ticks(gpu_device):
  """The 64-member mixed ensemble (32 T85 + 24 T42 + 8 TL127, moist L24, 10 min
  steps, T42 at 20 min every other tick): 4 graphed ticks on concurrent
  streams equal 4 eager ticks bitwise, and the states stay finite."""
  print('[test] config-3 ensemble: 32 T85 + 24 T42 + 8 TL127, L24 moist', flush=True)

  def groups():
    return [_group('T85', 24, 32, 0, 600.0, 1, gpu_device),
            _group('T42', 24, 24, 32, 1200.0, 2, gpu_device),
            _group('TL127', 24, 8, 56, 600.0, 1, gpu_device)]

  eager = ensemble.MixedTruncationEnsemble(groups(), graph=False)
  graphed = ensemble.MixedTruncationEnsemble(groups(), graph=True, concurrent=True)
  eager.run(4)
  torch.cuda.synchronize()
  t0 = time.perf_counter()
  graphed.run(4)
  torch.cuda.synchronize()
  print(f'  4 graphed ticks {1e3 * (time.perf_counter() - t0):.1f} ms', flush=True)
  for a, b in zip(_buffers(graphed), _buffers(eager)):
    assert torch.isfinite(b).all()
    assert torch.equal(a, b)
  days = [s.sim_time for s in graphed.states()]
  assert np.allclose(days, days[0]), days
