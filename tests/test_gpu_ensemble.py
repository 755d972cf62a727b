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


def test_ensemble_members_match_oracle(gpu_device):
  """One member of each group of a graph-captured mixed ensemble (T21, 3
  members, 20 min every tick + T42, 2 members, 40 min every other tick;
  moist L6 with 2 cloud tracers; FP64) against the oracle stepping that
  member alone from the same initial state with the same SIL3 + filter: 4
  ticks = 4 T21 steps and 2 T42 steps, every field within 1e-10."""
  from oracle import pe as ope, sht as osht, sigma as osig, time_integration as oti
  from oracle import units as ounits
  print('[test] mixed ensemble members vs oracle (FP64)', flush=True)
  clouds = ('specific_cloud_liquid_water_content', 'specific_cloud_ice_water_content')
  specs = (('T21', 3, 0, 1200.0, 1, 4), ('T42', 2, 3, 2400.0, 2, 2))
  groups, refs = [], []
  for name, members, first, dt_s, every, steps in specs:
    eq, st, coords, ps = workloads.build_primitive(name, 6, moist=True, members=members,
                                                   device=gpu_device, first_member=first,
                                                   dtype=torch.float64)
    dt = ps.nondimensionalize(dt_s * units_lib.scales.units.s)
    step = ti.step_with_filters(ti.imex_rk_sil3(eq, dt),
                                [ti.exponential_step_filter(coords.horizontal, dt)])
    flat = st.as_flat()
    groups.append(ensemble.Group(name, step, flat, every))
    og = getattr(osht.Grid, name)()
    oeq = ope.PrimitiveEquationsSigma(
        np.asarray(eq.reference_temperature, np.float64), eq.orography.double().cpu().numpy(),
        ope.Coords(og, osig.SigmaCoordinates.equidistant(6)), ounits.SimUnits.from_si(),
        humidity_key='specific_humidity', cloud_keys=clouds)
    ostep = oti.step_with_filters(oti.imex_rk_sil3(oeq, dt),
                                  [oti.exponential_step_filter(og, dt)])
    m = 1
    y = ope.State(*(getattr(st, f)[m].cpu().numpy() for f in
                    ('vorticity', 'divergence', 'temperature_variation',
                     'log_surface_pressure')),
                  tracers={k: v[m].cpu().numpy() for k, v in st.tracers.items()})
    for _ in range(steps):
      y = ostep(y)
    refs.append(y)
  ens = ensemble.MixedTruncationEnsemble(groups, graph=True, concurrent=True)
  ens.run(4)
  for (name, *_), got, ref in zip(specs, ens.states(), refs):
    for f in ('vorticity', 'divergence', 'temperature_variation', 'log_surface_pressure'):
      a, b = getattr(got, f)[1].cpu().numpy(), getattr(ref, f)
      err = np.abs(a - b).max() / np.abs(b).max()
      print(f'  {name} member 1 {f}: {err:.2e}', flush=True)
      assert err < 1e-10, (name, f, err)
    for k, b in ref.tracers.items():
      a = got.tracers[k][1].cpu().numpy()
      scale = np.abs(b).max()
      err = np.abs(a - b).max() / scale if scale else np.abs(a).max()
      print(f'  {name} member 1 {k}: {err:.2e}', flush=True)
      assert err < 1e-10, (name, k, err)
