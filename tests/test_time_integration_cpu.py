"""Host-side checks of the generic time-integration utilities (they run on
any torch tensors; the reference's own known answers,
time_integration_test.py:303-335)."""
import numpy as np
import pytest
import torch

from dinosaur_amd import time_integration as ti


def test_accumulate_repeated():
  result = ti.accumulate_repeated(lambda x: 2 * x, torch.arange(4.0), torch.tensor(1.0))
  assert result.item() == 0 * 2 + 1 * 4 + 2 * 8 + 3 * 16


def test_dfi_lanczos_weights():
  weights = ti._dfi_lanczos_weights(10, 10, 0.1)
  assert weights.size == 50
  assert (np.diff(weights) < 0).all()
  assert abs(weights[0] - 1.0) < 5e-3
  assert abs(weights[-1]) < 5e-3


def test_digital_filter_initialization_removes_oscillation():
  """Linear growth in x[0] plus a harmonic oscillator in x[1], x[2]: the
  filter keeps the slow mode and removes the oscillation (atol 1e-3)."""
  eq = ti.ImplicitExplicitODE.from_functions(
      lambda x: torch.stack([torch.tensor(0.1, dtype=x.dtype), x[2], -x[1]]),
      torch.zeros_like,
      lambda x, eta: x)
  dfi = ti.digital_filter_initialization(eq, ti.imex_rk_sil3, [], time_span=20.0,
                                         cutoff_period=20.0, dt=0.01)
  result = dfi(torch.tensor([1.0, 1.0, 1.0], dtype=torch.float64))
  np.testing.assert_allclose(result.numpy(), [1.0, 0.0, 0.0], atol=1e-3)


def test_time_reversed_ode_undoes_forward_steps():
  """One SIL3 step of the reversed harmonic oscillator after one forward step
  returns to the start to the scheme's local error."""
  eq = ti.ImplicitExplicitODE.from_functions(lambda x: torch.stack([x[1], -x[0]]),
                                             torch.zeros_like, lambda x, eta: x)
  x0 = torch.tensor([1.0, 0.0], dtype=torch.float64)
  fwd = ti.imex_rk_sil3(eq, 0.01)
  bwd = ti.imex_rk_sil3(ti.TimeReversedImExODE(eq), 0.01)
  x1 = bwd(fwd(x0))
  assert (x1 - x0).abs().max().item() < 1e-8


def test_maybe_fix_sim_time_roundoff():
  class S:
    sim_time = 0.30000000000000004
  s = ti.maybe_fix_sim_time_roundoff(S(), 0.1)
  assert abs(s.sim_time - 0.3) < 1e-15


def test_checkpointed_rollout_keep_graphs_policy():
  """Segments that record their graphs in the forward pass (keep_graphs) give
  bitwise the same state and gradient as recomputed segments and as plain
  autograd; the report counts them; 'auto' keeps none off the GPU; invalid
  policies are rejected."""
  torch.manual_seed(0)
  w = torch.randn(8, 8, dtype=torch.float64) * 0.3
  step = lambda y: torch.tanh(y @ w) + 0.1 * y  # noqa: E731
  x0 = torch.randn(4, 8, dtype=torch.float64)
  xa = x0.clone().requires_grad_(True)
  y_ref = xa
  for _ in range(6):
    y_ref = step(y_ref)
  g_ref, = torch.autograd.grad((y_ref ** 2).sum(), xa)
  for keep, kept in ((0, 0), (2, 2), (5, 3), ('auto', 0)):
    x = x0.clone().requires_grad_(True)
    rep = {}
    y = ti.checkpointed_rollout(step, x, 6, checkpoint_every=2, keep_graphs=keep, report=rep)
    g, = torch.autograd.grad((y ** 2).sum(), x)
    assert rep == {'segments': 3, 'kept': kept}, (keep, rep)
    assert torch.equal(y.detach(), y_ref.detach()) and torch.equal(g, g_ref), keep
  for bad in (-1, 'all', 1.5):
    with pytest.raises(ValueError):
      ti.checkpointed_rollout(step, x0.clone().requires_grad_(True), 6, 2, keep_graphs=bad)


def test_recompute_stage_plan():
  """What a recomputed IMEX step computes when its output value is unread:
  SIL3 records f0, f1, f2 (f2 graph-only: only the output combines it), keeps
  g0 (stages 1 and 2 read it) and skips g2, g3 and the last stage; with the
  output read every used tendency is computed.  A two-stage tableau whose
  second stage only feeds the output keeps just stage 0."""
  need_f, need_g, graph_only = ti._stage_plan(ti.SIL3_TABLEAU, need_output=True)
  assert need_f == [True, True, True, False]
  assert need_g == [True, False, True, True]
  assert graph_only == [False] * 4
  need_f, need_g, graph_only = ti._stage_plan(ti.SIL3_TABLEAU, need_output=False)
  assert need_f == [True, True, True, False]
  assert need_g == [True, False, False, False]
  assert graph_only == [False, False, True, False]
  # forward-backward Euler: y1 = y0 + dt f0 + dt g1 (solve), out = y0 + dt f0 + dt g1
  tab = ti.ImExButcherTableau(a_ex=[[1.0]], a_im=[[0.0, 1.0]], b_ex=[1.0, 0.0],
                              b_im=[0.0, 1.0])
  assert ti._stage_plan(tab, need_output=True) == ([True, False], [False, True],
                                                   [False, False])
  assert ti._stage_plan(tab, need_output=False) == ([True, False], [False, False],
                                                    [True, False])


class _FakeMemory:
  """A device of `capacity` bytes: kept segments leave `held` bytes allocated
  each and peak `transient` above that while they run.  `frag` bytes sit in
  the allocator's cache as free space inside partly used segments (never
  reusable for new blocks), `whole` bytes as cached segments holding no live
  block, which new allocations reuse before asking the driver."""

  def __init__(self, capacity, held, transient, frag=0, whole=0):
    self.capacity, self.held, self.transient = capacity, held, transient
    self.frag, self.whole = frag, whole
    self.alloc = 0
    self._peak = 0
    self.high_water = 0

  def _reserved(self, alloc):
    return self.frag + max(alloc, self.whole)

  def free(self):
    return self.capacity - self._reserved(self.alloc)

  def reusable(self):
    return max(0, self.whole - self.alloc)

  def allocated(self):
    return self.alloc

  def reset_peak(self):
    self._peak = self.alloc

  def peak(self):
    return self._peak

  def run_segment(self):
    top = self.alloc + self.held + self.transient
    self._peak = max(self._peak, top)
    self.high_water = max(self.high_water, self._reserved(top))
    self.alloc += self.held


GiB = 1 << 30


def _simulate(mem, segments, input_bytes=1 * GiB):
  leaves = [torch.empty(input_bytes // 8, dtype=torch.float64, device='meta')]
  budget = ti._GraphBudget('auto', leaves, 1, memory=mem)
  for _ in range(segments):
    if budget.keep_next():
      budget.run_kept(lambda *ls: mem.run_segment(), leaves)
  return budget


def test_graph_budget_stops_before_the_device_is_full():
  """The TL255 x 16 rollout's numbers (16 GiB graph, 6 GiB of temporaries per
  step, 288 GiB): 'auto' keeps segments while a segment and the backward
  pass's recomputation still fit, never beyond the device."""
  mem = _FakeMemory(288 * GiB, 16 * GiB, 6 * GiB)
  budget = _simulate(mem, 40)
  need = 16 * GiB + 2 * 22 * GiB + 4 * GiB
  assert budget.kept == (288 * GiB - need) // (16 * GiB) + 1
  assert mem.high_water <= mem.capacity
  # what the backward pass needs on top of the kept graphs: one recomputed
  # segment's peak and the cotangents' temporaries
  assert mem.alloc + 2 * 22 * GiB <= mem.capacity


def test_graph_budget_ignores_fragmented_cache():
  """Free space inside partly used cached segments is not available memory:
  with 39 GiB of it (the state the allocator was in when a policy that
  counted it ran out of memory) the budget keeps fewer segments, none
  beyond what the rest of the device admits, and the device never
  overflows."""
  clean = _simulate(_FakeMemory(288 * GiB, 16 * GiB, 6 * GiB), 40)
  mem = _FakeMemory(288 * GiB, 16 * GiB, 6 * GiB, frag=39 * GiB)
  frag = _simulate(mem, 40)
  assert frag.kept < clean.kept
  assert mem.high_water <= mem.capacity
  assert mem.frag + mem.alloc + 2 * 22 * GiB <= mem.capacity


def test_graph_budget_counts_wholly_free_cached_segments():
  """Cached segments with no live block (left by an earlier rollout of the
  same shapes) are reused as they stand: the budget keeps as many segments
  as on a clean device, without emptying the cache, and the device never
  overflows."""
  clean = _simulate(_FakeMemory(288 * GiB, 16 * GiB, 6 * GiB), 40)
  mem = _FakeMemory(288 * GiB, 16 * GiB, 6 * GiB, whole=150 * GiB)
  warm = _simulate(mem, 40)
  assert warm.kept == clean.kept
  assert mem.high_water <= mem.capacity
  assert not hasattr(ti._CudaMemory, 'release_cache')


def test_graph_budget_sizes_segments_by_their_peak():
  """A segment that keeps little but peaks high (temporaries much larger than
  its graph) is budgeted by its peak, not by what it leaves allocated."""
  mem = _FakeMemory(100 * GiB, 1 * GiB, 30 * GiB)
  budget = _simulate(mem, 40)
  # keep while free >= 1 + 2 * 31 + 4 GiB = 67 GiB: free 100, 99, ..., 67
  assert budget.kept == 34
  assert budget.peak == 31 * GiB and budget.held == 1 * GiB
  assert mem.high_water <= mem.capacity


def test_graph_budget_first_estimate_and_int_policy():
  """Before a measurement the estimate is 12x (held) and 16x (peak) the
  input bytes per step; an int policy ignores memory; off the GPU 'auto'
  keeps nothing."""
  leaves = [torch.empty(GiB // 8, dtype=torch.float64, device='meta')]
  tight = _FakeMemory(12 * GiB + 32 * GiB + 4 * GiB - 1, 0, 0)
  assert not ti._GraphBudget('auto', leaves, 1, memory=tight).keep_next()
  ok = _FakeMemory(12 * GiB + 32 * GiB + 4 * GiB, 0, 0)
  assert ti._GraphBudget('auto', leaves, 1, memory=ok).keep_next()
  b = ti._GraphBudget(2, leaves, 1, memory=tight)
  assert b.keep_next()
  b.run_kept(lambda *ls: None, leaves)
  b.run_kept(lambda *ls: None, leaves)
  assert not b.keep_next()
  assert not ti._GraphBudget('auto', [torch.zeros(4)], 1).keep_next()


def test_cuda_memory_queries(monkeypatch):
  """Free memory is what the driver reports; reusable cache is torch's
  reserved memory minus what is allocated minus the free space inside partly
  used segments (inactive splits)."""
  monkeypatch.setattr(torch.cuda, 'mem_get_info', lambda dev=None: (5 * GiB, 288 * GiB))
  monkeypatch.setattr(torch.cuda, 'memory_stats', lambda dev=None: {
      'reserved_bytes.all.current': 60 * GiB, 'allocated_bytes.all.current': 20 * GiB,
      'inactive_split_bytes.all.current': 15 * GiB})
  mem = ti._CudaMemory(torch.device('cuda', 0))
  assert mem.free() == 5 * GiB
  assert mem.reusable() == 25 * GiB


def test_checkpointed_rollout_auto_policy_with_device_memory(monkeypatch):
  """'auto' through checkpointed_rollout with the device memory stubbed: the
  segments the budget admits keep their graphs, the rest are recomputed, and
  the gradient is bitwise that of plain autograd."""
  torch.manual_seed(0)
  w = torch.randn(8, 8, dtype=torch.float64) * 0.3
  mem = _FakeMemory(100 * GiB, 10 * GiB, 5 * GiB)

  def step(y):
    mem.run_segment()
    return torch.tanh(y @ w) + 0.1 * y

  monkeypatch.setattr(ti, '_memory_for', lambda leaves: mem)
  x0 = torch.randn(4, 8, dtype=torch.float64)
  xa = x0.clone().requires_grad_(True)
  y_ref = xa
  for _ in range(8):
    y_ref = torch.tanh(y_ref @ w) + 0.1 * y_ref
  g_ref, = torch.autograd.grad((y_ref ** 2).sum(), xa)
  x = x0.clone().requires_grad_(True)
  rep = {}
  y = ti.checkpointed_rollout(step, x, 8, checkpoint_every=1, keep_graphs='auto', report=rep)
  g, = torch.autograd.grad((y ** 2).sum(), x)
  # first segment on the input-size estimate (tiny), then measured:
  # keep while free >= 10 + 2 * 15 + 4 = 44 GiB: free 100, 90, ..., 50 -> 6 kept
  assert rep == {'segments': 8, 'kept': 6}
  assert torch.equal(y.detach(), y_ref.detach()) and torch.equal(g, g_ref)


# ---------------------------------------------------------------- scan drivers


def _plain_scan(f, init, xs, length):
  carry, ys = init, []
  for i in range(length):
    carry, y = f(carry, None if xs is None else xs[i])
    ys.append(y)
  return carry, ys


@pytest.mark.parametrize('f,init,xs,lengths', [
    (lambda x, _: (x + 1, x ** 2), 1, None, [8]),
    (lambda x, _: (x + 1, x ** 2), 1, None, [2, 4]),
    (lambda x, _: (x + 1, x ** 2), 1, None, [1, 2, 2, 2]),
    (lambda x, y: (2 * x, x + y), 1, np.arange(8), [8]),
    (lambda x, y: (2 * x, x + y), 1, np.arange(8), [2, 4]),
    (lambda x, y: (2 * x, x + y), 1, np.arange(8), [2, 2, 2, 1]),
])
def test_nested_checkpoint_scan_matches_scan(f, init, xs, lengths):
  """time_integration_test.py:278-300: same carry and stacked outputs as an
  unnested scan."""
  length = int(np.prod(lengths))
  exp_c, exp_ys = _plain_scan(f, init, xs, length)
  c, ys = ti.nested_checkpoint_scan(f, init, xs, length, nested_lengths=lengths)
  np.testing.assert_allclose(np.asarray(c), np.asarray(exp_c))
  np.testing.assert_allclose(ys.numpy(), np.asarray([np.asarray(y) for y in exp_ys]))
  with pytest.raises(ValueError):
    ti.nested_checkpoint_scan(f, init, xs, length + 1, nested_lengths=lengths)


def test_nested_checkpoint_scan_gradients_and_recomputation():
  """Gradients of a loss on the final carry and on every per-step output,
  with respect to the initial carry, the scanned inputs and a parameter,
  equal those of plain autograd through an unnested loop; the backward pass
  recomputes the checkpointed inner scans (f runs again)."""
  torch.manual_seed(1)
  w = (torch.randn(8, 8, dtype=torch.float64) * 0.3).requires_grad_(True)
  calls = [0]

  def f(c, x):
    calls[0] += 1
    c = torch.tanh(c @ w) + 0.1 * c + x
    return c, (c ** 2).sum(-1)

  c0 = torch.randn(4, 8, dtype=torch.float64)
  xs = torch.randn(12, 4, 8, dtype=torch.float64) * 0.1
  ywt = torch.randn(12, 4, dtype=torch.float64)

  def loss_and_grads(runner):
    ci = c0.clone().requires_grad_(True)
    xi = xs.clone().requires_grad_(True)
    c, ys = runner(ci, xi)
    loss = (c ** 2).sum() + (ys * ywt).sum()
    return (c.detach(), ys.detach()) + torch.autograd.grad(loss, (ci, xi, w))

  def plain(ci, xi):
    c, ys = _plain_scan(f, ci, xi, 12)
    return c, torch.stack(ys)

  ref = loss_and_grads(plain)
  for lengths in ([3, 4], [2, 3, 2], [12]):
    calls[0] = 0
    got = loss_and_grads(lambda ci, xi: ti.nested_checkpoint_scan(
        f, ci, xi, 12, nested_lengths=lengths))
    for a, b in zip(got, ref):
      torch.testing.assert_close(a, b, rtol=0, atol=1e-14)
    # each level of checkpointing re-runs (at most) the forward pass once
    # more; torch's recomputation stops early once it has what it needs
    if len(lengths) == 1:
      assert calls[0] == 12
    else:
      assert 12 < calls[0] <= 12 * len(lengths), (lengths, calls[0])


def test_nested_checkpoint_scan_custom_scan_and_checkpoint_functions():
  """scan_fn and checkpoint_fn are used as given (not ignored)."""
  seen = {'scan': 0, 'checkpoint': 0}

  def my_scan(f, init, xs, length):
    seen['scan'] += 1
    return ti.scan(f, init, xs, length)

  def my_checkpoint(fn):
    seen['checkpoint'] += 1
    return fn

  c, ys = ti.nested_checkpoint_scan(lambda x, _: (x + 1, x), 0, None, 6, nested_lengths=[2, 3],
                                    scan_fn=my_scan, checkpoint_fn=my_checkpoint)
  assert c == 6 and ys.tolist() == [0, 1, 2, 3, 4, 5]
  assert seen == {'scan': 3, 'checkpoint': 1}


def test_repeated_and_trajectory_from_step():
  """time_integration.py:508-564: repeated applies fn `steps` times (the
  function itself for one step); trajectories have outer_steps frames,
  inner_steps apart, after each outer step or before it with
  start_with_input, through post_process_fn, stacked along a leading axis."""
  inc = lambda x: x + 1  # noqa: E731
  assert ti.repeated(inc, 1) is inc
  assert ti.repeated(inc, 5)(torch.tensor(0.0)).item() == 5.0
  final, frames = ti.trajectory_from_step(inc, 3, 2)(torch.tensor(0.0))
  assert final.item() == 6.0 and frames.tolist() == [2.0, 4.0, 6.0]
  final, frames = ti.trajectory_from_step(inc, 3, 2, start_with_input=True)(torch.tensor(0.0))
  assert final.item() == 6.0 and frames.tolist() == [0.0, 2.0, 4.0]
  _, frames = ti.trajectory_from_step(inc, 4, 1, post_process_fn=lambda x: {'y': 10 * x})(
      torch.tensor(1.0))
  assert frames['y'].tolist() == [20.0, 30.0, 40.0, 50.0]


def test_trajectory_of_states_stacks_leaves_and_sim_time():
  from dinosaur_amd import primitive_equations as pe
  st = pe.State(torch.zeros(2, 3, 4), torch.zeros(2, 3, 4), torch.zeros(2, 3, 4),
                torch.zeros(1, 3, 4), tracers={'q': torch.zeros(2, 3, 4)}, sim_time=0.0)

  def step(s):
    return pe.State(s.vorticity + 1, s.divergence, s.temperature_variation,
                    s.log_surface_pressure, tracers={'q': s.tracers['q'] + 2},
                    sim_time=s.sim_time + 0.5)
  final, traj = ti.trajectory_from_step(step, 3, 1)(st)
  assert final.sim_time == 1.5
  assert tuple(traj.vorticity.shape) == (3, 2, 3, 4)
  assert traj.vorticity[:, 0, 0, 0].tolist() == [1.0, 2.0, 3.0]
  assert traj.tracers['q'][:, 0, 0, 0].tolist() == [2.0, 4.0, 6.0]
  assert traj.sim_time.tolist() == [0.5, 1.0, 1.5]


def test_validate_state_shape_raises_state_shape_error():
  """primitive_equations.py:87-120: StateShapeError names the field."""
  from dinosaur_amd import coordinate_systems, primitive_equations as pe
  from dinosaur_amd import sigma_coordinates, spherical_harmonic as sh
  grid = sh.Grid.T21()
  coords = coordinate_systems.CoordinateSystem(grid, sigma_coordinates.SigmaCoordinates.equidistant(4))
  ms, ss = coords.modal_shape, coords.surface_modal_shape
  good = dict(vorticity=torch.zeros(ms), divergence=torch.zeros(ms),
              temperature_variation=torch.zeros(ms), log_surface_pressure=torch.zeros(ss))
  pe.validate_state_shape(pe.State(**good, tracers={'q': torch.zeros(ms)}), coords)
  # a leading member axis is allowed
  pe.validate_state_shape(pe.State(**{k: v[None] for k, v in good.items()}), coords)
  for field, shape in (('vorticity', (4, 43, 22)), ('divergence', (3,) + ms[1:]),
                       ('log_surface_pressure', ms)):
    bad = dict(good, **{field: torch.zeros(shape)})
    with pytest.raises(pe.StateShapeError, match=field):
      pe.validate_state_shape(pe.State(**bad), coords)
  with pytest.raises(pe.StateShapeError, match='tracer q'):
    pe.validate_state_shape(pe.State(**good, tracers={'q': torch.zeros(ss)}), coords)
