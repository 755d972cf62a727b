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
  each and peak `transient` above that while they run; `cache` bytes sit in
  the allocator's cache, fragmented (not usable for a new large block)."""

  def __init__(self, capacity, held, transient, cache=0):
    self.capacity, self.held, self.transient, self.cache = capacity, held, transient, cache
    self.alloc = 0
    self._peak = 0
    self.high_water = 0

  def free(self):
    return self.capacity - self.alloc - self.cache

  def allocated(self):
    return self.alloc

  def reset_peak(self):
    self._peak = self.alloc

  def peak(self):
    return self._peak

  def run_segment(self):
    self._peak = max(self._peak, self.alloc + self.held + self.transient)
    self.high_water = max(self.high_water, self.alloc + self.held + self.transient + self.cache)
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
  """Cached but fragmented blocks are not free memory: with 39 GiB of them
  (the state the allocator was in when a policy that counted them ran out of
  memory) fewer segments are kept and the device never overflows."""
  clean = _simulate(_FakeMemory(288 * GiB, 16 * GiB, 6 * GiB), 40)
  mem = _FakeMemory(288 * GiB, 16 * GiB, 6 * GiB, cache=39 * GiB)
  frag = _simulate(mem, 40)
  assert frag.kept < clean.kept
  assert mem.high_water <= mem.capacity
  assert mem.alloc + mem.cache + 2 * 22 * GiB <= mem.capacity


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


def test_cuda_memory_free_excludes_cached_blocks(monkeypatch):
  """Free memory is what the driver reports, without the allocator's cached
  (reserved but unallocated) blocks."""
  monkeypatch.setattr(torch.cuda, 'mem_get_info', lambda dev=None: (5 * GiB, 288 * GiB))
  monkeypatch.setattr(torch.cuda, 'memory_reserved', lambda dev=None: 60 * GiB)
  monkeypatch.setattr(torch.cuda, 'memory_allocated', lambda dev=None: 20 * GiB)
  assert ti._CudaMemory(torch.device('cuda', 0)).free() == 5 * GiB


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
