"""Ensembles of members at mixed spectral truncations, advanced together.

Members that share a truncation form a group: one batched `State` with a
leading member axis, one equation and one time step (`Group`).  A tick
advances every group whose `every` divides the tick count, so a T42 group at a
20 min step next to 10 min T85 and TL127 groups steps on every other tick and
all groups reach the same simulated time every second tick.

`MixedTruncationEnsemble` captures each distinct set of groups that step on
one tick into one HIP graph, so a tick is one graph launch.  Inside the
capture every group runs on a stream of its own, forked from and joined back
to the capture stream, so the kernels of a small grid overlap those of a large
one instead of queueing behind them.  Each group's state lives in one static
flat buffer that its captured step reads and overwrites; nothing is allocated
or copied from the host inside the graph (plans, implicit factors and fragment
tables are built by the eager warm-up ticks before capture).

The reference advances one resolution per traced function
(time_integration.py:494-564, a `lax.scan` over `step_fn`); mixed resolutions
there are separate programs.
"""
from __future__ import annotations

import dataclasses
import math
from typing import Any, Callable, Sequence

import torch


@dataclasses.dataclass
class Group:
  """Members sharing one truncation: `state` has a leading member axis (or
  none, for one member), `step_fn` advances it by its own time step, and the
  group steps on ticks divisible by `every`."""
  name: str
  step_fn: Callable[[Any], Any]
  state: Any
  every: int = 1


class MixedTruncationEnsemble:
  """Advances several `Group`s tick by tick, one HIP graph launch per tick.

  graph=False runs the same ticks eagerly (group after group on the current
  stream); concurrent=False captures the groups one after another on one
  stream.  `states()` returns the current states as views of the static
  buffers: clone what must survive the next tick.
  """

  def __init__(self, groups: Sequence[Group], *, graph: bool = True, concurrent: bool = True,
               warmup: int = 1):
    if not groups:
      raise ValueError('an ensemble needs at least one group')
    for g in groups:
      if g.every < 1:
        raise ValueError(f'group {g.name}: every must be >= 1, got {g.every}')
    self.groups = list(groups)
    self.graph_mode = graph
    self.concurrent = concurrent
    flats = [g.state.as_flat() for g in self.groups]
    self._cls = [type(f) for f in flats]
    self._layout = [f._flat[1:] for f in flats]
    self._buf = [f.flat_buffer().detach().clone() for f in flats]
    self._time = [f.sim_time for f in flats]
    self._dt = [None] * len(self.groups)
    self._period = math.lcm(*(g.every for g in self.groups))
    self._patterns = [tuple(i for i, g in enumerate(self.groups) if t % g.every == 0)
                      for t in range(self._period)]
    self.tick_count = 0
    self._streams = [torch.cuda.Stream() for _ in self.groups]
    self._graphs = {}
    with torch.no_grad():
      if graph:
        # eager warm-up on copies: builds every lazily created object and
        # measures each group's time increment, leaving the states untouched
        saved = [b.clone() for b in self._buf]
        for _ in range(warmup):
          for i in range(len(self.groups)):
            self._step_group(i, record_dt=True)
        torch.cuda.synchronize()
        for b, s in zip(self._buf, saved):
          b.copy_(s)
        for pat in sorted(set(self._patterns)):
          self._graphs[pat] = self._capture(pat)

  def _step_group(self, i, record_dt=False):
    y = self._cls[i].from_flat(self._buf[i], *self._layout[i], sim_time=self._time[i])
    y = self.groups[i].step_fn(y).as_flat()
    if record_dt and self._time[i] is not None and y.sim_time is not None:
      self._dt[i] = y.sim_time - self._time[i]
    self._buf[i].copy_(y.flat_buffer())

  def _capture(self, pat):
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
      cap = torch.cuda.current_stream()
      if self.concurrent:
        for i in pat:
          self._streams[i].wait_stream(cap)
          with torch.cuda.stream(self._streams[i]):
            self._step_group(i)
        for i in pat:
          cap.wait_stream(self._streams[i])
      else:
        for i in pat:
          self._step_group(i)
    return g

  def tick(self):
    """Advances the groups due on this tick."""
    pat = self._patterns[self.tick_count % self._period]
    if self.graph_mode:
      self._graphs[pat].replay()
    else:
      with torch.no_grad():
        for i in pat:
          self._step_group(i, record_dt=True)
    for i in pat:
      if self._time[i] is not None and self._dt[i] is not None:
        self._time[i] = self._time[i] + self._dt[i]
    self.tick_count += 1

  def run(self, ticks: int):
    for _ in range(ticks):
      self.tick()

  def states(self):
    """Current state of every group (views of the static buffers)."""
    return [c.from_flat(b, *lay, sim_time=t)
            for c, b, lay, t in zip(self._cls, self._buf, self._layout, self._time)]
