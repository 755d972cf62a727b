"""Autotuner for the launch schedules of the transform stages.

For one plan and batch size, `tune_plan` times every candidate schedule of
each stage (Legendre inverse and forward: row-resident or tiled kernels, wide
or narrow tiles, natural or paired m order; longitude DFT inverse and forward:
persistent or per-tile FFT workgroups) with HIP events on the current stream,
keeps the fastest, and records the choice on the plan.  Later transforms look
the choice up by batch size (same power-of-two bucket, else the nearest tuned
bucket) and launch with it, so a tuned plan needs no further calls.

Choices can be kept in a JSON file (`cache` argument or the DINO_AUTOTUNE_CACHE
environment variable), keyed by device architecture and CU count, grid,
spacing, data type, precision and batch bucket, so a later process loads them
without timing.

Every schedule computes the same contraction; they differ in speed and in the
FP32 accumulation order only.  FP64 data always runs the FP64 kernels, so
there is nothing to tune for it.  The reference has no counterpart: XLA picks
its einsum kernels internally (spherical_harmonic.py:262-280).
"""
from __future__ import annotations

import ctypes
import json
import math
import os
import threading

import torch

from . import _lib

LEGENDRE_CANDIDATES = (
    ('rows', _lib.SCHED_ROWS),
    ('rows_mpaired', _lib.SCHED_ROWS | _lib.SCHED_MPAIRED),
    ('rows_wide', _lib.SCHED_ROWS_WIDE),
    ('rows_wide_mpaired', _lib.SCHED_ROWS_WIDE | _lib.SCHED_MPAIRED),
    ('tiled', _lib.SCHED_TILED),
    ('tiled_mpaired', _lib.SCHED_TILED | _lib.SCHED_MPAIRED),
    ('tiled_narrow', _lib.SCHED_TILED_NARROW),
    ('tiled_narrow_mpaired', _lib.SCHED_TILED_NARROW | _lib.SCHED_MPAIRED),
)
DFT_CANDIDATES = (
    ('persistent', _lib.SCHED_PERSISTENT),
    ('per_tile', _lib.SCHED_PER_TILE),
)
STAGES = {
    _lib.STAGE_LEGENDRE_INVERSE: ('legendre_inverse', LEGENDRE_CANDIDATES),
    _lib.STAGE_LEGENDRE_FORWARD: ('legendre_forward', LEGENDRE_CANDIDATES),
    _lib.STAGE_DFT_INVERSE: ('dft_inverse', DFT_CANDIDATES),
    _lib.STAGE_DFT_FORWARD: ('dft_forward', DFT_CANDIDATES),
}

_cache_lock = threading.Lock()


def bucket(batch: int) -> int:
  """Smallest power of two >= batch (batches of one bucket share a choice)."""
  return 1 << max(0, int(batch) - 1).bit_length()


def lookup(tuned: dict, batch: int, dtype_code: int, precision: int):
  """Tuned {stage: schedule} for a batch: its bucket's entry, else the entry
  of the nearest tuned bucket (in log2) with the same dtype and precision."""
  key = (bucket(batch), dtype_code, precision)
  entry = tuned.get(key)
  if entry is not None:
    return entry
  near = [k for k in tuned if k[1:] == key[1:]]
  if not near:
    return None
  best = min(near, key=lambda k: abs(math.log2(k[0]) - math.log2(key[0])))
  return tuned[best]


def _file_key(plan, spacing, b, precision) -> str:
  # the architecture and CU count, not the marketing name, which some tool
  # environments (profilers) report as empty
  props = torch.cuda.get_device_properties(plan.device)
  dev = f'{getattr(props, "gcnArchName", "") or props.name}x{props.multi_processor_count}'
  return f'{dev}|{plan.M}x{plan.L}x{plan.I}x{plan.J}|{spacing}|f32|{precision}|b{b}'


def _load(path):
  if not path or not os.path.exists(path):
    return {}
  with open(path) as f:
    return json.load(f)


def _store(path, key, value):
  if not path:
    return
  with _cache_lock:
    data = _load(path)
    data[key] = value
    tmp = f'{path}.tmp{os.getpid()}'
    with open(tmp, 'w') as f:
      json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def tune_plan(plan, batch: int, *, precision: str = 'float32', spacing: str = 'gauss',
              reps: int = 5, cache: str | None = None, force: bool = False) -> dict:
  """Times the candidate schedules of every stage for FP32 data of `batch`
  fields and sets the fastest on the plan.  Returns, per stage, the choice,
  its schedule code and the measured ms per launch of every candidate that
  exists for this grid ('source': 'cache' when loaded from the cache file)."""
  if batch < 1:
    raise ValueError(f'batch must be >= 1, got {batch}')
  pcode = _lib.PRECISIONS[precision]
  dcode = _lib.DINO_F32
  b = bucket(batch)
  cache = cache if cache is not None else os.environ.get('DINO_AUTOTUNE_CACHE')
  fkey = _file_key(plan, spacing, b, precision)
  if cache and not force:
    hit = _load(cache).get(fkey)
    if hit is not None:
      choice = {int(k): int(v['schedule']) for k, v in hit.items()}
      plan.tuned[(b, dcode, pcode)] = choice
      for st, sc in choice.items():
        plan.set_schedule(st, sc)
      return {STAGES[int(k)][0]: dict(v, source='cache') for k, v in hit.items()}

  lib = _lib.lib()
  dev = plan.device
  R, L, I, J = plan.R, plan.L, plan.I, plan.J
  report, choice, raw = {}, {}, {}
  with torch.cuda.device(dev), torch.no_grad():
    gen = torch.Generator(device=dev).manual_seed(0)
    modal = torch.randn((batch, R, L), device=dev, generator=gen)
    fourier = torch.randn((batch, R, J), device=dev, generator=gen)
    nodal = torch.randn((batch, I, J), device=dev, generator=gen)
    fourier_out = torch.empty_like(fourier)
    stream = _stream(dev)
    p = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    calls = {
        _lib.STAGE_LEGENDRE_INVERSE: lambda: lib.dino_legendre_inverse(
            plan.handle, p(modal), p(fourier_out), batch, dcode, pcode, stream),
        _lib.STAGE_LEGENDRE_FORWARD: lambda: lib.dino_legendre_forward(
            plan.handle, p(fourier), p(modal), batch, dcode, pcode, stream),
        _lib.STAGE_DFT_INVERSE: lambda: lib.dino_dft_inverse(
            plan.handle, p(fourier), p(nodal), batch, dcode, 0, stream),
        _lib.STAGE_DFT_FORWARD: lambda: lib.dino_dft_forward(
            plan.handle, p(nodal), p(fourier_out), batch, dcode, 0, stream),
    }
    # order: the inverse Legendre reads `modal` before the forward one overwrites it
    for st in (_lib.STAGE_LEGENDRE_INVERSE, _lib.STAGE_LEGENDRE_FORWARD,
               _lib.STAGE_DFT_INVERSE, _lib.STAGE_DFT_FORWARD):
      name, cands = STAGES[st]
      times = {}
      for cname, sched in cands:
        if lib.dino_plan_set_schedule(plan.handle, st, sched) != _lib.DINO_OK:
          continue  # this family does not exist for the grid
        _lib.check(calls[st](), name)
        start = torch.cuda.Event(enable_timing=True)
        end = torch.cuda.Event(enable_timing=True)
        start.record()
        for _ in range(reps):
          _lib.check(calls[st](), name)
        end.record()
        end.synchronize()
        times[cname] = start.elapsed_time(end) / reps
      best = min(times, key=times.get)
      choice[st] = dict(cands)[best]
      plan.set_schedule(st, choice[st])
      report[name] = {'choice': best, 'schedule': choice[st], 'ms': times}
      raw[str(st)] = report[name]
    del modal, fourier, nodal, fourier_out
  plan.tuned[(b, dcode, pcode)] = choice
  _store(cache, fkey, raw)
  return report


def _stream(device):
  return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def tune_grid(grid, batches, *, device=None, reps: int = 5, cache: str | None = None,
              force: bool = False) -> dict:
  """Tunes the transform plan of `grid` (its spherical-harmonics precision,
  FP32 data) for each batch size; returns {batch: tune_plan report}."""
  sh = grid.spherical_harmonics
  if device is None:
    device = torch.device('cuda', torch.cuda.current_device())
  plan = sh.plan(device)
  return {int(bt): tune_plan(plan, int(bt), precision=sh.transform_precision,
                             spacing=sh.latitude_spacing, reps=reps, cache=cache, force=force)
          for bt in batches}
