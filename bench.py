#!/usr/bin/env python
"""Benchmark: member-simulated-days per second of a TL255 L37 moist ensemble,
forward plus checkpointed adjoint, on one MI355X.

Workload (the full chosen-form configuration): TL255 (512 x 256 Gaussian grid,
2M-1 = 511 x L = 257 modal), 37 equidistant sigma levels, moist primitive
equations with specific humidity and two cloud tracers, 16 members, SIL3 with
the exponential step filter at dt = 5 min, seeded synthetic initial states
(rest atmosphere over seeded 2 km orography + seeded perturbations and
humidity).  One benchmark step = one SIL3 step of the whole ensemble, forward
and its adjoint.  The rollout is cut into one-step checkpoint segments that
keep only their input and are recomputed in the backward pass, as the
configured 576-step (48 h) rollout must do for all but the ~12 steps whose
graphs fit in HBM; the JSON line also reports the rate with every graph kept
(--keep-graphs auto), which applies to rollouts that short.  The loss is
quadratic in the final level-18 (sigma ~ 0.5) nodal temperature and
vorticity.

Legs, all in one run:
* chosen   - the dinosaur_amd HIP pipeline (value of the JSON line);
* plain    - the reference's formulation in plain FP32 on the same GPU:
             dense (2M-1, J, L) Legendre and (I, 2M-1) Fourier matmuls
             through torch and the nodal terms as separate tensor ops,
             otherwise the same equations and rollout;
* accuracy - the chosen and plain legs against the same workload in FP64
             (2 members, as many steps as the timed rollout), state and
             gradients, per field;
* cpu      - the oracle's whole forward path (NumPy, FP32 state arrays) on
             the host cores: two member-steps of the workload, and the T21
             L12 day in full.
The roofline entry times the dominant kernel live with HIP events on the
stream the kernel runs on.  Every leg after the headline is guarded: a
failure is recorded under config.errors and the JSON line is still printed.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import platform
import sys
import time

import numpy as np

MEMBERS = 16
GRID = 'TL255'
LAYERS = 37
DT_SECONDS = 300.0


def _parse():
  ap = argparse.ArgumentParser()
  ap.add_argument('--gpus', type=int, default=1)
  ap.add_argument('--steps', type=int, default=4)
  ap.add_argument('--warmup', type=int, default=1)
  ap.add_argument('--members', type=int, default=MEMBERS)
  ap.add_argument('--precision', default='auto', help='float32 | bf16x3 | auto (= float32)')
  ap.add_argument('--plain-steps', type=int, default=1)
  ap.add_argument('--keep-graphs', default='0',
                  help="rollout segments that keep their graphs instead of being recomputed: "
                       "0 (the headline: a long rollout), 'auto' (as HBM allows) or a count")
  ap.add_argument('--no-plain', action='store_true')
  ap.add_argument('--no-cpu', action='store_true')
  ap.add_argument('--forward-only', action='store_true')
  ap.add_argument('--no-accuracy', action='store_true')
  ap.add_argument('--kernel-only', default=None, choices=STAGES + IMPLICIT_STAGES,
                  help='run one stage kernel at the bench shape, then a 256 MiB device '
                       'copy (the calibration kernel for hardware byte counters), and exit')
  ap.add_argument('--reps', type=int, default=10)
  ap.add_argument('--profile-step', action='store_true',
                  help='build the workload, tune, run one warm-up and one timed forward + '
                       'adjoint step (one-step checkpoint, recomputed; one idle second before '
                       'and after it) and exit: the process to put under rocprofv3 for '
                       'per-kernel times and counters')
  ap.add_argument('--workload', default='tl255', choices=['tl255', 'ensemble'],
                  help='tl255: the headline fwd+adjoint line; ensemble: the mixed-truncation '
                       'ensemble (32 T85 + 24 T42 + 8 TL127, moist L24), forward ticks')
  return ap.parse_args()


STAGES = ('legendre_forward', 'legendre_inverse', 'dft_forward', 'dft_inverse')
IMPLICIT_STAGES = ('implicit_solve', 'implicit_terms')
# fields per member in one explicit-terms evaluation of the moist L37 state
# (17 per level + ln ps forward, 10 per level + 3 inverse)
FWD_FIELDS, INV_FIELDS = 630, 374


def measure_implicit(torch, eq, state, stage, reps=10):
  """Average duration of one implicit kernel launch over the ensemble state
  (HIP events on the launch stream) and its algorithmic HBM bytes."""
  from dinosaur_amd import primitive_equations as pe
  flat = pe.to_flat(state)
  d, t, p = flat.divergence, flat.temperature_variation, flat.log_surface_pressure
  outs = [torch.empty_like(d), torch.empty_like(t), torch.empty_like(p)]
  members = int(np.prod(d.shape[:-3])) if d.ndim > 3 else 1
  fn = 'dino_implicit_solve' if stage == 'implicit_solve' else 'dino_implicit_terms'
  op = eq._implicit
  call = lambda: op.apply_into(fn, 0.05, members, d, t, p, *outs)
  call()
  torch.cuda.synchronize()
  stream = torch.cuda.current_stream()
  start = torch.cuda.Event(enable_timing=True)
  end = torch.cuda.Event(enable_timing=True)
  start.record(stream)
  for _ in range(reps):
    call()
  end.record(stream)
  end.synchronize()
  secs = start.elapsed_time(end) / 1e3 / reps
  nbytes = 2.0 * (d.numel() + t.numel() + p.numel()) * d.element_size()
  return secs, nbytes, 'byte'


def kernel_only(torch, stage, precision, members, reps):
  from dinosaur_amd import _lib
  from dinosaur_amd import spherical_harmonic as sh
  if stage in IMPLICIT_STAGES:
    eq, state, _, _, _ = build(torch, precision, members)
    secs, work, _ = measure_implicit(torch, eq, state, stage, reps)
    print(f'[bench] {stage} members={members}: {secs * 1e3:.3f} ms/launch, '
          f'{work / secs / 1e9:.1f} GB/s algorithmic', file=sys.stderr, flush=True)
    return
  grid = getattr(sh.Grid, GRID)()
  batch = members * (FWD_FIELDS if stage.endswith('forward') else INV_FIELDS)
  if os.environ.get('DINO_AUTOTUNE_CACHE'):
    # the schedules the full bench tuned (read from the cache: nothing is timed)
    from dinosaur_amd import autotune
    autotune.tune_grid(grid, [batch])
  secs, work, kind = measure_kernel(torch, grid, batch, stage, _lib.PRECISIONS[precision],
                                    reps)
  unit = 'GB/s'
  rate = work / secs / 1e9
  print(f'[bench] {stage} batch={batch}: {secs * 1e3:.3f} ms/launch, {rate:.1f} {unit} '
        f'algorithmic', file=sys.stderr, flush=True)
  src = torch.ones(64 << 20, device='cuda')
  dst = torch.empty_like(src)
  for _ in range(3):
    dst.copy_(src)
  torch.cuda.synchronize()
  print('[bench] calibration copy: 268435456 bytes read, 268435456 written per launch',
        file=sys.stderr, flush=True)


def _sync(torch):
  torch.cuda.synchronize()


def build(torch, precision, members, impl=None, fused_nodal=True, dtype=None):
  from dinosaur_amd import time_integration as ti
  from dinosaur_amd import units as units_lib
  from dinosaur_amd import workloads
  eq, state, coords, ps = workloads.build_primitive(
      GRID, LAYERS, moist=True, members=members, transform_precision=precision,
      impl=impl, fused_nodal=fused_nodal, dtype=dtype or torch.float32)
  dt = ps.nondimensionalize(DT_SECONDS * units_lib.scales.units.s)
  step = ti.step_with_filters(ti.imex_rk_sil3(eq, dt),
                              [ti.exponential_step_filter(coords.horizontal, dt)])
  rng = np.random.default_rng(7)
  grid = coords.horizontal
  t18 = grid.to_nodal(state.temperature_variation[..., 18, :, :])
  z18 = grid.to_nodal(state.vorticity[..., 18, :, :])
  targets = (t18 + torch.as_tensor(rng.normal(0, 0.5, t18.shape), device=t18.device,
                                   dtype=t18.dtype),
             z18 + torch.as_tensor(rng.normal(0, 1e-6, z18.shape), device=z18.device,
                                   dtype=z18.dtype))
  return eq, state, coords, step, targets


def rollout_fwd_adj(torch, step, state, coords, targets, steps, forward_only=False,
                    keep_graphs=0, report=None, graphed=False):
  from dinosaur_amd import primitive_equations as pe
  from dinosaur_amd import time_integration as ti
  grid = coords.horizontal
  leaves = {f: getattr(state, f).detach().clone().requires_grad_(not forward_only)
            for f in pe.State._ARRAYS}
  tracers = {k: v.detach().clone().requires_grad_(not forward_only)
             for k, v in state.tracers.items()}
  s0 = pe.State(**leaves, tracers=tracers)
  if forward_only:
    with torch.no_grad():
      y = s0
      for _ in range(steps):
        y = step(y)
    return None
  if graphed:
    y = ti.graphed_checkpointed_rollout(step, s0, steps, checkpoint_every=1)
  else:
    y = ti.checkpointed_rollout(step, s0, steps, checkpoint_every=1, keep_graphs=keep_graphs,
                                report=report)
  t18 = grid.to_nodal(y.temperature_variation[..., 18, :, :])
  z18 = grid.to_nodal(y.vorticity[..., 18, :, :])
  loss = ((t18 - targets[0]) ** 2).mean() + 1e12 * ((z18 - targets[1]) ** 2).mean()
  grads = torch.autograd.grad(loss, list(leaves.values()) + list(tracers.values()))
  return loss.detach(), grads


def accuracy(torch, precision, members=2, steps=1, impl=None, fused_nodal=True):
  """Normwise relative error of the chosen pipeline against the FP64 pipeline
  on the same workload (same seeded initial state, rounded to FP32, and the
  same loss): final-state fields and gradients of the loss with respect to
  the initial state, after `steps` SIL3 steps of forward + checkpointed
  adjoint.  Returns the worst max-abs ratio and RMS ratio over the leaves
  (leaves that are identically zero in FP64, e.g. cloud tracers, are
  skipped)."""
  from dinosaur_amd import primitive_equations as pe
  from dinosaur_amd import time_integration as ti
  _, st64, coords64, step64, tg64 = build(torch, 'float64', members, dtype=torch.float64)
  _, st32, coords32, step32, _ = build(torch, precision, members, impl=impl,
                                       fused_nodal=fused_nodal)

  def run(step, st, coords, targets, dtype):
    leaves = {f: getattr(st, f).detach().to(dtype).clone().requires_grad_(True)
              for f in pe.State._ARRAYS}
    tracers = {k: v.detach().to(dtype).clone().requires_grad_(True)
               for k, v in st.tracers.items()}
    y = ti.checkpointed_rollout(step, pe.State(**leaves, tracers=tracers), steps, 1)
    grid = coords.horizontal
    t18 = grid.to_nodal(y.temperature_variation[..., 18, :, :])
    z18 = grid.to_nodal(y.vorticity[..., 18, :, :])
    tt, tz = targets[0].to(dtype), targets[1].to(dtype)
    loss = ((t18 - tt) ** 2).mean() + 1e12 * ((z18 - tz) ** 2).mean()
    grads = torch.autograd.grad(loss, list(leaves.values()) + list(tracers.values()))
    outs = [getattr(y, f).detach() for f in pe.State._ARRAYS]
    outs += [y.tracers[k].detach() for k in tracers]
    return outs, grads

  y64, g64 = run(step64, st64, coords64, tg64, torch.float64)
  # FP32 leg starts from the FP64 initial state rounded to FP32
  y32, g32 = run(step32, st64, coords32, tg64, torch.float32)

  names = list(pe.State._ARRAYS) + list(st64.tracers.keys())

  def per_leaf(xs, refs):
    out = {}
    for n, a, b in zip(names, xs, refs):
      a, b = a.double(), b.double()
      scale = b.abs().max().item()
      if scale == 0.0:
        continue
      out[n] = ((a - b).abs().max().item() / scale,
                ((a - b).pow(2).mean() / b.pow(2).mean()).sqrt().item())
    return out

  st, gr = per_leaf(y32, y64), per_leaf(g32, g64)

  # The floor FP32 state storage alone sets: the FP64 pipeline with its state
  # rounded to FP32 after every step (same initial state, same steps).
  with torch.no_grad():
    y = pe.State(**{f: getattr(st64, f).detach().float().double() for f in pe.State._ARRAYS},
                 tracers={k: v.detach().float().double() for k, v in st64.tracers.items()})
    for _ in range(steps):
      y = step64(y)
      y = pe.State(**{f: getattr(y, f).float().double() for f in pe.State._ARRAYS},
                   tracers={k: v.float().double() for k, v in y.tracers.items()})
    yf = [getattr(y, f) for f in pe.State._ARRAYS] + [y.tracers[k] for k in st64.tracers]
  floor = per_leaf(yf, y64)
  return {'state_max_rel': max(v[0] for v in st.values()),
          'state_rms_rel': max(v[1] for v in st.values()),
          'grad_max_rel': max(v[0] for v in gr.values()),
          'grad_rms_rel': max(v[1] for v in gr.values()),
          'state_by_field': st, 'grad_by_field': gr, 'members': members, 'steps': steps,
          'fp32_storage_floor': {
              'state_max_rel': max(v[0] for v in floor.values()),
              'state_rms_rel': max(v[1] for v in floor.values()),
              'state_by_field': floor,
              'definition': 'the FP64 pipeline with its state rounded to FP32 after every '
                            'step: the error FP32 state storage alone leaves'},
          'reference': 'same pipeline in FP64 (FP64 transforms, FP64 nodal terms)'}


def time_leg(torch, step, state, coords, targets, steps, warmup, forward_only=False,
             keep_graphs=0, report=None, warm_len=1, graphed=False):
  # warm_len: steps per warm-up rollout (a rollout that keeps every graph
  # needs a full-length warm-up, or its first allocations land in the timing)
  for _ in range(warmup):
    rollout_fwd_adj(torch, step, state, coords, targets, warm_len, forward_only, keep_graphs,
                    graphed=graphed)
  _sync(torch)
  t0 = time.perf_counter()
  out = rollout_fwd_adj(torch, step, state, coords, targets, steps, forward_only, keep_graphs,
                        report, graphed=graphed)
  _sync(torch)
  wall = time.perf_counter() - t0
  return wall, out


class DenseFP32Harmonics:
  """Factory for the reference formulation of the transforms (dense einsums,
  spherical_harmonic.py:262-280), executed by torch matmuls in FP32."""

  def __new__(cls, **kw):
    import torch
    from dinosaur_amd import spherical_harmonic as sh

    class _Dense(sh.RealSphericalHarmonics):
      _cache = {}

      def _tables(self, device, dtype):
        key = (self.longitude_wavenumbers, self.total_wavenumbers, self.longitude_nodes,
               self.latitude_nodes, device, dtype)
        t = _Dense._cache.get(key)
        if t is None:
          b = self.basis
          f = torch.as_tensor(b.f, device=device, dtype=dtype)
          p = torch.as_tensor(b.p, device=device, dtype=dtype)
          w = torch.as_tensor(b.w, device=device, dtype=dtype)
          t = (f, f.t().contiguous(), p.contiguous(), p.transpose(1, 2).contiguous(), w)
          _Dense._cache[key] = t
        return t

      def inverse_transform(self, x):
        f, ft, p, pt, w = self._tables(x.device, x.dtype)
        lead = x.shape[:-2]
        R, L = self.modal_shape
        I, J = self.nodal_shape
        xb = x.reshape(-1, R, L).transpose(0, 1).contiguous()          # (R, B, L)
        px = torch.bmm(xb, pt)                                          # (R, B, J)
        B = px.shape[1]
        y = (f @ px.reshape(R, B * J)).reshape(I, B, J).transpose(0, 1)  # (B, I, J)
        return y.reshape(lead + (I, J))

      def transform(self, z):
        f, ft, p, pt, w = self._tables(z.device, z.dtype)
        lead = z.shape[:-2]
        R, L = self.modal_shape
        I, J = self.nodal_shape
        zb = (z.reshape(-1, I, J) * w).transpose(0, 1).contiguous()    # (I, B, J)
        B = zb.shape[1]
        fz = (ft @ zb.reshape(I, B * J)).reshape(R, B, J)              # (R, B, J)
        y = torch.bmm(fz, p)                                            # (R, B, L)
        return y.transpose(0, 1).reshape(lead + (R, L))

    return _Dense(**kw)


def measure_kernel(torch, grid, batch, stage, precision_code, reps=10):
  """Average duration of one stage kernel launch (HIP events on the launch
  stream) and its algorithmic work."""
  from dinosaur_amd import _lib
  sh = grid.spherical_harmonics
  dev = torch.device('cuda', torch.cuda.current_device())
  plan = sh.plan(dev)
  M, L, I, J, R = plan.M, plan.L, plan.I, plan.J, plan.R
  g = torch.Generator(device=dev).manual_seed(0)
  modal = torch.randn((batch, R, L), device=dev, generator=g)
  fourier = torch.randn((batch, R, J), device=dev, generator=g)
  nodal = torch.randn((batch, I, J), device=dev, generator=g)
  lib = _lib.lib()
  stream = torch.cuda.current_stream(dev)
  sptr = ctypes.c_void_p(stream.cuda_stream)
  P = lambda t: ctypes.c_void_p(t.data_ptr())
  plan.apply_tuned(batch, _lib.DINO_F32, precision_code)  # the autotuned schedule, if any
  if stage == 'legendre_inverse':
    call = lambda: lib.dino_legendre_inverse(plan.handle, P(modal), P(fourier), batch, 0,
                                             precision_code, sptr)
  elif stage == 'legendre_forward':
    call = lambda: lib.dino_legendre_forward(plan.handle, P(fourier), P(modal), batch, 0,
                                             precision_code, sptr)
  elif stage == 'dft_forward':
    call = lambda: lib.dino_dft_forward(plan.handle, P(nodal), P(fourier), batch, 0, 1, sptr)
  else:
    call = lambda: lib.dino_dft_inverse(plan.handle, P(fourier), P(nodal), batch, 0, 0, sptr)
  _lib.check(call(), stage)
  torch.cuda.synchronize()
  start = torch.cuda.Event(enable_timing=True)
  end = torch.cuda.Event(enable_timing=True)
  start.record(stream)
  for _ in range(reps):
    call()
  end.record(stream)
  end.synchronize()
  secs = start.elapsed_time(end) / 1e3 / reps
  # every stage is HBM-bound at these batch sizes: algorithmic bytes are one
  # read of the input and one write of the output, FP32
  if stage.startswith('legendre'):
    work = 4.0 * batch * R * (J + L)
  else:
    work = 4.0 * batch * (I * J + R * J)
  return secs, work, 'byte'


def traffic_from_profiles(kernel, workload):
  """HBM bytes per launch of `kernel` from the committed FETCH_SIZE /
  WRITE_SIZE counter runs (profiles/traffic.json), scaled by the calibration
  factors measured on a copy kernel of known size; None when absent."""
  path = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'profiles', 'traffic.json')
  try:
    with open(path) as fh:
      table = json.load(fh)
  except (OSError, ValueError):
    return None
  e = table.get(kernel)
  if not e or e.get('workload') != workload:
    return None
  return 1024.0 * (e['fetch_kib'] * e['fetch_scale'] + e['write_kib'] * e['write_scale'])


def oracle_workload(name, layers, moist, dtype=np.float32):
  """The seeded rest-atmosphere recipe built by the oracle alone (its own
  grid, transforms and closed-form initial state): (equation, state, grid)
  with FP32 state arrays."""
  sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
  from oracle import pe as ope, sht as osht, sigma as osig, states as ostates
  from oracle import units as ounits
  from dinosaur_amd import workloads
  og = getattr(osht.Grid, name)()
  vertical = osig.SigmaCoordinates.equidistant(layers)
  coords = ope.Coords(og, vertical)
  ps = ounits.SimUnits.from_si()
  oro = workloads.scale_orography(og.to_nodal(workloads.orography_modal_coefficients(og)))
  st, oro_nodal, tref = ostates.isothermal_rest_atmosphere(coords, ps,
                                                           surface_height_nondim=oro)
  modal_oro = ope.truncated_modal_orography(oro_nodal, coords)
  dz, dd, dt_ = workloads.perturbation_arrays(og, layers, seed=0)
  tracers, kw = {}, {}
  if moist:
    q = og.to_modal(workloads.humidity_nodal(og, vertical.centers, seed=200))
    tracers = {'specific_humidity': q,
               'specific_cloud_liquid_water_content': np.zeros_like(q),
               'specific_cloud_ice_water_content': np.zeros_like(q)}
    kw = dict(humidity_key='specific_humidity',
              cloud_keys=('specific_cloud_liquid_water_content',
                          'specific_cloud_ice_water_content'))
  cast = lambda a: np.ascontiguousarray(a, dtype)  # noqa: E731
  state = ope.State(cast(st.vorticity + dz), cast(st.divergence + dd),
                    cast(st.temperature_variation + dt_), cast(st.log_surface_pressure),
                    tracers={k: cast(v) for k, v in tracers.items()})
  eq = ope.PrimitiveEquationsSigma(cast(tref), cast(modal_oro), coords, ps, **kw)
  return eq, state, og


def cpu_baseline(torch, member_steps=2):
  """The oracle's whole forward path (explicit terms with the reference's
  dense transforms, implicit terms and inverse, SIL3, step filter) on the
  host cores with FP32 state arrays: the T21 L12 dry day (72 steps of 20 min)
  in full, and `member_steps` steps of one TL255 L37 moist member (q + 2
  clouds, dt = 5 min), the headline's member-step, forward only."""
  sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
  from oracle import time_integration as oti
  from dinosaur_amd import units as units_lib

  def timed(name, layers, moist, dt_s, steps):
    eq, state, og = oracle_workload(name, layers, moist)
    dt = units_lib.SimUnits.from_si().nondimensionalize(dt_s * units_lib.scales.units.s)
    step = oti.step_with_filters(oti.imex_rk_sil3(eq, dt), [oti.exponential_step_filter(og, dt)])
    og._basis_as(np.float32)  # the cached transform tables are built outside the timing
    og._basis_as(np.float64)
    t0 = time.perf_counter()
    for _ in range(steps):
      state = step(state)
    secs = time.perf_counter() - t0
    finite = all(np.isfinite(a).all() for a in state.arrays().values())
    return secs, finite

  c0_secs, c0_ok = timed('T21', 12, False, 1200.0, 72)
  print(f'[bench] cpu oracle T21 L12 day: {c0_secs:.2f} s', file=sys.stderr, flush=True)
  c4_secs, c4_ok = timed(GRID, LAYERS, True, DT_SECONDS, member_steps)
  per_member_step = c4_secs / member_steps
  print(f'[bench] cpu oracle {GRID} L{LAYERS} moist: {per_member_step:.2f} s per member-step',
        file=sys.stderr, flush=True)
  threads = torch.get_num_threads()
  try:
    affinity = len(os.sched_getaffinity(0))
  except AttributeError:
    affinity = None
  cpu_model = platform.processor() or platform.machine()
  try:
    with open('/proc/cpuinfo') as fh:
      for line in fh:
        if line.startswith('model name'):
          cpu_model = line.split(':', 1)[1].strip()
          break
  except OSError:
    pass
  blas_threads = os.environ.get('OMP_NUM_THREADS')
  return {
      'value': (DT_SECONDS / 86400.0) / per_member_step,
      'unit': 'member-sim-days/s',
      'cores': int(blas_threads) if blas_threads else affinity,
      'kind': 'port',
      'sample': (f'oracle (NumPy restatement of the reference, dense-einsum transforms as '
                 f'stacked BLAS GEMMs, FP32 state arrays) SIL3 + step filter: {member_steps} '
                 f'steps of one {GRID} L{LAYERS} moist member (q + 2 clouds, dt 5 min) in '
                 f'{c4_secs:.1f} s, forward only (no adjoint); host {cpu_model}, '
                 f'{affinity} cores visible, BLAS threads {blas_threads or "default"}'),
      'finite': bool(c4_ok and c0_ok),
      'config0_t21_l12_day': {'seconds': c0_secs,
                              'sim_days_per_s': 1.0 / c0_secs,
                              'steps': 72},
      'torch_threads': threads,
  }


ENSEMBLE = (('T85', 32, 600.0, 1), ('T42', 24, 1200.0, 2), ('TL127', 8, 600.0, 1))
ENSEMBLE_LAYERS = 24


def mixed_ensemble(torch, ticks, warmup):
  """Member-sim-days/s of the mixed-truncation ensemble advanced tick by tick:
  eager, graph-captured on one stream, and graph-captured with every group on
  its own stream (one graph launch per tick).  A tick is 10 simulated minutes
  for every member (the T42 group takes 20 min steps on even ticks)."""
  from dinosaur_amd import ensemble
  from dinosaur_amd import time_integration as ti
  from dinosaur_amd import units as units_lib
  from dinosaur_amd import workloads
  ticks = max(2, ticks - ticks % 2)

  def groups():
    out, first = [], 0
    for name, members, dt_s, every in ENSEMBLE:
      eq, st, coords, ps = workloads.build_primitive(name, ENSEMBLE_LAYERS, moist=True,
                                                     members=members, first_member=first)
      dt = ps.nondimensionalize(dt_s * units_lib.scales.units.s)
      step = ti.step_with_filters(ti.imex_rk_sil3(eq, dt),
                                  [ti.exponential_step_filter(coords.horizontal, dt)])
      out.append(ensemble.Group(name, step, st.as_flat(), every))
      first += members
    return out

  members = sum(g[1] for g in ENSEMBLE)
  res = {}
  for mode, kw in (('eager', dict(graph=False)), ('graph_serial', dict(concurrent=False)),
                   ('graph_concurrent', dict(concurrent=True))):
    ens = ensemble.MixedTruncationEnsemble(groups(), **kw)
    ens.run(2 * max(1, warmup))
    _sync(torch)
    t0 = time.perf_counter()
    ens.run(ticks)
    _sync(torch)
    wall = time.perf_counter() - t0
    res[mode] = {'ms_per_tick': wall / ticks * 1e3,
                 'member_sim_days_per_s': members * ticks * (600.0 / 86400.0) / wall}
    print(f'[bench] ensemble {mode}: {res[mode]}', file=sys.stderr, flush=True)
    del ens
    torch.cuda.empty_cache()
  best = res['graph_concurrent']
  print(json.dumps({
      'metric': 'mixed-truncation ensemble member-sim-days/s (forward ticks)',
      'value': best['member_sim_days_per_s'], 'unit': 'member-sim-days/s', 'n_gpus': 1,
      'steps': ticks, 'warmup': warmup, 'ms_per_step': best['ms_per_tick'],
      'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f32',
      'data': 'synthetic (seeded rest atmosphere + orography + perturbations + humidity)',
      'config': {'workload': '32 T85 + 24 T42 + 8 TL127 members, moist L24 (q + 2 clouds), '
                             'SIL3 10 min (T42 20 min every other tick), one graph launch '
                             'per tick', 'modes': res}}))


def main():
  args = _parse()
  if args.gpus != 1:
    print(json.dumps({'error': 'bench.py measures one GPU; multi-GPU runs are not part of '
                      'this build (the ensemble would shard by member with no exchange)'}))
    sys.exit(2)
  import torch
  from dinosaur_amd import _lib
  torch.backends.cuda.matmul.allow_tf32 = False
  dev = torch.device('cuda', 0)
  torch.cuda.set_device(dev)
  precision = args.precision
  if precision == 'auto':
    # split-BF16x3 carries ~16 mantissa bits per operand: within 1e-5 for one
    # transform, but its rollout error against FP64 is ~50x that of FP32
    # (the accuracy block below measures both), so FP32 is the default.
    precision = 'float32'
  if args.kernel_only:
    kernel_only(torch, args.kernel_only, precision, args.members, args.reps)
    return
  if args.workload == 'ensemble':
    mixed_ensemble(torch, args.steps, args.warmup)
    return
  members = args.members
  print(f'[bench] chosen leg: {GRID} L{LAYERS} moist+3 tracers, {members} members, '
        f'precision={precision}', file=sys.stderr, flush=True)
  eq, state, coords, step, targets = build(torch, precision, members)
  # launch schedules of the four transform stages, autotuned for both batches
  from dinosaur_amd import autotune
  tuned = autotune.tune_grid(coords.horizontal, [members * INV_FIELDS, members * FWD_FIELDS],
                             reps=3)
  torch.cuda.empty_cache()
  tune_summary = {str(b): {st: {'choice': r['choice'],
                                'ms': {k: round(v, 3) for k, v in r['ms'].items()}}
                           for st, r in rep.items()} for b, rep in tuned.items()}
  print(f'[bench] autotune: {tune_summary}', file=sys.stderr, flush=True)
  if args.profile_step:
    rollout_fwd_adj(torch, step, state, coords, targets, 1)
    _sync(torch)
    print('[bench] profile step: warm-up done', file=sys.stderr, flush=True)
    # an idle second on each side marks the timed step in a kernel trace
    time.sleep(1.0)
    t0 = time.perf_counter()
    rollout_fwd_adj(torch, step, state, coords, targets, 1)
    _sync(torch)
    elapsed = time.perf_counter() - t0
    time.sleep(1.0)
    print(f'[bench] profile step: {1e3 * elapsed:.1f} ms', file=sys.stderr, flush=True)
    return
  # One-step checkpoint segments.  The headline keeps no graph (every step is
  # recomputed in the backward pass): the configured 48 h rollout (576 steps)
  # can keep the graphs of only ~12 steps in HBM, so its rate is this one to
  # within ~2%.  The rate of rollouts short enough for every graph to stay
  # (keep_graphs='auto') is timed alongside.
  keep = 'auto' if args.keep_graphs == 'auto' else int(args.keep_graphs)
  rollout = {}
  wall, _ = time_leg(torch, step, state, coords, targets, args.steps, args.warmup,
                     args.forward_only, keep, rollout)
  days = DT_SECONDS / 86400.0
  value = members * args.steps * days / wall
  ms_per_step = wall / args.steps * 1e3
  print(f'[bench] chosen: {wall:.3f} s for {args.steps} steps, {ms_per_step:.1f} ms/step, '
        f'{value:.3f} member-sim-days/s, graphs kept for {rollout.get("kept")} of '
        f'{rollout.get("segments")} steps', file=sys.stderr, flush=True)
  torch.cuda.empty_cache()

  # Every leg after the headline is guarded: a failure is recorded in the
  # JSON line (config.errors) and the line is printed regardless.
  errors = {}

  def leg(name, fn):
    try:
      return fn()
    except Exception as e:  # pylint: disable=broad-except
      errors[name] = f'{type(e).__name__}: {e}'[:400]
      print(f'[bench] {name} failed: {errors[name]}', file=sys.stderr, flush=True)
      return None
    finally:
      torch.cuda.empty_cache()

  out = {
      'metric': 'TL255 L37 sim-days/s fwd+adjoint, chosen vs plain FP32, <=1e-5 rel err vs FP64',
      'value': value,
      'unit': 'member-sim-days/s',
      'n_gpus': 1,
      'steps': args.steps,
      'warmup': args.warmup,
      'ms_per_step': ms_per_step,
      'higher_is_better': True,
      'scaling': 'weak',
      'vs_baseline': None,
      'dtype': 'f32' if precision == 'float32' else f'f32 ({precision} Legendre)',
      'data': 'synthetic (seeded rest atmosphere + orography + perturbations + humidity)',
      'config': {'workload': f'{GRID} L{LAYERS} moist (q + 2 clouds), {members} members, '
                             f'SIL3 dt=5 min, forward + checkpointed adjoint, timed over '
                             f'{args.steps} steps of the rollout (one-step checkpoint '
                             'segments, each recomputed in the backward pass, as in a '
                             '576-step 48 h rollout)',
                 'rollout': {'steps': args.steps, 'keep_graphs': keep,
                             'graphs_kept': rollout.get('kept'),
                             'recomputed': (rollout.get('segments') or 0) -
                                           (rollout.get('kept') or 0)},
                 'members': members, 'grid': GRID, 'levels': LAYERS,
                 'ensemble_sim_days_per_s': args.steps * days / wall,
                 'forward_only': args.forward_only, 'autotune': tune_summary,
                 'errors': errors},
      'roofline': None,
      'cpu_baseline': None,
  }
  cfg = out['config']
  try:
    if not args.forward_only:
      def other_leg():
        other_keep = 0 if rollout.get('kept') else 'auto'
        orep = {}
        owall, _ = time_leg(torch, step, state, coords, targets, args.steps, 1, False,
                            other_keep, orep, warm_len=args.steps)
        res = {'keep_graphs': other_keep, 'ms_per_step': owall / args.steps * 1e3,
               'graphs_kept': orep.get('kept'),
               'member_sim_days_per_s': members * args.steps * days / owall}
        print(f'[bench] chosen, keep_graphs={other_keep!r}: {res["ms_per_step"]:.1f} ms/step, '
              f'graphs kept for {orep.get("kept")} of {orep.get("segments")} steps',
              file=sys.stderr, flush=True)
        return res
      cfg['rollout']['other_policy'] = leg('other_policy', other_leg)

    if not args.forward_only:
      def graphed_leg():
        gwall, _ = time_leg(torch, step, state, coords, targets, args.steps, 1, graphed=True)
        res = {'ms_per_step': gwall / args.steps * 1e3,
               'member_sim_days_per_s': members * args.steps * days / gwall,
               'form': 'forward step and adjoint step (forward recording + reverse sweep) '
                       'replayed from two HIP graphs, one-step checkpoints'}
        print(f'[bench] chosen, graph-captured fwd+adjoint: {res["ms_per_step"]:.1f} ms/step',
              file=sys.stderr, flush=True)
        return res
      cfg['rollout']['graphed_adjoint'] = leg('graphed_adjoint', graphed_leg)

    if not args.forward_only:
      cfg['step_roofline'] = leg('step_roofline', lambda: step_roofline(
          torch, step, state, coords, targets, ms_per_step))

    out['roofline'] = leg('roofline', lambda: dominant_kernel_roofline(
        torch, coords.horizontal, members, precision))
    cfg['graphed_forward_ms_per_step'] = leg('graphed_forward', lambda: graphed_forward_ms(
        torch, step, state))
    del eq, state, step, targets
    torch.cuda.empty_cache()
    plain = None
    if not args.no_plain:
      plain = leg('plain', lambda: plain_leg(torch, members, args.plain_steps,
                                             args.forward_only, keep, days))
    cfg['plain_fp32'] = plain
    cfg['chosen_over_plain'] = (value / plain['value']) if plain else None
    if not args.no_accuracy:
      steps_acc = args.steps
      print(f'[bench] accuracy vs FP64 (2 members, {steps_acc} steps)', file=sys.stderr,
            flush=True)
      acc = leg('accuracy', lambda: accuracy(torch, precision, steps=steps_acc))
      print(f'[bench] accuracy: {acc}', file=sys.stderr, flush=True)
      if acc is not None and not args.no_plain:
        pacc = leg('accuracy_plain', lambda: accuracy(
            torch, 'float32', steps=steps_acc, impl=DenseFP32Harmonics, fused_nodal=False))
        if pacc is not None:
          acc['plain_fp32'] = {k: pacc[k] for k in ('state_max_rel', 'state_rms_rel',
                                                     'grad_max_rel', 'grad_rms_rel')}
          print(f'[bench] plain accuracy: {acc["plain_fp32"]}', file=sys.stderr, flush=True)
      cfg['accuracy_vs_fp64'] = acc
    if not args.no_cpu:
      out['cpu_baseline'] = leg('cpu_baseline', lambda: cpu_baseline(torch))
  finally:
    print(json.dumps(out), flush=True)


def step_roofline(torch, step, state, coords, targets, ms_per_step):
  """Sum over the kernels of one forward + checkpointed-adjoint step (the
  headline's rollout, one step, with the loss) of each launch's time at its
  binding roof -- algorithmic bytes at 8 TB/s, or flops at the FP32 MFMA /
  FP64 peak -- against the measured time per step."""
  from dinosaur_amd import _lib
  with _lib.work_tally() as t:
    rollout_fwd_adj(torch, step, state, coords, targets, 1)
    _sync(torch)
  roof = t.roof_seconds()
  total = sum(roof.values())
  fams = {k: {'launches': e['launches'], 'gbytes': e['bytes'] / 1e9,
              'gflops': e['flops'] / 1e9, 'roof': e['roof'],
              't_roof_ms': roof[k] * 1e3}
          for k, e in sorted(t.families.items(), key=lambda kv: -roof[kv[0]])}
  res = {'t_roof_ms': total * 1e3, 'wall_ms': ms_per_step, 'frac': total * 1e3 / ms_per_step,
         'gbytes': sum(e['bytes'] for e in t.families.values()) / 1e9,
         'families': fams,
         'definition': 'sum over launches of max(algorithmic bytes / 8 TB/s, flops / '
                       'peak of the binding unit) divided by the measured ms per step'}
  print(f'[bench] step roofline: {res["t_roof_ms"]:.1f} ms at the roofs vs '
        f'{ms_per_step:.1f} ms measured (frac {res["frac"]:.3f})', file=sys.stderr, flush=True)
  return res


def dominant_kernel_roofline(torch, grid, members, precision):
  """The longest of the four transform-stage kernels at the workload's batch,
  timed live on its launch stream with HIP events."""
  from dinosaur_amd import _lib
  pcode = _lib.PRECISIONS[precision]
  kern = {}
  for stage in STAGES:
    batch = members * (FWD_FIELDS if stage.endswith('forward') else INV_FIELDS)
    kern[stage] = measure_kernel(torch, grid, batch, stage, pcode)
    print(f'[bench] {stage}: {kern[stage][0] * 1e3:.3f} ms/launch', file=sys.stderr, flush=True)
  dom = max(kern, key=lambda k: kern[k][0])
  secs, work, _ = kern[dom]
  roof = {'bound': 'hbm', 'achieved': work / secs / 1e9, 'peak': 8000.0, 'unit': 'GB/s'}
  roof['frac'] = roof['achieved'] / roof['peak']
  batch = members * (FWD_FIELDS if dom.endswith('forward') else INV_FIELDS)
  roof['traffic'] = traffic_from_profiles(dom, f'{GRID} batch={batch} {precision}')
  roof['kernel'] = dom
  roof['algorithmic_bytes'] = work
  roof['ms_per_launch'] = secs * 1e3
  return roof


def graphed_forward_ms(torch, step, state):
  """Forward-only steps replayed from a HIP graph (inference path)."""
  from dinosaur_amd import primitive_equations as pe_g
  from dinosaur_amd import time_integration as ti_g
  graphed = ti_g.GraphedStep(step, pe_g.to_flat(state))
  y = pe_g.to_flat(state)
  _sync(torch)
  g0 = time.perf_counter()
  for _ in range(4):
    y = graphed(y)
  _sync(torch)
  ms = (time.perf_counter() - g0) / 4 * 1e3
  print(f'[bench] graphed forward step: {ms:.1f} ms', file=sys.stderr, flush=True)
  return ms


def plain_leg(torch, members, steps, forward_only, keep, days):
  print('[bench] plain FP32 leg (dense reference formulation)', file=sys.stderr, flush=True)
  _, pstate, pcoords, pstep, ptargets = build(torch, 'float32', members,
                                              impl=DenseFP32Harmonics, fused_nodal=False)
  prollout = {}
  pwall, _ = time_leg(torch, pstep, pstate, pcoords, ptargets, steps, 1, forward_only, keep,
                      prollout)
  pval = members * steps * days / pwall
  print(f'[bench] plain: {pwall:.3f} s, {pval:.3f} member-sim-days/s', file=sys.stderr,
        flush=True)
  return {'value': pval, 'ms_per_step': pwall / steps * 1e3, 'steps': steps,
          'graphs_kept': prollout.get('kept')}


if __name__ == '__main__':
  main()
