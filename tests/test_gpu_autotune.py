"""Launch schedules and the autotuner on the GPU.

Every schedule of every transform stage (row-resident / tiled / narrow-tiled
Legendre kernels in natural or paired m order, persistent / per-tile FFT)
keeps the FP32 transforms within the FP32 tolerance of the FP64 oracle and
keeps the adjoints exact transposes (dot-product test).  The autotuner picks
an existing candidate for each stage, later transforms of the same batch
bucket launch the tuned schedule, and a cache file restores the choice
without timing.
"""
import numpy as np
import pytest
import torch

from oracle import sht as osht
from dinosaur_amd import _lib, autotune, spherical_harmonic as sh

pytestmark = pytest.mark.gpu

ALL_STAGES = (_lib.STAGE_LEGENDRE_INVERSE, _lib.STAGE_LEGENDRE_FORWARD,
              _lib.STAGE_DFT_INVERSE, _lib.STAGE_DFT_FORWARD)


def _rel(a, b):
  a = np.asarray(a, np.float64)
  b = np.asarray(b, np.float64)
  return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def _grids():
  return {
      'T21': (sh.Grid.T21(), osht.Grid.T21()),
      'TL127': (sh.Grid.TL127(), osht.Grid.TL127()),
      'equiangular': (sh.Grid(24, 24, 96, 48, latitude_spacing='equiangular'),
                      osht.Grid(24, 24, 96, 48, latitude_spacing='equiangular')),
      'odd_j25': (sh.Grid(12, 13, 32, 25), osht.Grid(12, 13, 32, 25)),
      'j36_l16': (sh.Grid(16, 16, 48, 36), osht.Grid(16, 16, 48, 36)),
  }


def _modal(og, n, seed):
  rs = np.random.RandomState(seed)
  _, l = og.modal_mesh
  return rs.normal(size=(n,) + og.modal_shape) * og.mask / (l + 1) ** 1.5


def _reset(plan):
  plan.tuned.clear()
  for st in ALL_STAGES:
    plan.set_schedule(st, _lib.SCHED_AUTO)


@pytest.mark.parametrize('name', ['T21', 'TL127', 'equiangular', 'odd_j25', 'j36_l16'])
def test_every_schedule_matches_oracle_and_adjoint(name, gpu_device):
  g, og = _grids()[name]
  print(f'[test] every schedule {name}', flush=True)
  plan = g.spherical_harmonics.plan(gpu_device)
  lib = _lib.lib()
  x = _modal(og, 3, 4)
  y_ref = og.to_nodal(x)
  z_ref = og.to_modal(y_ref)
  xt = torch.tensor(x, dtype=torch.float32, device=gpu_device)
  yt = torch.tensor(y_ref, dtype=torch.float32, device=gpu_device)
  rs = np.random.RandomState(5)
  wn = torch.tensor(rs.normal(size=y_ref.shape), dtype=torch.float32, device=gpu_device)
  wm = torch.tensor(rs.normal(size=x.shape), dtype=torch.float32, device=gpu_device)
  tried = 0
  try:
    for leg_name, leg in autotune.LEGENDRE_CANDIDATES:
      for dft_name, dft in autotune.DFT_CANDIDATES:
        _reset(plan)
        ok = [lib.dino_plan_set_schedule(plan.handle, st, leg) == _lib.DINO_OK
              for st in (_lib.STAGE_LEGENDRE_INVERSE, _lib.STAGE_LEGENDRE_FORWARD)]
        if not any(ok):
          continue
        plan.set_schedule(_lib.STAGE_DFT_INVERSE, dft)
        plan.set_schedule(_lib.STAGE_DFT_FORWARD, dft)
        tried += 1
        y = g.to_nodal(xt).cpu().numpy()
        z = g.to_modal(yt).cpu().numpy()
        ey, ez = _rel(y, y_ref), _rel(z, z_ref)
        # adjoints: <A x, w> = <x, A^T w> for both directions
        xa = xt.clone().requires_grad_(True)
        ga, = torch.autograd.grad((g.to_nodal(xa) * wn).sum(), xa)
        ya = yt.clone().requires_grad_(True)
        gb, = torch.autograd.grad((g.to_modal(ya) * wm).sum(), ya)
        lhs_a = (g.to_nodal(xt).double() * wn.double()).sum().item()
        rhs_a = (xt.double() * ga.double()).sum().item()
        lhs_b = (g.to_modal(yt).double() * wm.double()).sum().item()
        rhs_b = (yt.double() * gb.double()).sum().item()
        da = abs(lhs_a - rhs_a) / abs(lhs_a)
        db = abs(lhs_b - rhs_b) / abs(lhs_b)
        print(f'  {leg_name}/{dft_name} (rows ok {ok}): inverse {ey:.1e} forward {ez:.1e} '
              f'adjoint dot {da:.1e} {db:.1e}', flush=True)
        assert ey < 2e-6 and ez < 2e-6, (leg_name, dft_name, ey, ez)
        assert da < 1e-5 and db < 1e-5, (leg_name, dft_name, da, db)
  finally:
    _reset(plan)
  assert tried >= 8


def test_autotuner_picks_applies_and_caches(gpu_device, tmp_path):
  g = sh.Grid.TL127()
  print('[test] autotuner TL127 batch 64', flush=True)
  plan = g.spherical_harmonics.plan(gpu_device)
  cache = str(tmp_path / 'tune.json')
  try:
    _reset(plan)
    rep = autotune.tune_plan(plan, 64, reps=2, cache=cache)
    for st in ALL_STAGES:
      name, cands = autotune.STAGES[st]
      r = rep[name]
      print(f'  {name}: {r["choice"]} {r["ms"]}', flush=True)
      assert r['choice'] in dict(cands) and r['choice'] in r['ms']
      assert r['ms'][r['choice']] == min(r['ms'].values())
      assert plan.schedule(st) == r['schedule']
    # later calls of the same bucket launch the tuned schedules
    for st in ALL_STAGES:
      plan.set_schedule(st, _lib.SCHED_AUTO)
    x = torch.zeros((60,) + g.modal_shape, device=gpu_device)
    g.to_nodal(x)
    for st in ALL_STAGES:
      assert plan.schedule(st) == rep[autotune.STAGES[st][0]]['schedule']
    # the cache restores the same choices without timing
    _reset(plan)
    again = autotune.tune_plan(plan, 64, cache=cache)
    for st in ALL_STAGES:
      name = autotune.STAGES[st][0]
      assert again[name]['source'] == 'cache'
      assert again[name]['schedule'] == rep[name]['schedule'] == plan.schedule(st)
  finally:
    _reset(plan)
