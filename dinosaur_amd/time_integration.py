"""IMEX time stepping (mirrors dinosaur/time_integration.py:56-161, 313-637).

Steppers are plain Python closures over torch operations, so a rollout can be
recorded into a HIP graph with ``torch.cuda.graphs`` and differentiated with
``torch.autograd`` (``checkpointed_rollout`` recomputes each segment of steps
in the backward pass, the analogue of ``nested_checkpoint_scan``).
"""
from __future__ import annotations

import ctypes
import dataclasses
import functools
import math
import threading
from typing import Any, Callable, Sequence

import torch
import torch.utils.checkpoint

from . import filtering


class ExplicitODE:
  """time_integration.py:56-71."""

  def explicit_terms(self, state):
    raise NotImplementedError

  @classmethod
  def from_functions(cls, explicit_terms):
    ode = cls()
    ode.explicit_terms = explicit_terms
    return ode


class ImplicitExplicitODE:
  """dx/dt = explicit_terms(x) + implicit_terms(x) (time_integration.py:74-114)."""

  def explicit_terms(self, state):
    raise NotImplementedError

  def implicit_terms(self, state):
    raise NotImplementedError

  def implicit_inverse(self, state, step_size: float):
    raise NotImplementedError

  @classmethod
  def from_functions(cls, explicit_terms, implicit_terms, implicit_inverse):
    ode = cls()
    ode.explicit_terms = explicit_terms
    ode.implicit_terms = implicit_terms
    ode.implicit_inverse = implicit_inverse
    return ode


def _is_imex(eq):
  return hasattr(eq, 'implicit_inverse') and hasattr(eq, 'implicit_terms')


def _tree_sum(items):
  items = [x for x in items if x is not None]
  out = items[0]
  for x in items[1:]:
    out = _add_trees(out, x)
  return out


def _add_trees(a, b):
  if hasattr(a, '_ARRAYS'):  # primitive_equations.State
    from .primitive_equations import State  # pylint: disable=g-import-not-at-top
    kw = {f: getattr(a, f) + getattr(b, f) for f in a._ARRAYS}
    kw['tracers'] = {k: v + b.tracers[k] if k in b.tracers else v for k, v in a.tracers.items()}
    for k, v in b.tracers.items():
      if k not in kw['tracers']:
        kw['tracers'][k] = v
    st = a.sim_time if a.sim_time is not None else b.sim_time
    kw['sim_time'] = st
    return State(**kw)
  return a + b


def compose_equations(equations: Sequence[Any]) -> ImplicitExplicitODE:
  """Sums the explicit terms of several equations, keeping the single
  implicit-explicit one for the implicit parts (time_integration.py:141-161)."""
  imex = [e for e in equations if _is_imex(e)]
  if len(imex) != 1:
    raise ValueError('compose_equations supports at most 1 ImplicitExplicitODE '
                     f'got {len(imex)}')
  (main,) = imex

  def explicit_fn(x):
    return _tree_sum([e.explicit_terms(x) for e in equations])

  return ImplicitExplicitODE.from_functions(explicit_fn, main.implicit_terms,
                                            main.implicit_inverse)


@dataclasses.dataclass
class ImExButcherTableau:
  a_ex: Sequence[Sequence[float]]
  a_im: Sequence[Sequence[float]]
  b_ex: Sequence[float]
  b_im: Sequence[float]

  def __post_init__(self):
    if len({len(self.a_ex) + 1, len(self.a_im) + 1, len(self.b_ex), len(self.b_im)}) > 1:
      raise ValueError('inconsistent Butcher tableau')


def _weighted_sum(terms):
  """sum(a * x) in the reference's order (time_integration.py:357-358)."""
  acc = None
  for a, x in terms:
    t = a * x
    acc = t if acc is None else acc + t
  return acc


def _flat_of(x):
  """(buffer, layout) of a flat-backed state, or None."""
  fb = getattr(x, 'flat_buffer', None)
  buf = fb() if fb is not None else None
  return None if buf is None else (buf, x._flat[1:])


class _ImexStage(torch.autograd.Function):
  """y0 + dt * sum(a f) + dt * sum(b g) over flat buffers in one pass
  (dino_imex_stage); the backward writes every term's cotangent in one pass."""

  @staticmethod
  def forward(ctx, n_ex, coefs, dt, y0, *xs):
    from . import _lib
    y0 = y0.contiguous()
    xs = [x.contiguous() for x in xs]
    out = torch.empty_like(y0)
    n_im = len(xs) - n_ex
    P = lambda ts: (ctypes.c_void_p * max(len(ts), 1))(*[t.data_ptr() for t in ts])
    D = lambda cs: (ctypes.c_double * max(len(cs), 1))(*cs)
    code = _lib.DINO_F64 if y0.dtype == torch.float64 else _lib.DINO_F32
    stream = ctypes.c_void_p(torch.cuda.current_stream(y0.device).cuda_stream)
    with torch.cuda.device(y0.device):
      rc = _lib.lib().dino_imex_stage(
          ctypes.c_void_p(y0.data_ptr()), n_ex, P(xs[:n_ex]), D(coefs[:n_ex]), n_im,
          P(xs[n_ex:]), D(coefs[n_ex:]), float(dt), ctypes.c_void_p(out.data_ptr()),
          y0.numel(), code, stream)
    _lib.check(rc, 'dino_imex_stage')
    ctx.n_ex, ctx.coefs, ctx.dt, ctx.n = n_ex, coefs, dt, len(xs)
    return out

  @staticmethod
  def backward(ctx, g):
    from . import _lib
    g = g.contiguous()
    n_ex, n_im = ctx.n_ex, ctx.n - ctx.n_ex
    grads = [torch.empty_like(g) for _ in range(ctx.n)]
    P = lambda ts: (ctypes.c_void_p * max(len(ts), 1))(*[t.data_ptr() for t in ts])
    D = lambda cs: (ctypes.c_double * max(len(cs), 1))(*cs)
    code = _lib.DINO_F64 if g.dtype == torch.float64 else _lib.DINO_F32
    stream = ctypes.c_void_p(torch.cuda.current_stream(g.device).cuda_stream)
    with torch.cuda.device(g.device):
      rc = _lib.lib().dino_imex_stage_adjoint(
          ctypes.c_void_p(g.data_ptr()), n_ex, D(ctx.coefs[:n_ex]), n_im,
          D(ctx.coefs[n_ex:]), float(ctx.dt), P(grads[:n_ex]), P(grads[n_ex:]), g.numel(),
          code, stream)
    _lib.check(rc, 'dino_imex_stage_adjoint')
    return (None, None, None, g) + tuple(grads)


def _stage_time(y0, dt, groups):
  """sim_time of the stage value, by the rules of elementwise state algebra
  (None as soon as an operand has none)."""
  st = y0.sim_time
  for terms in groups:
    if not terms:
      continue
    s = None
    for a, x in terms:
      if x.sim_time is None:
        s = None
        break
      t = a * x.sim_time
      s = t if s is None else s + t
    st = None if (st is None or s is None) else st + dt * s
  return st


def _stage(y0, dt, ex, im):
  """y0 + dt * sum(a_ex f) + dt * sum(a_im g) (time_integration.py:351-370)."""
  base = _flat_of(y0)
  if base is not None and base[0].is_cuda and 0 < len(ex) + len(im) and len(ex) <= 4 and len(im) <= 4:
    flats = [_flat_of(x) for _, x in ex + im]
    if all(f is not None and f[1] == base[1] for f in flats):
      coefs = [float(a) for a, _ in ex + im]
      buf = _ImexStage.apply(len(ex), coefs, float(dt), base[0], *[f[0] for f in flats])
      return type(y0).from_flat(buf, *base[1], sim_time=_stage_time(y0, dt, (ex, im)))
  out = y0
  for terms in (ex, im):
    if terms:
      out = out + dt * _weighted_sum(terms)
  return out


def _ptrs(ts):
  return (ctypes.c_void_p * max(len(ts), 1))(*[t.data_ptr() for t in ts])


def _dbls(cs):
  return (ctypes.c_double * max(len(cs), 1))(*[float(c) for c in cs])


def _code(t):
  from . import _lib  # pylint: disable=g-import-not-at-top
  return _lib.DINO_F64 if t.dtype == torch.float64 else _lib.DINO_F32


def _tally_elementwise(family, ops, like, n_full, n_range):
  """Algorithmic bytes of a flat elementwise pass: n_full full-length and
  n_range implicit-range operands (inputs and output), each read or written
  once on the triangular entries of its fields."""
  from . import _lib  # pylint: disable=g-import-not-at-top
  frac = 1.0
  if ops.degrees:
    rows = ops.field_size // ops.degrees
    frac = _lib.triangle_entries(rows, ops.degrees) / ops.field_size
  elems = n_full * ops.n + n_range * (ops.ie - ops.ib)
  _lib.tally(family, like.element_size() * elems * frac)


def _fused_stage(ops, y0, dt, ex, im, scale=None, triangle_only=False):
  """y0 + dt sum(a f) + dt sum(b g) with compact implicit terms g (one pass),
  times the per-degree `scale` (a step filter) when given.  triangle_only:
  the entries below the triangle l < m of the Real layout are left
  uninitialised (a stage value that only the implicit solve and terms,
  column-local, and the explicit terms, which read l >= m, consume)."""
  from . import _lib  # pylint: disable=g-import-not-at-top
  out = torch.empty_like(y0)
  code = _code(y0) | (_lib.STAGE_TRIANGLE_ONLY if triangle_only and ops.degrees else 0)
  fs, gs = [x for _, x in ex], [x for _, x in im]
  stream = ctypes.c_void_p(torch.cuda.current_stream(y0.device).cuda_stream)
  args = (ctypes.c_void_p(y0.data_ptr()), len(fs), _ptrs(fs), _dbls([a for a, _ in ex]),
          len(gs), _ptrs(gs), _dbls([b for b, _ in im]), float(dt),
          ctypes.c_void_p(out.data_ptr()), y0.numel(), ops.ib, ops.ie, ops.go, ops.field_size,
          ops.degrees)
  if scale is None:
    rc = _lib.lib().dino_imex_stage_ranged(*args, code, stream)
    _lib.check(rc, 'dino_imex_stage_ranged')
  else:
    rc = _lib.lib().dino_imex_stage_ranged_scaled(
        *args, ctypes.c_void_p(scale.data_ptr()), scale.numel(), code, stream)
    _lib.check(rc, 'dino_imex_stage_ranged_scaled')
  if _lib._active is not None:
    _tally_elementwise('stage', ops, y0, 2 + len(fs), len(gs))
  return out


def _fused_stage_implicit_range(ops, y0, dt, ex, im):
  """The stage combination over the implicit range [ib, ie) only -- all that
  the implicit solve and terms read -- for a stage value that feeds nothing
  else; the rest of the returned buffer is left uninitialised."""
  from . import _lib  # pylint: disable=g-import-not-at-top
  ib, ie, go = ops.ib, ops.ie, ops.go
  out = torch.empty_like(y0)
  sub = lambda t: t.view(-1)[ib:ie]  # noqa: E731
  fs = [sub(x) for _, x in ex]
  gs = [x[ib - go:ie - go] for _, x in im]
  stream = ctypes.c_void_p(torch.cuda.current_stream(y0.device).cuda_stream)
  n = ie - ib
  rc = _lib.lib().dino_imex_stage_ranged(
      ctypes.c_void_p(sub(y0).data_ptr()), len(fs), _ptrs(fs), _dbls([a for a, _ in ex]),
      len(gs), _ptrs(gs), _dbls([b for b, _ in im]), float(dt),
      ctypes.c_void_p(sub(out).data_ptr()), n, 0, n, 0, ops.field_size, ops.degrees,
      _code(y0), stream)
  _lib.check(rc, 'dino_imex_stage_ranged')
  if _lib._active is not None:
    _tally_elementwise('stage', ops, y0, 0, 2 + len(fs) + len(gs))
  return out


def _fused_gather(ops, sources, dt, like, ranged=False, scale=None, triangle_only=False):
  """One cotangent from its pending (source, coef, identity, scaled,
  implicit_only) contributions, summed in list order in one pass (at most 8
  sources: _fused_imex_step takes tableaus of at most 7 stages); scaled
  sources are multiplied by the per-degree `scale` first, implicit_only
  sources hold values on the implicit range only.  triangle_only: the
  entries below the triangle are left uninitialised (a cotangent that only
  the fused explicit terms' adjoint, which reads l >= m, consumes).  None
  when there are none."""
  from . import _lib  # pylint: disable=g-import-not-at-top
  if not sources:
    return None
  if len(sources) > 8:
    raise ValueError(f'{len(sources)} cotangent sources; at most 8 are gathered')
  out = ops.new_compact(like) if ranged else torch.empty_like(like)
  stream = ctypes.c_void_p(torch.cuda.current_stream(like.device).cuda_stream)
  flags = (ctypes.c_int * len(sources))(*[
      (_lib.GATHER_IDENTITY if i else 0) | (_lib.GATHER_SCALED if sc else 0) |
      (_lib.GATHER_IMPLICIT_RANGE if r else 0)
      for _, _, i, sc, r in sources])
  rc = _lib.lib().dino_imex_adjoint_gather(
      _ptrs([t for t, _, _, _, _ in sources]), _dbls([c for _, c, _, _, _ in sources]), flags,
      len(sources), float(dt), ctypes.c_void_p(out.data_ptr()), like.numel(),
      1 if ranged else 0, ops.ib, ops.ie, ops.go, ops.field_size, ops.degrees,
      ctypes.c_void_p(scale.data_ptr() if scale is not None else 0),
      scale.numel() if scale is not None else 0,
      _code(like) | (_lib.STAGE_TRIANGLE_ONLY if triangle_only and ops.degrees else 0), stream)
  _lib.check(rc, 'dino_imex_adjoint_gather')
  if _lib._active is not None:
    n_range = sum(1 for _, _, _, _, r in sources if r) + (1 if ranged else 0)
    _tally_elementwise('gather', ops, like, len(sources) + 1 - n_range, n_range)
  return out


def _uses(tab, n):
  """(f_used, g_used): whether stage i's explicit / implicit tendency enters a
  later stage or the output (the condition for computing it at all)."""
  a_ex, a_im, b_ex, b_im = tab.a_ex, tab.a_im, tab.b_ex, tab.b_im
  f_used = [bool(any(a_ex[j][i] for j in range(max(i, 0), n - 1)) or b_ex[i]) for i in range(n)]
  g_used = [bool(any(a_im[j][i] for j in range(max(i, 0), n - 1)) or b_im[i]) for i in range(n)]
  return f_used, g_used


def _stage_plan(tab, need_output=True):
  """(need_f, need_g, graph_only) per stage of one IMEX step: compute the
  explicit / implicit tendency of stage i, and record explicit tendency i
  for differentiation without computing its value.  With the output read,
  every tendency the step's formula uses is computed.  Without it (a
  recomputation), stage i is needed if its explicit tendency enters the
  step at all (its graph is differentiated) or its implicit tendency feeds a
  needed later stage; an implicit tendency is computed only for a needed
  later stage (its adjoint needs no forward value), and a recorded explicit
  tendency that no needed stage combines is graph-only."""
  a_ex, a_im = tab.a_ex, tab.a_im
  n = len(tab.b_ex)
  f_used, g_used = _uses(tab, n)
  if need_output:
    return list(f_used), list(g_used), [False] * n
  need_stage = [False] * n
  need_g = [False] * n
  for i in range(n - 1, -1, -1):
    feeds = any(need_stage[j + 1] and a_im[j][i] for j in range(i, n - 1))
    need_stage[i] = f_used[i] or feeds
    need_g[i] = g_used[i] and feeds
  need_f = [f_used[i] and need_stage[i] for i in range(n)]
  graph_only = [need_f[i] and not any(need_stage[j + 1] and a_ex[j][i]
                                      for j in range(i, n - 1)) for i in range(n)]
  return need_f, need_g, graph_only


def _fused_forward(ops, tab, dt, y0, keep, out_scale=None, need_output=True):
  """The stages of one IMEX step on flat buffers.  keep=True records each
  explicit evaluation's autograd graph (leaf input, output) for the reverse
  sweep; out_scale, a per-degree step filter, scales the output in the
  output combination's pass.  need_output=False (a recomputation whose
  output value nobody reads) computes only what the reverse sweep reads --
  the recorded explicit evaluations and the stage values and implicit terms
  they depend on -- and returns an uninitialised output."""
  a_ex, a_im, b_ex, b_im = tab.a_ex, tab.a_im, tab.b_ex, tab.b_im
  n = len(b_ex)
  f, g, leaves = [None] * n, [None] * n, [None] * n
  need_f, need_g, graph_only = _stage_plan(tab, need_output)

  def explicit(i, y):
    if keep:
      leaves[i] = y.detach().requires_grad_(True)
      with torch.enable_grad():
        if graph_only[i]:
          from . import _lib  # pylint: disable=g-import-not-at-top
          _lib.values_unread.active = True
          try:
            f[i] = ops.explicit(leaves[i])
          finally:
            _lib.values_unread.active = False
        else:
          f[i] = ops.explicit(leaves[i])
    else:
      f[i] = ops.explicit(y)

  explicit(0, y0)
  if need_g[0]:
    g[0] = ops.terms(y0)
  for i in range(1, n):
    if not (need_f[i] or need_g[i]):
      continue
    ex = [(a_ex[i - 1][j], f[j]) for j in range(i) if a_ex[i - 1][j]]
    im = [(a_im[i - 1][j], g[j]) for j in range(i) if a_im[i - 1][j]]
    if need_f[i]:
      # read by the solve (in place), the implicit terms and the explicit
      # terms only: the entries below the triangle are never read
      y = _fused_stage(ops, y0, dt, ex, im,
                       triangle_only=getattr(ops, 'stage_triangle_only', False))
    else:  # only the implicit terms read this stage value
      y = _fused_stage_implicit_range(ops, y0, dt, ex, im)
    ops.solve_(y, dt * a_im[i - 1][i])
    if need_f[i]:
      explicit(i, y)
    if need_g[i]:
      g[i] = ops.terms(y)
  if not need_output:
    return torch.empty_like(y0), f, g, leaves
  ex = [(b_ex[j], f[j]) for j in range(n) if b_ex[j]]
  im = [(b_im[j], g[j]) for j in range(n) if b_im[j]]
  return _fused_stage(ops, y0, dt, ex, im, out_scale), f, g, leaves


class _FusedImexStep(torch.autograd.Function):
  """One IMEX step with a hand-written reverse sweep.

  The forward keeps, for each explicit evaluation, its autograd graph; the
  backward walks the stages in reverse: explicit adjoints through those
  graphs, implicit-term adjoints accumulated in place, the solve adjoint in
  place, and each term's cotangent formed once, when its stage is reached,
  from all its contributions in one pass (dino_imex_adjoint_gather) instead
  of being written at the output and updated at every later stage.  The
  result equals autograd through the same stages up to FP rounding of the
  sums.
  """

  @staticmethod
  def forward(ctx, ops, tab, dt, y0, out_scale, need_output):
    out, f, g, leaves = _fused_forward(ops, tab, dt, y0, keep=True, out_scale=out_scale,
                                       need_output=need_output)
    ctx.ops, ctx.tab, ctx.dt, ctx.out_scale = ops, tab, dt, out_scale
    ctx.f, ctx.leaves = f, leaves
    # which implicit tendencies the step's formula contains (their adjoints
    # need no forward values, so they apply whether or not g was computed)
    ctx.has_g = [i == 0 or u for i, u in enumerate(_uses(tab, len(tab.b_ex))[1])]
    return out

  @staticmethod
  def backward(ctx, obar):
    ops, tab, dt, f, leaves = ctx.ops, ctx.tab, ctx.dt, ctx.f, ctx.leaves
    a_ex, a_im, b_ex, b_im = tab.a_ex, tab.a_im, tab.b_ex, tab.b_im
    obar = obar.contiguous()
    n = len(b_ex)
    sc = ctx.out_scale
    # Each cotangent is formed once, when it is first needed, from the list of
    # its contributions (output cotangent -- through the output filter's
    # per-degree scale, if any -- then the stage cotangents in reverse stage
    # order): the same sums, in the same order, as writing it at the output
    # combination and accumulating at every stage.
    out_sc = sc is not None
    pend_f = {j: [(obar, b_ex[j], False, out_sc, False)] for j in range(n)
              if b_ex[j] and f[j] is not None}
    pend_g = {j: [(obar, b_im[j], False, out_sc, False)] for j in range(n)
              if b_im[j] and ctx.has_g[j]}
    pend_y0 = [(obar, 1.0, True, out_sc, False)]
    # the explicit cotangents feed only the explicit terms' adjoint
    tri_f = getattr(ops, 'stage_triangle_only', False)
    for i in range(n - 1, 0, -1):
      fbar = _fused_gather(ops, pend_f.pop(i, None), dt, obar, scale=sc, triangle_only=tri_f)
      gbar = _fused_gather(ops, pend_g.pop(i, None), dt, obar, ranged=True, scale=sc)
      ybar = None
      implicit_only = False
      if f[i] is not None and fbar is not None:
        (ybar,) = torch.autograd.grad(f[i], leaves[i], fbar)
      if ctx.has_g[i] and gbar is not None:
        if ybar is None:
          # only the implicit terms read this stage value: its cotangent lives
          # on the implicit range (the rest of the buffer is never read)
          ybar = torch.empty_like(obar)
          ops.terms_adjoint_into(gbar, ybar)
          implicit_only = True
        else:
          ops.terms_adjoint_acc(gbar, ybar)
      f[i] = leaves[i] = None
      del fbar, gbar
      if ybar is None:
        continue
      ybar = ybar.contiguous()
      ops.solve_adjoint_(ybar, dt * a_im[i - 1][i])
      pend_y0.append((ybar, 1.0, True, False, implicit_only))
      for j in range(i):
        if a_ex[i - 1][j] and f[j] is not None:
          pend_f.setdefault(j, []).append((ybar, a_ex[i - 1][j], False, False, implicit_only))
        if a_im[i - 1][j] and ctx.has_g[j]:
          pend_g.setdefault(j, []).append((ybar, a_im[i - 1][j], False, False, implicit_only))
    fbar0 = _fused_gather(ops, pend_f.pop(0, None), dt, obar, scale=sc, triangle_only=tri_f)
    gbar0 = _fused_gather(ops, pend_g.pop(0, None), dt, obar, ranged=True, scale=sc)
    if f[0] is not None and fbar0 is not None:
      (gy0,) = torch.autograd.grad(f[0], leaves[0], fbar0)
      pend_y0.append((gy0.contiguous(), 1.0, True, False, False))
    del fbar0
    ybar0 = _fused_gather(ops, pend_y0, dt, obar, scale=sc)
    if ctx.has_g[0] and gbar0 is not None:
      ops.terms_adjoint_acc(gbar0, ybar0)
    ctx.f = ctx.leaves = None
    return None, None, None, ybar0, None, None


# Set while a checkpoint recomputation runs the last step of its segment,
# whose output value nobody reads.
_recompute = threading.local()


def _fused_imex_step(ops, tab, dt, y0, out_scale=None):
  """One IMEX step on a flat state through the fused kernels, or None when
  the tableau does not fit them (at most 4 terms per group, 7 stages).
  out_scale: per-degree step filter applied to the output in its pass."""
  n = len(tab.b_ex)
  rows = [tab.a_ex[i] for i in range(n - 1)] + [tab.b_ex]
  rows_im = [tab.a_im[i] for i in range(n - 1)] + [tab.b_im]
  if n > 7 or any(sum(1 for c in r if c) > 4 for r in rows + rows_im):
    return None
  buf = y0.flat_buffer()
  if torch.is_grad_enabled() and buf.requires_grad:
    need = not getattr(_recompute, 'output_unread', False)
    out = _FusedImexStep.apply(ops, tab, dt, buf, out_scale, need)
  else:
    with torch.no_grad():
      out = _fused_forward(ops, tab, dt, buf, keep=False, out_scale=out_scale)[0]
  t0 = y0.sim_time
  st = None
  if t0 is not None:
    s = None
    for b in tab.b_ex:
      if b:
        s = b if s is None else s + b
    st = t0 if s is None else t0 + dt * s
  return type(y0).from_flat(out, *y0._flat[1:], sim_time=st)


def imex_runge_kutta(tableau: ImExButcherTableau, equation, time_step: float):
  """IMEX Runge-Kutta step (time_integration.py:333-370)."""
  dt = time_step
  F, G, G_inv = equation.explicit_terms, equation.implicit_terms, equation.implicit_inverse
  a_ex, a_im, b_ex, b_im = tableau.a_ex, tableau.a_im, tableau.b_ex, tableau.b_im
  n = len(b_ex)

  hook = getattr(equation, '_flat_imex_ops', None)

  def fused(y0, filter_fn):
    """(flat y0, fused step output or None)."""
    as_flat = getattr(y0, 'as_flat', None)
    if as_flat is not None:
      y0 = as_flat()
    ops = hook(y0) if (hook is not None and as_flat is not None) else None
    if ops is None:
      return y0, None
    scale = None
    if filter_fn is not None:
      scale = filter_fn._degree_scale_tensor(y0.flat_buffer())
      if scale is None:
        return y0, None
    return y0, _fused_imex_step(ops, tableau, dt, y0, scale)

  def step_fn(y0):
    # one flat buffer per state: every stage combination is a single pass
    y0, out = fused(y0, None)
    if out is not None:
      return out
    return unfused(y0)

  def unfused(y0):
    f = [None] * n
    g = [None] * n
    f[0] = F(y0)
    g[0] = G(y0)
    for i in range(1, n):
      ex = [(a_ex[i - 1][j], f[j]) for j in range(i) if a_ex[i - 1][j]]
      im = [(a_im[i - 1][j], g[j]) for j in range(i) if a_im[i - 1][j]]
      y_star = _stage(y0, dt, ex, im)
      y = G_inv(y_star, dt * a_im[i - 1][i])
      if any(a_ex[j][i] for j in range(i, n - 1)) or b_ex[i]:
        f[i] = F(y)
      if any(a_im[j][i] for j in range(i, n - 1)) or b_im[i]:
        g[i] = G(y)
    ex = [(b_ex[j], f[j]) for j in range(n) if b_ex[j]]
    im = [(b_im[j], g[j]) for j in range(n) if b_im[j]]
    return _stage(y0, dt, ex, im)

  def with_filter(filter_fn):
    """The step followed by `filter_fn` (a per-degree scaling step filter),
    the scaling done in the fused output pass when the fused path runs."""
    def filtered_step_fn(u):
      y0, out = fused(u, filter_fn)
      if out is not None:
        return out
      return filter_fn(u, step_fn(u))
    filtered_step_fn._output_from_fused_step = True
    return filtered_step_fn

  step_fn._with_scaling_filter = with_filter
  step_fn._output_from_fused_step = True
  return step_fn


SIL3_TABLEAU = ImExButcherTableau(
    a_ex=[[1 / 3], [1 / 6, 1 / 2], [1 / 2, -1 / 2, 1]],
    a_im=[[1 / 6, 1 / 6], [1 / 3, 0, 1 / 3], [3 / 8, 0, 3 / 8, 1 / 4]],
    b_ex=[1 / 2, -1 / 2, 1, 0],
    b_im=[3 / 8, 0, 3 / 8, 1 / 4],
)


def imex_rk_sil3(equation, time_step: float):
  """Whitaker & Kar (2013) SIL3 (time_integration.py:373-405)."""
  return imex_runge_kutta(SIL3_TABLEAU, equation, time_step)


def runge_kutta_step_filter(state_filter):
  def _filter(u, u_next):
    del u
    return state_filter(u_next)
  # a per-degree scaling filter can run inside an IMEX step's output pass
  scaling = getattr(state_filter, '_degree_scaling', None)
  if scaling is not None:
    cache = {}

    def degree_scale_tensor(buf):
      """The scaling as a tensor for a flat buffer whose last axis it spans
      (rounded to the buffer's type, as the elementwise product rounds it),
      or None."""
      if buf.shape[-1] != scaling.shape[-1]:
        return None
      key = (buf.device, buf.dtype)
      t = cache.get(key)
      if t is None:
        t = torch.as_tensor(scaling, device=buf.device).to(buf.dtype).contiguous()
        cache[key] = t
      return t

    _filter._degree_scale_tensor = degree_scale_tensor
  return _filter


def exponential_step_filter(grid, dt: float, tau: float = 0.010938, order: int = 18,
                            cutoff: float = 0):
  """time_integration.py:427-455."""
  return runge_kutta_step_filter(filtering.exponential_filter(grid, dt / tau, order, cutoff))


def horizontal_diffusion_step_filter(grid, dt: float, tau: float, order: int = 1):
  """time_integration.py:459-486."""
  eigenvalues = grid.laplacian_eigenvalues
  scale = dt / (tau * abs(eigenvalues[-1]) ** order)
  return runge_kutta_step_filter(filtering.horizontal_diffusion_filter(grid, scale, order))


def step_with_filters(step_fn, filters):
  # one per-degree scaling filter after an IMEX step: fused into the step's
  # output pass (same roundings: one product per element)
  fuse = getattr(step_fn, '_with_scaling_filter', None)
  if (fuse is not None and len(filters) == 1 and
      getattr(filters[0], '_degree_scale_tensor', None) is not None):
    return fuse(filters[0])

  def _step_fn(u):
    u_next = step_fn(u)
    for filter_fn in filters:
      u_next = filter_fn(u, u_next)
    return u_next
  return _step_fn


def _is_state(x):
  return hasattr(x, '_ARRAYS') and hasattr(x, 'tracers')


def _tree_map_leaves(fn, *trees):
  """fn over matching leaves of pytrees made of tensors, arrays, numbers,
  None, States (array leaves and tracers; sim_time is treated as a leaf),
  tuples, lists and dicts."""
  first = trees[0]
  if first is None:
    return None
  if _is_state(first):
    kw = {f: fn(*[getattr(t, f) for t in trees]) for f in first._ARRAYS}
    kw['tracers'] = {k: fn(*[t.tracers[k] for t in trees]) for k in first.tracers}
    kw['sim_time'] = (None if first.sim_time is None
                      else fn(*[t.sim_time for t in trees]))
    return type(first)(**kw)
  if isinstance(first, (tuple, list)):
    return type(first)(_tree_map_leaves(fn, *xs) for xs in zip(*trees))
  if isinstance(first, dict):
    return {k: _tree_map_leaves(fn, *[t[k] for t in trees]) for k in first}
  return fn(*trees)


def _as_stackable(x):
  if isinstance(x, torch.Tensor):
    return x
  return torch.as_tensor(x)


def _stack_outputs(ys):
  """Per-step outputs -> one pytree with a leading step axis (lax.scan's ys)."""
  if not ys:
    return None
  return _tree_map_leaves(lambda *leaves: torch.stack([_as_stackable(v) for v in leaves]), *ys)


def _concat_outputs(chunks):
  if not chunks or chunks[0] is None:
    return None
  return _tree_map_leaves(lambda *leaves: torch.cat(leaves), *chunks)


def _index_xs(xs, i):
  return None if xs is None else _tree_map_leaves(lambda a: a[i], xs)


def scan(f, init, xs=None, length=None):
  """lax.scan semantics on the host loop: f(carry, x) -> (carry, y); returns
  (carry, ys) with ys stacked along a new leading axis (None if f returns
  None)."""
  if length is None:
    if xs is None:
      raise ValueError('scan needs `length` when xs is None')
    leaves = []
    _tree_map_leaves(lambda a: leaves.append(a), xs)
    length = len(leaves[0])
  carry, ys = init, []
  for i in range(length):
    carry, y = f(carry, _index_xs(xs, i))
    ys.append(y)
  if all(y is None for y in ys):
    return carry, None
  return carry, _stack_outputs(ys)


def repeated(fn, steps: int, scan_fn=scan):
  """fn applied `steps` times (time_integration.py:508-520)."""
  if steps == 1:
    return fn

  def f_repeated(x):
    x_final, _ = scan_fn(lambda c, _: (fn(c), None), x, None, steps)
    return x_final
  return f_repeated


def trajectory_from_step(step_fn, outer_steps: int, inner_steps: int, *,
                         start_with_input: bool = False, post_process_fn=lambda x: x,
                         outer_scan_fn=scan, inner_scan_fn=scan):
  """x -> (final state, trajectory): `outer_steps` frames, `inner_steps`
  steps apart, stacked along a leading axis (time_integration.py:523-564).
  Frames are the states after each outer step, or before it with
  start_with_input, passed through post_process_fn."""
  if inner_steps != 1:
    step_fn = repeated(step_fn, inner_steps, inner_scan_fn)

  def step(carry_in, _):
    carry_out = step_fn(carry_in)
    frame = carry_in if start_with_input else carry_out
    return carry_out, post_process_fn(frame)

  def multistep(x):
    return outer_scan_fn(step, x, None, outer_steps)
  return multistep


def _checkpoint(fn):
  """Non-reentrant activation checkpointing of fn (torch.utils.checkpoint):
  the forward keeps no intermediate values; the backward pass recomputes
  them.  Outputs may be any pytree."""

  def wrapped(*args):
    if torch.is_grad_enabled():
      return torch.utils.checkpoint.checkpoint(fn, *args, use_reentrant=False)
    return fn(*args)
  return wrapped


def nested_checkpoint_scan(f, init, xs=None, length=None, *, nested_lengths,
                           scan_fn=scan, checkpoint_fn=_checkpoint):
  """scan with recursive checkpointing (time_integration.py:573-637): the
  same (carry, stacked outputs) as scan(f, init, xs, length), computed as
  nested scans whose inner scans are checkpointed, so gradients keep
  O(max(nested_lengths)) intermediate states and recompute each level's
  forward pass once.  scan_fn and checkpoint_fn replace the scan and the
  checkpoint transform (defaults: `scan` and non-reentrant
  torch.utils.checkpoint)."""
  total = math.prod(nested_lengths)
  if length is not None and length != total:
    raise ValueError(f'inconsistent {length=} and {nested_lengths=}')

  def nested_reshape(a):
    a = torch.as_tensor(a)
    return a.reshape(tuple(nested_lengths) + tuple(a.shape[1:]))
  sub_xs = None if xs is None else _tree_map_leaves(nested_reshape, xs)
  return _inner_nested_scan(f, init, sub_xs, list(nested_lengths), scan_fn, checkpoint_fn)


def _inner_nested_scan(f, init, xs, lengths, scan_fn, checkpoint_fn):
  if len(lengths) == 1:
    return scan_fn(f, init, xs, lengths[0])

  @checkpoint_fn
  def sub_scans(carry, sub_xs):
    return _inner_nested_scan(f, carry, sub_xs, lengths[1:], scan_fn, checkpoint_fn)

  carry, out = scan_fn(sub_scans, init, xs, lengths[0])
  if out is None:
    return carry, None
  return carry, _tree_map_leaves(lambda a: a.reshape((-1,) + tuple(a.shape[2:])), out)


class GraphedStep:
  """`steps` applications of `step_fn` captured into one HIP graph.

  The graph owns a static input buffer and a static output buffer (the flat
  state layout of `example`).  Calling it copies the state into the input
  buffer, replays the graph and returns a state whose leaves are views of the
  output buffer: the next call overwrites them, so keep a `clone` of any
  output that must survive.  `sim_time` advances on the host by the increment
  observed at capture.  Lazily created objects (transform plans, implicit
  factors, fragment tables) are built by the warm-up calls, before capture,
  so nothing allocates or copies from the host inside the graph.  The steps
  run without autograd (a forward rollout / inference path).
  """

  def __init__(self, step_fn, example, steps: int = 1, warmup: int = 2):
    flat = example.as_flat()
    self._layout = flat._flat[1:]
    self._cls = type(flat)
    self._in = flat.flat_buffer().detach().clone()
    t0 = flat.sim_time

    def run():
      y = self._cls.from_flat(self._in, *self._layout, sim_time=t0)
      for _ in range(steps):
        y = step_fn(y)
      return y.as_flat()

    with torch.no_grad():
      side = torch.cuda.Stream()
      side.wait_stream(torch.cuda.current_stream())
      with torch.cuda.stream(side):
        for _ in range(warmup):
          run()
      torch.cuda.current_stream().wait_stream(side)
      self.graph = torch.cuda.CUDAGraph()
      with torch.cuda.graph(self.graph):
        y = run()
    self._out = y.flat_buffer()
    self._dt = None if (t0 is None or y.sim_time is None) else y.sim_time - t0

  def __call__(self, state):
    buf = state.as_flat().flat_buffer()
    if buf.shape != self._in.shape or buf.dtype != self._in.dtype:
      raise ValueError(f'graphed step expects a state buffer of shape {tuple(self._in.shape)} '
                       f'and dtype {self._in.dtype}, got {tuple(buf.shape)} {buf.dtype}')
    self._in.copy_(buf)
    self.graph.replay()
    t = None if (state.sim_time is None or self._dt is None) else state.sim_time + self._dt
    return self._cls.from_flat(self._out, *self._layout, sim_time=t)


class GraphedAdjointStep:
  """The vector-Jacobian product of one step captured into a HIP graph.

  Calling it with a state buffer x and the cotangent ybar of step(x)
  returns xbar = (d step / d x)^T ybar.  The graph holds the step's forward
  evaluation -- recorded for differentiation only: when the step is exactly
  a fused IMEX step, the work whose values the reverse sweep never reads is
  skipped, as in a checkpoint recomputation -- followed by its reverse sweep
  (the hand-written adjoint kernels), on static input, cotangent and output
  buffers.  Gradients equal those of the eager recomputation bitwise.  The
  returned cotangent is a view of the static output: the next call
  overwrites it.  Lazily created objects are built by the warm-up calls
  before capture, so nothing allocates or copies from the host inside the
  graph."""

  def __init__(self, step_fn, example, warmup: int = 2):
    flat = example.as_flat()
    self._layout = flat._flat[1:]
    self._cls = type(flat)
    self._t0 = flat.sim_time
    self._x = flat.flat_buffer().detach().clone()
    self._ybar = torch.zeros_like(self._x)
    skip = getattr(step_fn, '_output_from_fused_step', False)

    def run():
      x = self._x.detach().requires_grad_(True)
      st = self._cls.from_flat(x, *self._layout, sim_time=self._t0)
      with torch.enable_grad():
        _recompute.output_unread = skip
        try:
          y = step_fn(st)
        finally:
          _recompute.output_unread = False
        yb = y.as_flat().flat_buffer()
        (xbar,) = torch.autograd.grad(yb, x, self._ybar)
      return xbar

    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
      for _ in range(warmup):
        run()
    torch.cuda.current_stream().wait_stream(side)
    self.graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(self.graph):
      self._xbar = run()

  def __call__(self, x_buf: torch.Tensor, ybar_buf: torch.Tensor) -> torch.Tensor:
    if x_buf.shape != self._x.shape or x_buf.dtype != self._x.dtype:
      raise ValueError(f'graphed adjoint step expects buffers of shape {tuple(self._x.shape)} '
                       f'and dtype {self._x.dtype}, got {tuple(x_buf.shape)} {x_buf.dtype}')
    self._x.copy_(x_buf)
    self._ybar.copy_(ybar_buf)
    self.graph.replay()
    return self._xbar


def _graphs_for(step_fn, flat):
  """(GraphedStep, GraphedAdjointStep) of step_fn for this buffer layout,
  captured on first use and kept on the step function."""
  cache = getattr(step_fn, '_graph_cache', None)
  if cache is None:
    cache = {}
    try:
      step_fn._graph_cache = cache
    except AttributeError:  # pragma: no cover  (builtins)
      pass
  buf = flat.flat_buffer()
  key = (tuple(buf.shape), buf.dtype, buf.device, flat._flat[1:], flat.sim_time)
  g = cache.get(key)
  if g is None:
    g = (GraphedStep(step_fn, flat), GraphedAdjointStep(step_fn, flat))
    cache[key] = g
  return g


class _GraphedSegment(torch.autograd.Function):
  """`steps` steps replayed from a forward HIP graph; the backward pass
  replays the forward graph to rebuild the segment's inner states and then
  the captured adjoint graph once per step, in reverse."""

  @staticmethod
  def forward(ctx, buf, graphs, steps, layout, cls, t0):
    fwd, _ = graphs
    ctx.graphs, ctx.steps, ctx.layout, ctx.cls, ctx.t0 = graphs, steps, layout, cls, t0
    ctx.save_for_backward(buf)
    y = cls.from_flat(buf.detach(), *layout, sim_time=t0)
    for _ in range(steps):
      y = fwd(y)
    return y.flat_buffer().clone()

  @staticmethod
  def backward(ctx, ybar):
    fwd, adj = ctx.graphs
    (buf,) = ctx.saved_tensors
    xs = [buf.detach()]
    y = ctx.cls.from_flat(xs[0], *ctx.layout, sim_time=ctx.t0)
    for _ in range(ctx.steps - 1):
      y = fwd(y)
      xs.append(y.flat_buffer().clone())
    g = ybar.contiguous()
    for x in reversed(xs):
      g = adj(x, g).clone()
    return g, None, None, None, None, None


def graphed_checkpointed_rollout(step_fn, x, steps: int, checkpoint_every: int):
  """checkpointed_rollout with every step replayed from HIP graphs: the
  forward pass replays a captured forward step, and the backward pass
  replays the captured forward steps of each segment and then, per step in
  reverse, the captured adjoint step (GraphedAdjointStep) -- the
  differentiated scan body of the reference's jitted rollout
  (time_integration.py:508-520 under jax.vjp), as two HIP graphs.  The
  state must be flat-representable and on the GPU; the rollout keeps
  steps / checkpoint_every segment inputs.  Gradients equal those of
  checkpointed_rollout bitwise."""
  if steps % checkpoint_every:
    raise ValueError(f'inconsistent steps={steps} and checkpoint_every={checkpoint_every}')
  flat = x.as_flat()
  buf = flat.flat_buffer()
  layout, cls = flat._flat[1:], type(flat)
  graphs = _graphs_for(step_fn, flat)
  t = flat.sim_time
  for _ in range(steps // checkpoint_every):
    if torch.is_grad_enabled() and buf.requires_grad:
      buf = _GraphedSegment.apply(buf, graphs, checkpoint_every, layout, cls, t)
    else:
      y = cls.from_flat(buf, *layout, sim_time=t)
      for _ in range(checkpoint_every):
        y = graphs[0](y)
      buf = y.flat_buffer().clone()
    if t is not None and graphs[0]._dt is not None:
      t = t + checkpoint_every * graphs[0]._dt
  return cls.from_flat(buf, *layout, sim_time=t)


def _flatten_state(x):
  """State -> (list of tensors, rebuild)."""
  if hasattr(x, '_ARRAYS') and x.flat_buffer() is not None:
    layout, st, cls = x._flat[1:], x.sim_time, type(x)
    return [x.flat_buffer()], lambda ls: cls.from_flat(ls[0], *layout, sim_time=st)
  if hasattr(x, '_ARRAYS'):
    names = list(x._ARRAYS)
    keys = list(x.tracers.keys())
    leaves = [getattr(x, n) for n in names] + [x.tracers[k] for k in keys]
    st = x.sim_time
    cls = type(x)

    def rebuild(ls):
      kw = dict(zip(names, ls[:len(names)]))
      kw['tracers'] = dict(zip(keys, ls[len(names):]))
      kw['sim_time'] = st
      return cls(**kw)
    return leaves, rebuild
  if isinstance(x, torch.Tensor):
    return [x], lambda ls: ls[0]
  raise TypeError(f'cannot flatten {type(x)}')


def checkpointed_rollout(step_fn, x, steps: int, checkpoint_every: int, keep_graphs=0,
                         report: dict | None = None):
  """Runs `steps` steps keeping only every `checkpoint_every`-th state for the
  backward pass; each segment is recomputed when gradients are taken (the
  role of nested_checkpoint_scan, time_integration.py:573-637).

  keep_graphs: segments that record their graphs in the forward pass instead
  of being recomputed -- an int (the first `keep_graphs` segments) or 'auto'
  (decided per segment by `_GraphBudget` from the device's free memory and
  the measured high-water mark of the segments kept so far).  Kept segments
  skip the recomputation; gradients are bitwise the same.  `report`, if
  given, receives {'segments', 'kept'}."""
  if steps % checkpoint_every:
    raise ValueError(f'inconsistent steps={steps} and checkpoint_every={checkpoint_every}')
  if not (keep_graphs == 'auto' or (isinstance(keep_graphs, int) and keep_graphs >= 0)):
    raise ValueError(f"keep_graphs must be 'auto' or an int >= 0, got {keep_graphs!r}")
  as_flat = getattr(x, 'as_flat', None)
  if as_flat is not None:
    x = as_flat()
  leaves, rebuild = _flatten_state(x)
  structure = {}

  def segment(rebuild_in, *ls, output_unread=False):
    # rebuild_in is bound per segment, so the recomputation in the backward
    # pass rebuilds the same structure as the forward call did
    y = rebuild_in(list(ls))
    for i in range(checkpoint_every):
      # The recomputation's last output is only a root for differentiation:
      # a step that is exactly a fused IMEX step may skip the work whose
      # values the reverse sweep does not read.
      if (output_unread and i == checkpoint_every - 1 and
          getattr(step_fn, '_output_from_fused_step', False)):
        _recompute.output_unread = True
        try:
          y = step_fn(y)
        finally:
          _recompute.output_unread = False
      else:
        y = step_fn(y)
    out, structure['rebuild'] = _flatten_state(y)
    return tuple(out)

  n_seg = steps // checkpoint_every
  budget = _GraphBudget(keep_graphs, leaves, checkpoint_every)
  for _ in range(n_seg):
    fn = functools.partial(segment, rebuild)
    if torch.is_grad_enabled() and any(t.requires_grad for t in leaves):
      if budget.keep_next():
        leaves = list(budget.run_kept(fn, leaves))
      else:
        leaves = list(_CheckpointSegment.apply(fn, *leaves))
    else:
      leaves = list(fn(*leaves))
    rebuild = structure['rebuild']
  if report is not None:
    report.update(segments=n_seg, kept=budget.kept)
  return rebuild(leaves)


class _CudaMemory:
  """Device memory queries used by `_GraphBudget` (a seam for host tests)."""

  def __init__(self, dev):
    self.dev = dev

  def free(self) -> int:
    # memory the driver can still hand out
    return torch.cuda.mem_get_info(self.dev)[0]

  def reusable(self) -> int:
    # torch's cached segments with nothing allocated in them: the allocator
    # serves new requests from them (or returns them to the driver when a
    # request is larger).  Free space inside segments that still hold live
    # blocks (inactive splits) is fragmented and not counted.
    st = torch.cuda.memory_stats(self.dev)
    return max(0, st.get('reserved_bytes.all.current', 0) -
               st.get('allocated_bytes.all.current', 0) -
               st.get('inactive_split_bytes.all.current', 0))

  def allocated(self) -> int:
    return torch.cuda.memory_allocated(self.dev)

  def reset_peak(self) -> None:
    torch.cuda.reset_peak_memory_stats(self.dev)

  def peak(self) -> int:
    return torch.cuda.max_memory_allocated(self.dev)


def _memory_for(leaves):
  dev = leaves[0].device
  return _CudaMemory(dev) if dev.type == 'cuda' else None


class _GraphBudget:
  """Which rollout segments record their graphs in the forward pass.

  An int policy keeps the first `policy` segments.  'auto' keeps a segment
  while the memory available for it covers

      held + 2 * peak + 4 GiB

  where `held` is what a kept segment leaves allocated (its graph) and
  `peak` its high-water mark above the allocation it started from (the
  graph plus the step's temporaries: transform workspaces, nodal buffers).
  One `peak` is the segment itself; the other is reserved for the backward
  pass, which recomputes the segments that were not kept while every kept
  graph is still alive, and allocates cotangents of the same sizes.  Before
  any segment has been measured both are estimated from the input bytes
  (12x per step held, 16x peak).  Measurements take the largest seen so
  far.  Available memory is the driver's free memory plus torch's cached
  segments that hold no live block (reused as they stand); fragmented free
  space inside partly used segments is not counted.  The cache is never
  emptied here: returning it to the driver makes every later step allocate
  its buffers from the driver again, which costs far more than the
  recomputation a kept graph saves.  Resets the device's peak-memory
  statistics around kept segments."""

  RESERVE = 4 << 30

  def __init__(self, policy, leaves, steps_per_segment, memory=None):
    self.policy = policy
    self.kept = 0
    self.mem = memory if memory is not None else (
        _memory_for(leaves) if policy == 'auto' else None)
    nbytes = sum(t.numel() * t.element_size() for t in leaves)
    self.held = 12 * steps_per_segment * nbytes
    self.peak = 16 * steps_per_segment * nbytes
    self.measured = False

  def keep_next(self) -> bool:
    if self.policy != 'auto':
      return self.kept < self.policy
    if self.mem is None:
      return False
    need = self.held + 2 * self.peak + self.RESERVE
    return self.mem.free() + self.mem.reusable() >= need

  def run_kept(self, fn, leaves):
    mem = self.mem
    if mem is not None:
      before = mem.allocated()
      mem.reset_peak()
    out = fn(*leaves)
    if mem is not None:
      held = mem.allocated() - before
      peak = mem.peak() - before
      if self.measured:
        self.held, self.peak = max(self.held, held), max(self.peak, peak)
      else:
        self.held, self.peak, self.measured = held, peak, True
    self.kept += 1
    return out


class _CheckpointSegment(torch.autograd.Function):
  """Runs a segment of steps without recording anything; the backward pass
  re-runs it from the saved input with gradients enabled and differentiates
  that recomputation.  Steps that record their own graphs when gradients are
  enabled (the fused IMEX step) therefore record them once per segment, in
  the backward pass, and never during the first forward.  `fn` takes
  output_unread=True in the recomputation: the values of its outputs are
  never read there (the cotangents arrive from outside)."""

  @staticmethod
  def forward(ctx, fn, *leaves):
    ctx.fn = fn
    ctx.save_for_backward(*leaves)
    with torch.no_grad():
      outs = fn(*leaves)
    # outputs must not alias inputs (autograd would see them as views)
    ids = {t.data_ptr() for t in leaves}
    return tuple(o.clone() if o.data_ptr() in ids else o for o in outs)

  @staticmethod
  def backward(ctx, *gouts):
    saved = ctx.saved_tensors
    inputs = [t.detach().requires_grad_(t.requires_grad) for t in saved]
    with torch.enable_grad():
      outs = ctx.fn(*inputs, output_unread=True)
    pairs = [(o, g) for o, g in zip(outs, gouts) if g is not None and o.requires_grad]
    want = [t for t in inputs if t.requires_grad]
    grads = torch.autograd.grad([o for o, _ in pairs], want, [g for _, g in pairs],
                                allow_unused=True) if pairs and want else [None] * len(want)
    it = iter(grads)
    return (None,) + tuple(next(it) if t.requires_grad else None for t in inputs)


# ---------------------------------------------------------------- digital filter initialization


def _tree_map(fn, *trees):
  """fn over matching array leaves of States (scalars such as sim_time are
  taken from the first tree) or of plain tensors / arrays."""
  first = trees[0]
  if hasattr(first, '_ARRAYS'):
    kw = {f: fn(*[getattr(t, f) for t in trees]) for f in first._ARRAYS}
    kw['tracers'] = {k: fn(*[t.tracers[k] for t in trees]) for k in first.tracers}
    kw['sim_time'] = first.sim_time
    return type(first)(**kw)
  if isinstance(first, (tuple, list)):
    return type(first)(_tree_map(fn, *xs) for xs in zip(*trees))
  if isinstance(first, dict):
    return {k: _tree_map(fn, *[t[k] for t in trees]) for k in first}
  return fn(*trees)


class TimeReversedImExODE(ImplicitExplicitODE):
  """dx/dt = -explicit_terms(x) - implicit_terms(x) (time_integration.py:118-140)."""

  def __init__(self, forward_eq):
    self.forward_eq = forward_eq

  def explicit_terms(self, state):
    return _tree_map(lambda x: -x, self.forward_eq.explicit_terms(state))

  def implicit_terms(self, state):
    return _tree_map(lambda x: -x, self.forward_eq.implicit_terms(state))

  def implicit_inverse(self, state, step_size: float):
    return self.forward_eq.implicit_inverse(state, -step_size)


def accumulate_repeated(step_fn, weights, state, scan_fn=None):
  """sum_i weights[i] * step_fn^(i+1)(state) (time_integration.py:645-659).
  `scan_fn` is accepted for signature compatibility; the loop runs in Python."""
  del scan_fn
  averaged = _tree_map(lambda x: x * 0, state)
  for w in weights:
    state = step_fn(state)
    averaged = _tree_map(lambda s, a, w=float(w): a + w * s, state, averaged)
  return averaged


def _dfi_lanczos_weights(time_span: float, cutoff_period: float, dt: float):
  """Lanczos-windowed sinc weights for times 1..N (time_integration.py:662-669)."""
  import numpy as np  # pylint: disable=g-import-not-at-top
  N = round(time_span / (2 * dt))
  n = np.arange(1, N + 1)
  return np.sinc(n / (N + 1)) * np.sinc(n * time_span / (cutoff_period * N))


def digital_filter_initialization(equation, ode_solver, filters, time_span: float,
                                  cutoff_period: float, dt: float):
  """State -> digitally filtered state: the Lanczos-weighted average of the
  trajectory over [-time_span/2, time_span/2], integrated forward and, with
  the time-reversed equation, backward from the same state (Lynch & Huang
  1992; time_integration.py:672-720)."""

  def f(state):
    forward_step = step_with_filters(ode_solver(equation, dt), filters)
    backward_step = step_with_filters(ode_solver(TimeReversedImExODE(equation), dt), filters)
    weights = _dfi_lanczos_weights(time_span, cutoff_period, dt)
    init_weight = 1.0
    total_weight = init_weight + 2 * weights.sum()
    init_weight /= total_weight
    weights = weights / total_weight
    init_term = _tree_map(lambda x: x * init_weight, state)
    forward_term = accumulate_repeated(forward_step, weights, state)
    backward_term = accumulate_repeated(backward_step, weights, state)
    return _tree_map(lambda a, b, c: a + b + c, init_term, forward_term, backward_term)

  return f


def maybe_fix_sim_time_roundoff(state, dt: float):
  """`state` with sim_time rounded to a whole number of steps (:723-730)."""
  if hasattr(state, 'sim_time') and state.sim_time is not None:
    state.sim_time = dt * round(state.sim_time / dt)
  return state
