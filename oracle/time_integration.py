"""IMEX time stepping and filters (oracle restatement).

Restates dinosaur/time_integration.py:
  compose_equations              :141-161
  ImExButcherTableau / imex_runge_kutta / imex_rk_sil3   :313-405
  exponential_step_filter        :427-455
  step_with_filters / repeated / trajectory_from_step    :494-564
and dinosaur/filtering.py:38-89 (exponential_filter, shape-gated scaling).

Pinned in tests/test_oracle_time.py by closed-form ODE solutions (the SIL3
third-order explicit / second-order implicit convergence on a split linear
ODE) and the filter KAT exp(-16) at the top wavenumber.
"""
from __future__ import annotations

import dataclasses

import numpy as np

from .pe import State


@dataclasses.dataclass
class Tableau:
  a_ex: list
  a_im: list
  b_ex: list
  b_im: list


SIL3 = Tableau(
    a_ex=[[1 / 3], [1 / 6, 1 / 2], [1 / 2, -1 / 2, 1]],
    a_im=[[1 / 6, 1 / 6], [1 / 3, 0, 1 / 3], [3 / 8, 0, 3 / 8, 1 / 4]],
    b_ex=[1 / 2, -1 / 2, 1, 0],
    b_im=[3 / 8, 0, 3 / 8, 1 / 4],
)


def _stage(y0, dt, ex, im):
  """y0 + dt * sum(a * f) + dt * sum(a * g) (time_integration.py:355-369)."""
  out = y0
  for terms in (ex, im):
    if terms:
      acc = None
      for a, x in terms:
        t = a * x
        acc = t if acc is None else acc + t
      out = out + dt * acc
  return out


def imex_runge_kutta(tableau, equation, time_step):
  """time_integration.py:333-370."""
  dt = time_step
  F, G, G_inv = equation.explicit_terms, equation.implicit_terms, equation.implicit_inverse
  a_ex, a_im, b_ex, b_im = tableau.a_ex, tableau.a_im, tableau.b_ex, tableau.b_im
  n = len(b_ex)

  def step_fn(y0):
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

  return step_fn


def imex_rk_sil3(equation, time_step):
  return imex_runge_kutta(SIL3, equation, time_step)


class Composed:
  """compose_equations (time_integration.py:141-161)."""

  def __init__(self, equations):
    imex = [e for e in equations if hasattr(e, 'implicit_inverse')]
    if len(imex) != 1:
      raise ValueError('compose_equations supports at most 1 ImplicitExplicitODE')
    self.equations = equations
    self.imex = imex[0]

  def explicit_terms(self, x):
    terms = [e.explicit_terms(x) for e in self.equations]
    out = terms[0]
    for t in terms[1:]:
      out = State(
          out.vorticity + t.vorticity, out.divergence + t.divergence,
          out.temperature_variation + t.temperature_variation,
          out.log_surface_pressure + t.log_surface_pressure,
          {k: v + t.tracers.get(k, 0) for k, v in out.tracers.items()},
          out.sim_time if out.sim_time is not None else t.sim_time)
    return out

  def implicit_terms(self, x):
    return self.imex.implicit_terms(x)

  def implicit_inverse(self, x, eta):
    return self.imex.implicit_inverse(x, eta)


def exponential_filter_scaling(grid, attenuation=16, order=18, cutoff=0):
  """filtering.py:38-89: scaling over total wavenumber."""
  _, l = grid.modal_axes
  k = l / l.max()
  return np.exp((k > cutoff) * (-attenuation * (((k - cutoff) / (1 - cutoff)) ** (2 * order))))


def exponential_filter(grid, attenuation=16, order=18, cutoff=0):
  scaling = exponential_filter_scaling(grid, attenuation, order, cutoff)

  def rescale(x):
    if x is None:
      return x
    shp = np.shape(x)
    if shp == np.broadcast_shapes(shp, scaling.shape):
      return (x * scaling).astype(x.dtype)
    return x

  def apply(state):
    if isinstance(state, State):
      return state.map(rescale)
    return rescale(state)

  return apply


def exponential_step_filter(grid, dt, tau=0.010938, order=18, cutoff=0):
  """time_integration.py:427-455."""
  f = exponential_filter(grid, dt / tau, order, cutoff)
  return lambda u, u_next: f(u_next)


def step_with_filters(step_fn, filters):
  def _step(u):
    u_next = step_fn(u)
    for flt in filters:
      u_next = flt(u, u_next)
    return u_next
  return _step


def repeated(fn, steps):
  def f(x):
    for _ in range(steps):
      x = fn(x)
    return x
  return f


def trajectory_from_step(step_fn, outer_steps, inner_steps, start_with_input=False):
  inner = repeated(step_fn, inner_steps)

  def multistep(x):
    frames = []
    for _ in range(outer_steps):
      y = inner(x)
      frames.append(x if start_with_input else y)
      x = y
    return x, frames

  return multistep


# ---------------------------------------------------------------- digital filter initialisation


def _negate(x):
  if isinstance(x, State):
    return x.map(lambda v: None if v is None else -v)
  return -x


class TimeReversed:
  """dx/dt = -F(x) - G(x); the inverse takes -eta (time_integration.py:
  118-140)."""

  def __init__(self, forward_eq):
    self.forward_eq = forward_eq

  def explicit_terms(self, x):
    return _negate(self.forward_eq.explicit_terms(x))

  def implicit_terms(self, x):
    return _negate(self.forward_eq.implicit_terms(x))

  def implicit_inverse(self, x, eta):
    return self.forward_eq.implicit_inverse(x, -eta)


def accumulate_repeated(step_fn, weights, state):
  """sum_n w_n step^n(state) for n = 1..N (time_integration.py:645-663)."""
  acc = None
  for w in weights:
    state = step_fn(state)
    term = state * w if isinstance(state, State) else w * state
    acc = term if acc is None else acc + term
  return acc


def dfi_lanczos_weights(time_span, cutoff_period, dt):
  """Lanczos-windowed sinc weights for n = 1..N, N = round(span / (2 dt))
  (time_integration.py:666-674)."""
  n_steps = round(time_span / (2 * dt))
  n = np.arange(1, n_steps + 1)
  return np.sinc(n / (n_steps + 1)) * np.sinc(n * time_span / (cutoff_period * n_steps))


def digital_filter_initialization(equation, ode_solver, filters, time_span, cutoff_period, dt):
  """Weighted mean of the states N steps forward and N steps backward in
  time, weights normalised with 1 for the initial state (Lynch & Huang 1992;
  time_integration.py:677-720)."""

  def f(state):
    fwd = step_with_filters(ode_solver(equation, dt), filters)
    bwd = step_with_filters(ode_solver(TimeReversed(equation), dt), filters)
    w = dfi_lanczos_weights(time_span, cutoff_period, dt)
    total = 1.0 + 2 * w.sum()
    w0, w = 1.0 / total, w / total
    out = state * w0 if isinstance(state, State) else w0 * state
    return out + accumulate_repeated(fwd, w, state) + accumulate_repeated(bwd, w, state)

  return f


def horizontal_diffusion_step_filter(grid, dt, tau, order=1):
  """exp(-scale (l (l + 1) / a^2)^order) per degree with scale = dt /
  (tau |lambda_max|^order), applied after each step (time_integration.py:
  459-486, filtering.py:92-100)."""
  ev = grid.laplacian_eigenvalues
  scale = dt / (tau * abs(ev[-1]) ** order)
  scaling = np.exp(-scale * (-ev) ** order)

  def rescale(x):
    if x is None:
      return x
    shp = np.shape(x)
    if shp == np.broadcast_shapes(shp, scaling.shape):
      return (x * scaling).astype(x.dtype)
    return x

  def apply(u, u_next):
    return u_next.map(rescale) if isinstance(u_next, State) else rescale(u_next)

  return apply
