"""ctypes binding of libdinosaur_amd.so (the gfx950 HIP kernels).

The library is built in-tree by ``python -m dinosaur_amd.build``.  There is
no fallback: if the library is missing, loading raises ImportError, and every
compute path of the package goes through it.
"""
from __future__ import annotations

import ctypes
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, '_lib', 'libdinosaur_amd.so')

DINO_OK, DINO_EINVAL, DINO_ESHAPE, DINO_EHIP, DINO_ENOMEM = range(5)
DINO_F32, DINO_F64 = 0, 1
PREC_NATIVE, PREC_BF16X3 = 0, 2
# transform stages and launch schedules (dino_plan_set_schedule)
STAGE_LEGENDRE_INVERSE, STAGE_LEGENDRE_FORWARD, STAGE_DFT_INVERSE, STAGE_DFT_FORWARD = 0, 1, 2, 3
SCHED_AUTO, SCHED_ROWS, SCHED_TILED, SCHED_TILED_NARROW, SCHED_ROWS_WIDE = 0, 1, 2, 3, 4
SCHED_PERSISTENT, SCHED_PER_TILE = 1, 2
SCHED_MPAIRED = 0x100
TARGET_ACCUMULATE, TARGET_IDENTITY, TARGET_RANGED = 1, 2, 4
GATHER_IDENTITY, GATHER_SCALED, GATHER_IMPLICIT_RANGE = 1, 2, 4
SHT_TRIANGLE_ONLY = 0x100
STAGE_TRIANGLE_ONLY = 0x100  # dtype flag of dino_imex_stage_ranged[_scaled]
PRECISIONS = {'native': PREC_NATIVE, 'float32': PREC_NATIVE, 'float64': PREC_NATIVE,
              'bf16x3': PREC_BF16X3}

_c_p = ctypes.c_void_p
_c_i = ctypes.c_int
_c_i64 = ctypes.c_int64
_c_d = ctypes.c_double
_c_dp = ctypes.POINTER(ctypes.c_double)

# name -> (restype, argtypes); every symbol declared in include/dinosaur_amd.h
SIGNATURES = {
    'dino_last_error': (ctypes.c_char_p, []),
    'dino_version': (ctypes.c_char_p, []),
    'dino_gauss_nodes_host': (_c_i, [_c_i, _c_dp, _c_dp]),
    'dino_plan_create': (_c_i, [_c_i, _c_i, _c_i, _c_i, _c_dp, _c_dp, ctypes.POINTER(_c_p)]),
    'dino_plan_destroy': (_c_i, [_c_p]),
    'dino_plan_nodes': (_c_i, [_c_p, _c_dp, _c_dp]),
    'dino_plan_table_size': (_c_i64, [_c_p]),
    'dino_plan_table': (_c_i, [_c_p, _c_dp]),
    'dino_plan_set_schedule': (_c_i, [_c_p, _c_i, _c_i]),
    'dino_plan_get_schedule': (_c_i, [_c_p, _c_i]),
    'dino_sht_workspace_bytes': (ctypes.c_size_t, [_c_p, _c_i64, _c_i]),
    'dino_sht_inverse': (_c_i, [_c_p, _c_p, _c_p, _c_p, _c_i64, _c_i, _c_i, _c_p]),
    'dino_sht_forward': (_c_i, [_c_p, _c_p, _c_p, _c_p, _c_i64, _c_i, _c_i, _c_p]),
    'dino_sht_inverse_scratch': (_c_i, [_c_p, _c_p, _c_p, _c_i64, _c_i, _c_i, _c_p]),
    'dino_sht_inverse_adjoint': (_c_i, [_c_p, _c_p, _c_p, _c_p, _c_i64, _c_i, _c_i, _c_p]),
    'dino_sht_forward_adjoint': (_c_i, [_c_p, _c_p, _c_p, _c_p, _c_i64, _c_i, _c_i, _c_p]),
    'dino_dft_forward': (_c_i, [_c_p, _c_p, _c_p, _c_i64, _c_i, _c_i, _c_p]),
    'dino_dft_inverse': (_c_i, [_c_p, _c_p, _c_p, _c_i64, _c_i, _c_i, _c_p]),
    'dino_legendre_inverse': (_c_i, [_c_p, _c_p, _c_p, _c_i64, _c_i, _c_i, _c_p]),
    'dino_legendre_forward': (_c_i, [_c_p, _c_p, _c_p, _c_i64, _c_i, _c_i, _c_p]),
    'dino_implicit_create': (_c_i, [_c_i, _c_i, _c_i, _c_dp, _c_dp, _c_dp, _c_dp, _c_dp,
                                    _c_dp, _c_dp, _c_dp, _c_d, _c_d, ctypes.POINTER(_c_p)]),
    'dino_implicit_destroy': (_c_i, [_c_p]),
    'dino_implicit_solve': (_c_i, [_c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_i64, _c_i, _c_p]),
    'dino_implicit_terms': (_c_i, [_c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_i64, _c_i, _c_p]),
    'dino_imex_stage': (_c_i, [_c_p, _c_i, _c_p, _c_p, _c_i, _c_p, _c_p, _c_d, _c_p, _c_i64, _c_i,
                               _c_p]),
    'dino_imex_stage_adjoint': (_c_i, [_c_p, _c_i, _c_p, _c_i, _c_p, _c_d, _c_p, _c_p, _c_i64,
                                       _c_i, _c_p]),
    'dino_implicit_solve_adjoint': (_c_i, [_c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_i64, _c_i,
                                           _c_p]),
    'dino_implicit_terms_adjoint_accumulate': (_c_i, [_c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p,
                                                      _c_i64, _c_i, _c_p]),
    'dino_imex_stage_ranged': (_c_i, [_c_p, _c_i, _c_p, _c_p, _c_i, _c_p, _c_p, _c_d, _c_p, _c_i64,
                                      _c_i64, _c_i64, _c_i64, _c_i64, _c_i, _c_i, _c_p]),
    'dino_imex_stage_adjoint_targets': (_c_i, [_c_p, _c_i, _c_p, _c_p, _c_p, _c_d, _c_i64, _c_i64,
                                               _c_i64, _c_i64, _c_i64, _c_i, _c_i, _c_p]),
    'dino_imex_adjoint_gather': (_c_i, [_c_p, _c_p, _c_p, _c_i, _c_d, _c_p, _c_i64, _c_i, _c_i64,
                                        _c_i64, _c_i64, _c_i64, _c_i, _c_p, _c_i64, _c_i, _c_p]),
    'dino_imex_stage_ranged_scaled': (_c_i, [_c_p, _c_i, _c_p, _c_p, _c_i, _c_p, _c_p, _c_d, _c_p,
                                             _c_i64, _c_i64, _c_i64, _c_i64, _c_i64, _c_i, _c_p,
                                             _c_i64, _c_i, _c_p]),
    'dino_implicit_terms_adjoint': (_c_i, [_c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_p, _c_i64, _c_i,
                                           _c_p]),
    'dino_nodal_create': (_c_i, [_c_i, _c_i, _c_i, _c_i, _c_i, ctypes.POINTER(_c_i), _c_i, _c_i,
                                 _c_dp, _c_dp, _c_dp, _c_dp, _c_dp, _c_dp, _c_dp, _c_dp, _c_dp,
                                 _c_d, _c_d, _c_d, _c_d, _c_d, ctypes.POINTER(_c_p)]),
    'dino_nodal_destroy': (_c_i, [_c_p]),
    'dino_nodal_fields': (_c_i64, [_c_p, _c_i64, _c_i]),
    'dino_nodal_explicit': (_c_i, [_c_p, _c_p, _c_p, _c_i64, _c_i, _c_p]),
    'dino_nodal_explicit_adjoint': (_c_i, [_c_p, _c_p, _c_p, _c_p, _c_i64, _c_i, _c_p]),
    'dino_modal_create': (_c_i, [_c_i, _c_i, _c_i, _c_i, _c_i, _c_d, _c_d, _c_dp, _c_dp, _c_dp,
                                 _c_dp, _c_dp, _c_dp, _c_dp, ctypes.POINTER(_c_p)]),
    'dino_modal_destroy': (_c_i, [_c_p]),
    'dino_modal_fields': (_c_i64, [_c_p, _c_i64, _c_i]),
    'dino_modal_prep': (_c_i, [_c_p, _c_p, _c_p, _c_p, _c_i64, _c_i, _c_p]),
    'dino_modal_prep_adjoint': (_c_i, [_c_p, _c_p, _c_p, _c_p, _c_i64, _c_i, _c_p]),
    'dino_modal_epilogue': (_c_i, [_c_p, _c_p, _c_p, _c_p, _c_i64, _c_i, _c_p]),
    'dino_modal_epilogue_adjoint': (_c_i, [_c_p, _c_p, _c_p, _c_p, _c_i64, _c_i, _c_p]),
    'dino_modal_create_ex': (_c_i, [_c_i, _c_i, _c_i, _c_i, _c_i, _c_d, _c_d, _c_dp, _c_dp, _c_dp,
                                    _c_dp, _c_dp, _c_dp, _c_dp, _c_i, ctypes.POINTER(_c_p)]),
    'dino_nodal_hybrid_create': (_c_i, [_c_i, _c_i, _c_i, _c_i, _c_i, ctypes.POINTER(_c_i), _c_i,
                                        _c_i, _c_dp, _c_dp, _c_dp, _c_dp, _c_dp, _c_d, _c_d, _c_d,
                                        _c_d, _c_d, _c_d, ctypes.POINTER(_c_p)]),
    'dino_nodal_hybrid_destroy': (_c_i, [_c_p]),
    'dino_nodal_hybrid_fields': (_c_i64, [_c_p, _c_i64, _c_i]),
    'dino_nodal_hybrid_explicit': (_c_i, [_c_p, _c_p, _c_p, _c_i64, _c_i, _c_p]),
    'dino_vregrid_create': (_c_i, [_c_i, _c_i, _c_dp, _c_dp, _c_i, _c_dp, _c_dp,
                                   ctypes.POINTER(_c_p)]),
    'dino_vregrid_destroy': (_c_i, [_c_p]),
    'dino_vregrid_apply': (_c_i, [_c_p, _c_p, _c_p, _c_p, _c_i64, _c_i64, _c_i64, _c_i, _c_p]),
}
VREGRID_CONSERVATIVE_SIGMA, VREGRID_CONSERVATIVE_PRESSURE, VREGRID_LINEAR_SIGMA = 0, 1, 2
MODAL_HYBRID = 1

_lib = None
_lock = threading.Lock()


class StateShapeError(Exception):
  """Exceptions for unexpected state shapes (primitive_equations.py:87-88)."""



# Set (per thread) while a recomputation evaluates terms whose values nobody
# reads (only their recorded graph is differentiated): autograd Functions
# whose backward reads neither their output nor anything the skipped kernel
# computes may then return an uninitialised output without launching it.
values_unread = threading.local()


def skip_values() -> bool:
  return getattr(values_unread, 'active', False)


class WorkTally:
  """Algorithmic work of the kernels launched while active, per kernel
  family: launches, bytes (each input read once and each output written
  once, modal arrays counted on the triangular truncation only) and flops,
  with the roof that binds them ('hbm', 'f32_mfma' or 'f64').  A reporting
  aid for roofline fractions: the counts come from the call sizes, not from
  hardware counters."""

  def __init__(self):
    self.families = {}

  def add(self, family: str, nbytes: float, flops: float = 0.0, roof: str = 'hbm'):
    e = self.families.setdefault(family, {'launches': 0, 'bytes': 0.0, 'flops': 0.0,
                                          'roof': roof})
    e['launches'] += 1
    e['bytes'] += float(nbytes)
    e['flops'] += float(flops)

  # MI355X published peaks
  PEAK = {'hbm': 8.0e12, 'f32_mfma': 157.3e12, 'f64': 78.6e12}

  def roof_seconds(self) -> dict:
    """Per family, the time at its binding roof: max(bytes / HBM peak,
    flops / compute peak)."""
    out = {}
    for name, e in self.families.items():
      t_mem = e['bytes'] / self.PEAK['hbm']
      t_cmp = e['flops'] / self.PEAK[e['roof']] if e['roof'] != 'hbm' else 0.0
      out[name] = max(t_mem, t_cmp)
    return out


_active = None


def triangle_entries(rows: int, degrees: int) -> int:
  """Entries of a Real-layout (2M-1, L) modal array inside the triangular
  truncation |m| <= l (cos and sin rows of each m > 0, one row for m = 0)."""
  m_count = (rows + 1) // 2
  n_tri = sum(max(0, degrees - m) for m in range(m_count))
  return 2 * n_tri - degrees


def tally(family: str, nbytes: float, flops: float = 0.0, roof: str = 'hbm') -> None:
  """Adds one launch to the active WorkTally, if any."""
  t = _active
  if t is not None:
    t.add(family, nbytes, flops, roof)


class work_tally:
  """Context manager that makes a new WorkTally active for the process (the
  backward passes of autograd run on other threads, so the tally is not
  per thread); not re-entrant."""

  def __enter__(self) -> WorkTally:
    global _active
    self.t = WorkTally()
    _active = self.t
    return self.t

  def __exit__(self, *exc):
    global _active
    _active = None
    return False


def lib() -> ctypes.CDLL:
  """Loads the HIP library (raises ImportError if it was not built)."""
  global _lib
  if _lib is not None:
    return _lib
  with _lock:
    if _lib is None:
      if not os.path.exists(LIB_PATH):
        raise ImportError(
            f'{LIB_PATH} is missing: build it with `python -m dinosaur_amd.build` '
            '(dinosaur_amd has no CPU fallback).')
      handle = ctypes.CDLL(LIB_PATH)
      for name, (res, args) in SIGNATURES.items():
        fn = getattr(handle, name)
        fn.restype = res
        fn.argtypes = args
      _lib = handle
  return _lib


# Library objects whose destruction waits for the end of a graph capture.
_released = []


def _capturing() -> bool:
  try:
    import torch  # pylint: disable=g-import-not-at-top
    return torch.cuda.is_available() and torch.cuda.is_current_stream_capturing()
  except Exception:  # pylint: disable=broad-except  (interpreter shutdown)
    return False


def release(destroy: str, handle) -> None:
  """Destroys a library object (`destroy` names its dino_*_destroy), or
  defers that while the current stream is being captured into a HIP graph.
  The destroy calls free device memory, which a capture does not permit:
  the capture would be invalidated and the error would surface at the next
  launch check.  Python's garbage collector may drop an object anywhere,
  also inside a capture, so every destructor goes through here; deferred
  objects are destroyed at the next release outside a capture."""
  if _lib is None:
    return
  if _capturing():
    _released.append((destroy, handle))
    return
  while _released:
    name, h = _released.pop()
    getattr(_lib, name)(h)
  getattr(_lib, destroy)(handle)


def check(rc: int, what: str = '') -> None:
  if rc == DINO_OK:
    return
  msg = lib().dino_last_error().decode()
  if what:
    msg = f'{what}: {msg}'
  if rc == DINO_EINVAL:
    raise ValueError(msg)
  if rc == DINO_ESHAPE:
    raise StateShapeError(msg)
  if rc == DINO_ENOMEM:
    raise MemoryError(msg)
  raise RuntimeError(msg)


def dptr(a: np.ndarray):
  assert a.dtype == np.float64 and a.flags['C_CONTIGUOUS']
  return a.ctypes.data_as(_c_dp)


def gauss_nodes_host(n: int):
  x = np.empty(n)
  w = np.empty(n)
  check(lib().dino_gauss_nodes_host(n, dptr(x), dptr(w)), 'gauss nodes')
  return x, w
