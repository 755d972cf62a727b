"""CPU oracle for the spectral dynamical-core hot path.

TEST INFRASTRUCTURE ONLY.  This package is a NumPy (FP64 by default)
restatement of neuralgcm/dinosaur's algorithm for the primitive-equations
path: Gauss nodes, associated Legendre tables, the real Fourier basis, the
Real-layout spherical harmonic transform, the Grid spectral operators, sigma
coordinate operators, PrimitiveEquationsSigma explicit/implicit/inverse terms,
Held-Suarez forcing, the SIL3 IMEX Runge-Kutta stepper, filters and the
initial-state recipes.  Each function cites the reference file:line it
restates.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it, and only as the checker (or as the timed CPU
baseline).  The product package ``dinosaur_amd`` never imports it: it computes
on the GPU through its own HIP kernels and fails loudly without them.

Pinning (what checks each part independently of its own formulas) is listed
in the module docstrings and exercised in ``tests/test_oracle_*.py``.
"""
