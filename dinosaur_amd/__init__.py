"""dinosaur_amd: the spectral dynamical core of neuralgcm/dinosaur on MI355X.

Reference-shaped modules (spherical_harmonic, sigma_coordinates,
coordinate_systems, primitive_equations, held_suarez, filtering,
time_integration, units, scales, primitive_equations_states) on torch tensors,
computing through hand-written gfx950 HIP kernels in libdinosaur_amd.so.
"""
__version__ = '0.1.0'
