"""MI355X (gfx950) implementation of the Strang-splitting path of
GeneralizedGrossPitaevskii.jl: ``solve(prob, StrangSplitting(), tspan; dt, nsaves)``.

The arithmetic runs in libgpe_hip.so (HIP kernels for gfx950, C ABI in
include/gpe_hip.h); this package is the host-side mirror of the reference's
user API.  The directory name is not a Python identifier, so load it with
``__graft_entry__.load_package()`` (registered as ``gpe_hip_amd``).
"""
from .problem import (GrossPitaevskiiProblem, PhiloxRNG, SMatrix, StrangSplitting, StrangSplittingIterator,
                      SVector, device_empty, direct_grid, init, pump_times, reciprocal_grid,
                      resolve_fixed_timestepping, solve, solve_sweep, steady_state, SweepIterator,
                      wigner_density_g2)
from ._lib import GpeError, GpeUnsupported, load as load_library

__all__ = [
    "GrossPitaevskiiProblem", "StrangSplitting", "StrangSplittingIterator", "solve", "init", "SVector", "SMatrix",
    "PhiloxRNG", "solve_sweep", "steady_state", "SweepIterator", "wigner_density_g2", "direct_grid", "reciprocal_grid", "resolve_fixed_timestepping", "pump_times", "device_empty",
    "GpeError", "GpeUnsupported", "load_library",
]
