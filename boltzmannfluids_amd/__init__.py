"""MI355X (gfx950) lattice Boltzmann solver.

The solver is the HIP library libbfamd.so behind the C ABI in include/bf/bf.h;
`capi` binds it for Python callers.  Importing this package does not touch
the GPU.
"""
from .capi import (BFError, Solver, bits_per_boundary, load_library, marching_cubes,  # noqa: F401
                   marching_cubes_table, pack_boundaries, unpack_boundaries, velocity_table, voxelize_mesh)

__all__ = ["BFError", "Solver", "load_library", "velocity_table", "bits_per_boundary",
           "pack_boundaries", "unpack_boundaries", "marching_cubes", "marching_cubes_table", "voxelize_mesh"]
