"""octree_hip: MI355X-native Barnes-Hut gravity solver (HIP kernels behind a C ABI).

Python surface over libbh_hip.so (include/bh_hip.h). Importing the package
does not touch the GPU; creating an `OctreeHip` does.
"""
from ._abi import EXPORTS, LIB_PATH, Params, Timings, default_params, lib  # noqa: F401
from .solver import (BhHipError, OctreeHip, hernquist, hernquist_merger,  # noqa: F401
                     normalize_solver, plummer, sort_pairs, uniform_cube,
                     zeldovich_lattice)

__all__ = ["OctreeHip", "BhHipError", "Params", "Timings", "default_params", "lib", "plummer",
           "hernquist", "hernquist_merger", "uniform_cube", "zeldovich_lattice", "sort_pairs",
           "normalize_solver",
           "EXPORTS", "LIB_PATH"]
