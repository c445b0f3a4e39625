"""moldyn_amd: MI355X (gfx950) engine for moldyn's per-step MD path.

The compute path is libmoldyn_hip.so (hand-written HIP kernels behind the C
ABI in include/moldyn_hip.h); this package is its Python front end.
"""
from .engine import (Engine, barostat, lattice, lj_potential_and_force, mb_velocities,
                     new_lennard_jones, thermostat)
from ._lib import MdError, load as load_library

__all__ = ["Engine", "thermostat", "barostat", "lattice", "mb_velocities",
           "lj_potential_and_force", "new_lennard_jones", "MdError", "load_library"]
