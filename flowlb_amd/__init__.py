"""MI355X-native core of FlowLB's D2Q9 lattice-Boltzmann solver.

The compute path is libflbcore.so (hand-written HIP for gfx950 behind the C ABI
of include/flb_core.h); this package is its host-side binding.
"""
from . import core
from .core import (FLUID, GAS, INTERFACE, WALL, MOVING_WALL, INLET, OUTLET, GAS_INTERFACE,
                   MONOFLUID, FREE_SURFACE, SRT, TRT, Deck, FlbError, Solver,
                   build_channel_flags, build_culvert, default_options,
                   init_free_surface_fields, seed_free_surface, units)

__all__ = ["core", "FLUID", "GAS", "INTERFACE", "WALL", "MOVING_WALL", "INLET", "OUTLET",
           "GAS_INTERFACE", "MONOFLUID", "FREE_SURFACE", "SRT", "TRT", "Deck", "FlbError", "Solver", "build_channel_flags", "build_culvert",
           "default_options", "init_free_surface_fields", "seed_free_surface", "units"]
