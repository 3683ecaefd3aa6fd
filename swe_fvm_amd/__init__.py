"""MI355X (gfx950) explicit finite-volume step of SWE_FVM.

The numerics run in hand-written HIP kernels behind the C ABI of
include/swe_fvm.h (libswe_fvm_amd.so, built in-tree by swe_fvm_amd.build).
This package mirrors the reference's operator interface on top of it.
"""
from .core import (DomainError, DeviceError, Domain, Fluxer, Fluxes, SolverError, SpaceDisc,
                   Solvers, StructTriangMesh, TimeDisc, Topology, Wavespeeds, jittered_mesh,
                   read_msh, OPT_HEAD_GRADIENT, OPT_HEAD_PARTWET2, OPT_FRICTION, OPT_DRAIN_LENGTH,
                   OPT_EXACT_BALANCE, SOLID_WALL, FREE_FLOW, PERIODIC, periodic_pairs, stack_domains,
                   morton_order, reorder_cells)

__all__ = [
    "DomainError", "DeviceError", "Domain", "Fluxer", "Fluxes", "SolverError", "SpaceDisc",
    "Solvers", "StructTriangMesh", "TimeDisc", "Topology", "Wavespeeds", "jittered_mesh",
    "read_msh", "OPT_HEAD_GRADIENT", "OPT_HEAD_PARTWET2", "OPT_FRICTION", "OPT_DRAIN_LENGTH",
    "OPT_EXACT_BALANCE", "SOLID_WALL", "FREE_FLOW", "PERIODIC", "periodic_pairs", "stack_domains",
    "morton_order", "reorder_cells",
]
