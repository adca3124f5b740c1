"""femhip -- MI355X (gfx950) backend for the FEM time-stepping hot path of
tvumcc/fea-visualizer: GPU assembly of P1 stiffness / mass / advection into
CSR and persistent implicit time stepping (heat, wave, advection-diffusion,
Gray-Scott, Newmark-beta) behind the reference's Solver interface.
"""
from . import _lib
from ._lib import FemError
from .meshes import (boundary_from_edges, icosphere, jittered_square, mesh_stats, read_obj,
                     structured_square)
from .solver import (Adjoint, AdvectionDiffusionParameters, Batch, BoundaryCondition, Equation, FEMContext,
                     HeatParameters, HIPSolver, NewmarkParameters, Operator, ReactionDiffusionParameters,
                     SolverOptions, WaveParameters)

__all__ = [
    "FemError", "FEMContext", "HIPSolver", "Batch", "Operator", "Adjoint", "Equation", "BoundaryCondition", "SolverOptions",
    "HeatParameters", "AdvectionDiffusionParameters", "WaveParameters", "ReactionDiffusionParameters",
    "NewmarkParameters", "structured_square", "jittered_square", "icosphere", "read_obj",
    "boundary_from_edges", "mesh_stats",
]
