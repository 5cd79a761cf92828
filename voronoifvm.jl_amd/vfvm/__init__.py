"""MI355X (gfx950) assemble-and-solve path of the Voronoi finite-volume method.

Host-side mirror of VoronoiFVM.jl's System / Physics / solve surface; the work
is done by libvfvm_hip.so (HIP kernels behind the C ABI in include/vfvm_hip.h).
"""
from ._capi import AssemblyError, ConvergenceError, DeviceArray, LinearSolverError, VFVMError
from .control import SolverControl, fixed_timesteps
from .grids import Grid, circular_symmetric, linspace, read_sg, simplexgrid, spherical_symmetric
from .physics import (BoundaryFlux, BoundaryReaction, BoundaryReactionField, BoundarySource, BoundaryStorage,
                      Dirichlet, GenericMatrix, DriftDiffusionFlux, DriftDiffusionReaction, ExchangeReaction, IdentityStorage,
                      LinearFlux, Neumann, NonlinearDiffusionFlux, Physics, PowerFlux, PowerReaction, ReactionField,
                      Robin, SinhReaction)
from .system import (Batch, NewtonSolverHistory, System, TransientSolution, boundary_dirichlet, boundary_neumann,
                     boundary_robin, dofs, enable_boundary_species, enable_species, eval_jacobian, eval_rhs, evaluate_residual_and_jacobian,
                     ilu0_factor, initialize_dirichlet, integrate, linsolve, mass_matrix, solve, unknowns)

from . import amg
from .impedance import (ImpedanceSystem, assemble_dudp, assemble_dudp_at, impedance, integrate_testfunction, measurement_derivatives,
                        parameter_derivative, testfunction)

__all__ = [n for n in dir() if not n.startswith("_")]
