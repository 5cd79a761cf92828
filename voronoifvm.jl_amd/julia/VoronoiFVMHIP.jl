# VoronoiFVMHIP.jl -- the Julia side of the drop-in boundary: ccall bindings of
# libvfvm_hip.so (include/vfvm_hip.h) and the two replacements a VoronoiFVM
# maintainer plugs in.  It cannot run in this repository's environment (no
# Julia); it documents, as code, how the reference reaches the library.
#
#   eval_and_assemble (src/vfvm_assembly.jl:520-643)  -> hip_eval_and_assemble!
#   solve_step!       (src/vfvm_solver.jl:13-222)     -> hip_solve_step!
#   _solve_linear!    (src/vfvm_linsolve.jl:6-61)     -> hip_solve_linear!
#   solve_transient!  (src/vfvm_solver.jl:267-537)    -> hip_solve_transient!
#   update_grid!      (src/vfvm_system.jl:607-612)    -> hip_update_grid!
#   eval_rhs!, eval_jacobian!, mass_matrix
#                     (src/vfvm_diffeq_interface.jl)  -> hip_eval_rhs!, hip_eval_jacobian!, hip_mass_matrix
#   ImpedanceSystem, impedance, measurement_derivative
#                     (src/vfvm_impedance.jl)         -> hip_impedance, hip_measurement_derivatives
# A grid with circular_symmetric! / spherical_symmetric! goes to vfvm_system_create_cs.
#
# The Physics callbacks are Julia closures the device cannot call; a system is
# mapped to a registry entry with `hip_physics` (flux/reaction/storage kinds and
# parameters) and its boundary conditions to a table with `hip_bc_table`.
module VoronoiFVMHIP

using LinearAlgebra, SparseArrays
const libvfvm = joinpath(@__DIR__, "..", "libvfvm_hip.so")

# status codes (src/vfvm_logging_exceptions.jl:6-29)
const VFVM_OK, VFVM_ERR_CONVERGENCE, VFVM_ERR_ASSEMBLY, VFVM_ERR_LINEAR = 0, 1, 2, 3

struct PhysicsT
    flux::Int32
    reaction::Int32
    storage::Int32
    pad::Int32
    p::NTuple{16, Float64}
end

struct BCT
    kind::Int32
    pad::Int32
    value::Float64
    value_lambda::Float64
    factor::Float64
end

mutable struct ControlT
    abstol::Float64; reltol::Float64; tol_round::Float64; tol_mono::Float64
    damp_initial::Float64; damp_growth::Float64
    maxiters::Int32; max_round::Int32; updatecontrol::Int32; is_linear::Int32
    reltol_linear::Float64; abstol_linear::Float64
    maxiters_linear::Int32; krylov::Int32; precon::Int32; precision::Int32
    gmres_restart::Int32; max_refinements::Int32; reuse_precon::Int32; check_every::Int32
    inner_reltol::Float64
    ControlT() = new()
end

mutable struct HistoryT
    niter::Int32; nlin::Int32; nlu::Int32; converged::Int32
    time::Float64; tasm::Float64; tlinsolve::Float64
    last_update_norm::Float64; residual_norm::Float64
    updatenorm::NTuple{64, Float64}
    fp64_fallbacks::Int32; pad::Int32
    HistoryT() = new()
end

function check(rc, ctx)
    rc == VFVM_OK && return nothing
    msg = unsafe_string(ccall((:vfvm_last_error, libvfvm), Cstring, (Ptr{Cvoid},), ctx))
    rc == VFVM_ERR_CONVERGENCE && throw(VoronoiFVM.ConvergenceError())
    rc == VFVM_ERR_ASSEMBLY && throw(VoronoiFVM.AssemblyError())
    rc == VFVM_ERR_LINEAR && throw(VoronoiFVM.LinearSolverError())
    error("vfvm_hip: $msg")
end

function context(device = 0)
    ctx = Ref{Ptr{Cvoid}}(C_NULL)
    rc = ccall((:vfvm_ctx_create, libvfvm), Cint, (Cint, Ptr{Ptr{Cvoid}}), device, ctx)
    rc == 0 || error(unsafe_string(ccall((:vfvm_last_error, libvfvm), Cstring, (Ptr{Cvoid},), C_NULL)))
    return ctx[]
end

"""
    hip_system(ctx, grid, physics::PhysicsT, bc::Vector{BCT}, nspec)

Uploads an ExtendableGrid (1-based Int64 CellNodes / BFaceNodes, as stored) and
builds the device system: what System(grid; ...) + _complete! do for the path.
"""
function hip_system(ctx, grid, physics::PhysicsT, bc::Vector{BCT}, nspec)
    coord = grid[Coordinates]
    cn = Int64.(grid[CellNodes]); cr = Int64.(grid[CellRegions])
    bn = Int64.(grid[BFaceNodes]); br = Int64.(grid[BFaceRegions])
    nbreg = maximum(br)
    sys = Ref{Ptr{Cvoid}}(C_NULL)
    rc = ccall((:vfvm_system_create, libvfvm), Cint,
               (Ptr{Cvoid}, Cint, Int64, Ptr{Float64}, Int64, Ptr{Int64}, Ptr{Int64}, Int64, Ptr{Int64}, Ptr{Int64},
                Cint, Ref{PhysicsT}, Cint, Ptr{BCT}, Ptr{Ptr{Cvoid}}),
               ctx, size(coord, 1), size(coord, 2), coord, size(cn, 2), cn, cr, size(bn, 2), bn, br,
               nspec, physics, nbreg, bc, sys)
    check(rc, ctx)
    return sys[]
end

"""
    hip_pattern(sys) -> SparseMatrixCSC

The Jacobian pattern, identical to SparseMatrixCSC(state.matrix) of the edgewise
reference assembly (1-based CSC, rows sorted).
"""
function hip_pattern(sys)
    ndof, nnz, ne = Ref{Int64}(0), Ref{Int64}(0), Ref{Int64}(0)
    ccall((:vfvm_system_sizes, libvfvm), Cint, (Ptr{Cvoid}, Ref{Int64}, Ref{Int64}, Ref{Int64}), sys, ndof, nnz, ne)
    colptr = zeros(Int64, ndof[] + 1)
    rowval = zeros(Int64, nnz[])
    ccall((:vfvm_pattern, libvfvm), Cint, (Ptr{Cvoid}, Ptr{Int64}, Ptr{Int64}), sys, colptr, rowval)
    return SparseMatrixCSC(ndof[], ndof[], colptr, rowval, zeros(nnz[]))
end

"""
    hip_eval_and_assemble!(ctx, sys, U, UOld, F, matrix::SparseMatrixCSC, time, tstep, λ)

Replacement of eval_and_assemble (src/vfvm_assembly.jl:520-643): residual into F,
Jacobian values into nonzeros(matrix) (pattern from hip_pattern).
"""
function hip_eval_and_assemble!(ctx, sys, U, UOld, F, matrix::SparseMatrixCSC, time, tstep, λ)
    rc = ccall((:vfvm_assemble, libvfvm), Cint,
               (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Float64, Float64, Float64, Ptr{Float64}, Ptr{Float64}),
               sys, U, UOld, time, tstep, λ, F, nonzeros(matrix))
    check(rc, ctx)
    return matrix
end

"""
    hip_solve_linear!(ctx, sys, u, b, control::ControlT)

Replacement of _solve_linear! (src/vfvm_linsolve.jl:6-61) on the last assembled
Jacobian: BiCGStab/GMRES with Jacobi, block-Jacobi or ILU(0), FP64 or FP32 with
FP64 iterative refinement.
"""
function hip_solve_linear!(ctx, sys, u, b, control::ControlT)
    stats = zeros(UInt8, 32)
    rc = ccall((:vfvm_linsolve, libvfvm), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ref{ControlT}, Ptr{UInt8}),
               sys, b, u, control, stats)
    check(rc, ctx)
    return u
end

"""
    hip_solve_step!(ctx, sys, solution, oldsol, control::ControlT, time, tstep, λ)

Replacement of solve_step! (src/vfvm_solver.jl:13-222): Dirichlet initialisation,
damped Newton, update-norm / round-off / monotonicity tests, all on the device.
"""
function hip_solve_step!(ctx, sys, solution, oldsol, control::ControlT, time, tstep, λ)
    hist = HistoryT()
    rc = ccall((:vfvm_newton, libvfvm), Cint,
               (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Float64, Float64, Float64, Ref{ControlT}, Ref{HistoryT}),
               sys, solution, oldsol, time, tstep, λ, control, hist)
    check(rc, ctx)
    return solution, hist
end

"""
    hip_update_grid!(ctx, sys, grid)

Replacement of update_grid!(system; grid) (src/vfvm_system.jl:607-612, 690-755) for a
grid with the system's cells and boundary faces and new coordinates: the form factors
are recomputed on the device.  Throws if the new coordinates would need a different
sparsity pattern (the system keeps its previous geometry then).
"""
function hip_update_grid!(ctx, sys, grid)
    coord = Float64.(grid[Coordinates])
    rc = ccall((:vfvm_update_grid, libvfvm), Cint, (Ptr{Cvoid}, Ptr{Float64}), sys, coord)
    check(rc, ctx)
    return sys
end

"""
    hip_eval_rhs!(ctx, sys, du, u, t)
    hip_eval_jacobian!(ctx, sys, J::SparseMatrixCSC, u, t)
    hip_mass_matrix(ctx, sys) -> Diagonal

The ODE/DAE form (src/vfvm_diffeq_interface.jl:31-96) for an ODEFunction: du = -F(u)
and J = -dF/du assembled with tstep = Inf; the mass matrix is diagonal for every
registry physics.  J has the pattern of hip_pattern(sys).
"""
function hip_eval_rhs!(ctx, sys, du, u, t)
    rc = ccall((:vfvm_eval_rhs, libvfvm), Cint, (Ptr{Cvoid}, Ptr{Float64}, Float64, Ptr{Float64}), sys, u, t, du)
    check(rc, ctx)
    return du
end

function hip_eval_jacobian!(ctx, sys, J::SparseMatrixCSC, u, t)
    rc = ccall((:vfvm_eval_jacobian, libvfvm), Cint, (Ptr{Cvoid}, Ptr{Float64}, Float64, Ptr{Float64}), sys, u, t,
               nonzeros(J))
    check(rc, ctx)
    return J
end

function hip_mass_matrix(ctx, sys, ndof)
    d = zeros(ndof)
    rc = ccall((:vfvm_mass_matrix, libvfvm), Cint, (Ptr{Cvoid}, Ptr{Float64}), sys, d)
    check(rc, ctx)
    return Diagonal(d)
end

"""
    hip_impedance(ctx, sys, U0, ω, dmeas_stdy, dmeas_tran; excited_spec, excited_bc, λ0)

ImpedanceSystem + impedance (src/vfvm_impedance.jl) for a whole frequency sweep at
once: boundary-excited (excited_spec, excited_bc given) or excited through the
parameter λ (ImpedanceSystem(sys, U0; λ0)).  Returns the complex impedances.
"""
function hip_impedance(ctx, sys, U0, ω::Vector{Float64}, dmeas_stdy, dmeas_tran;
                       excited_spec = 0, excited_bc = 0, λ0 = 0.0, reltol = 1.0e-12, maxiters = 20000)
    isys = Ref{Ptr{Cvoid}}(C_NULL)
    rc = excited_bc > 0 ?
        ccall((:vfvm_impedance_create, libvfvm), Cint, (Ptr{Cvoid}, Ptr{Float64}, Cint, Cint, Ref{Ptr{Cvoid}}),
              sys, U0, excited_spec, excited_bc, isys) :
        ccall((:vfvm_impedance_create_param, libvfvm), Cint, (Ptr{Cvoid}, Ptr{Float64}, Float64, Ref{Ptr{Cvoid}}),
              sys, U0, λ0, isys)
    check(rc, ctx)
    Z = zeros(ComplexF64, length(ω))
    rc = ccall((:vfvm_impedance_sweep, libvfvm), Cint,
               (Ptr{Cvoid}, Int64, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Float64, Int32, Ptr{ComplexF64},
                Ptr{ComplexF64}, Ptr{Int32}),
               isys[], length(ω), ω, dmeas_stdy, dmeas_tran, reltol, maxiters, C_NULL, Z, C_NULL)
    ccall((:vfvm_impedance_destroy, libvfvm), Cint, (Ptr{Cvoid},), isys[])
    check(rc, ctx)
    return Z
end

"""
    hip_measurement_derivatives(ctx, sys, U0, T, spec) -> (dmeas_stdy, dmeas_tran)

measurement_derivative of -integrate_stdy / -integrate_tran with test function T.
"""
function hip_measurement_derivatives(ctx, sys, U0, T, spec, ndof)
    ds, dt = zeros(ndof), zeros(ndof)
    rc = ccall((:vfvm_measurement_derivatives, libvfvm), Cint,
               (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Cint, Ptr{Float64}, Ptr{Float64}), sys, U0, T, spec, ds, dt)
    check(rc, ctx)
    return ds, dt
end

"""
    hip_integrate(ctx, sys, T, U; Uold = nothing, tstep = Inf) -> Vector (one value per species)

integrate(system, T, U) / integrate(system, T, U, Uold, tstep)
(src/vfvm_testfunctions.jl:110-210): test-function integrals, e.g. terminal currents.
"""
function hip_integrate(ctx, sys, T, U, nspec; Uold = nothing, tstep = Inf)
    out = zeros(nspec)
    rc = ccall((:vfvm_integrate_testfunction, libvfvm), Cint,
               (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Float64, Ptr{Float64}, Ptr{Float64}),
               sys, U, Uold === nothing ? C_NULL : Uold, tstep, T, out)
    check(rc, ctx)
    return out
end

"""
    coordsys_code(grid) -> Cint

The VFVM_COORD_* code of grid[CoordinateSystem] (circular_symmetric! /
spherical_symmetric!); pass it to vfvm_system_create_cs instead of calling
vfvm_system_create.
"""
function coordsys_code(grid)
    cs = grid[CoordinateSystem]
    cs <: Polar1D && return Cint(1)
    cs <: Spherical1D && return Cint(2)
    cs <: Cylindrical2D && return Cint(3)
    return Cint(0)
end

struct StepControlT
    d::Float64; d_min::Float64; d_max::Float64; d_grow::Float64; d_decrease::Float64
    du_opt::Float64; du_max_factor::Float64
    num_final_steps::Int32; force_first_step::Int32
end

"""
    step_control(control::VoronoiFVM.SolverControl, transient) -> StepControlT

The Δt (time stepping) or Δp (embedding) fields of a SolverControl
(src/vfvm_solvercontrol.jl:138-211).
"""
function step_control(c, transient)
    transient ?
        StepControlT(c.Δt, c.Δt_min, c.Δt_max, c.Δt_grow, c.Δt_decrease, c.Δu_opt, c.Δu_max_factor,
                     c.num_final_steps, c.force_first_step) :
        StepControlT(c.Δp, c.Δp_min, c.Δp_max, c.Δp_grow, c.Δp_decrease, c.Δu_opt, c.Δu_max_factor,
                     c.num_final_steps, c.force_first_step)
end

"""
    hip_solve_transient!(ctx, sys, inival, lambdas, control::ControlT, steps::StepControlT; transient = true)

Replacement of solve_transient! (src/vfvm_solver.jl:267-537): returns the accepted
times (or parameters) and solutions, the first entry being the start.
"""
function hip_solve_transient!(ctx, sys, inival, lambdas, control::ControlT, steps::StepControlT; transient = true,
                              max_out = 100_000)
    ndof = length(inival)
    U = similar(inival)
    ts = zeros(max_out)
    us = zeros(ndof, max_out)
    nsteps = Ref{Int64}(0)
    hist = HistoryT()
    rc = ccall((:vfvm_solve_transient, libvfvm), Cint,
               (Ptr{Cvoid}, Ptr{Float64}, Int64, Ptr{Float64}, Cint, Ref{ControlT}, Ref{StepControlT}, Ptr{Float64},
                Int64, Ptr{Float64}, Ptr{Float64}, Ref{Int64}, Ref{HistoryT}),
               sys, inival, length(lambdas), collect(Float64, lambdas), transient, control, steps, U, max_out, ts, us,
               nsteps, hist)
    check(rc, ctx)
    n = min(nsteps[] + 1, max_out)
    return ts[1:n], [us[:, k] for k in 1:n], hist
end

"""
    hip_batch(ctx, systems) -> handle

Packs independent systems (the one-SystemState-per-thread pattern of
examples/Example440_ParallelState.jl:36-45) into one batch.
"""
function hip_batch(ctx, systems::Vector{Ptr{Cvoid}})
    b = Ref{Ptr{Cvoid}}(C_NULL)
    rc = ccall((:vfvm_batch_create, libvfvm), Cint, (Ptr{Cvoid}, Int64, Ptr{Ptr{Cvoid}}, Ref{Ptr{Cvoid}}),
               ctx, length(systems), systems, b)
    check(rc, ctx)
    return b[]
end

"""
    hip_batch_solve_device!(ctx, batch, U, Uold, control; time = 0.0, tstep = Inf, λ = 0.0)

solve_step! for every member with start values and solutions in device memory:
`U` and `Uold` are device pointers (e.g. `pointer(A)` of an AMDGPU.jl ROCArray, its
stream synchronised) to the members' dofs packed in member order.
"""
function hip_batch_solve_device!(ctx, batch, U::Ptr{Float64}, Uold::Ptr{Float64}, control::ControlT;
                                 time = 0.0, tstep = Inf, λ = 0.0)
    # per-member histories are not requested here (NULL); the return code reports a
    # member that failed to converge (ConvergenceError)
    rc = ccall((:vfvm_batch_newton_device, libvfvm), Cint,
               (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Float64, Float64, Float64, Ref{ControlT}, Ptr{Cvoid}),
               batch, U, Uold, time, tstep, λ, control, C_NULL)
    check(rc, ctx)
    return nothing
end

function default_control()
    c = ControlT()
    ccall((:vfvm_default_control, libvfvm), Cvoid, (Ref{ControlT},), c)
    return c
end

end # module
