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
# hip_system reads the grid's coordinate system (circular_symmetric! /
# spherical_symmetric!), the species activity (enable_species! /
# enable_boundary_species!) and the assembly variant of the System.
#
# The Physics callbacks are Julia closures the device cannot call.  The registry terms
# below (LinearFlux, SinhReaction, ..., BoundaryConditions) are ordinary callables with
# the reference's callback signatures, so a Physics built from them runs unchanged on
# the reference's CPU path; `hip_physics` and `hip_bc_table` recognise them and map the
# system onto the device registry.  A Physics with any other callback maps to
# `nothing`: such a system stays on the reference CPU path (`hip_enabled`).
module VoronoiFVMHIP

using LinearAlgebra, SparseArrays
using VoronoiFVM, ExtendableGrids
const libvfvm = joinpath(@__DIR__, "..", "libvfvm_hip.so")

# status codes (src/vfvm_logging_exceptions.jl:6-29)
const VFVM_OK, VFVM_ERR_CONVERGENCE, VFVM_ERR_ASSEMBLY, VFVM_ERR_LINEAR = 0, 1, 2, 3

# ---- registry terms: callables with VoronoiFVM's callback signatures ----
# (include/vfvm_hip.h lists the formulas and the examples they restate)
abstract type RegistryTerm end

struct LinearFlux <: RegistryTerm  # Example105_NonlinearPoisson1D.jl:41-44, per species D
    D::Vector{Float64}
end
function (t::LinearFlux)(f, u, edge, data)
    for i in eachindex(t.D)
        f[i] = t.D[i] * (u[i, 1] - u[i, 2])
    end
    return nothing
end

struct PowerFlux <: RegistryTerm  # Example106_NonlinearDiffusion1D.jl:57-60
    m::Int
    D::Float64
end
(t::PowerFlux)(f, u, edge, data) = (f[1] = t.D * (u[1, 1]^t.m - u[1, 2]^t.m); nothing)

struct NonlinearDiffusionFlux <: RegistryTerm  # D(ū) (u_k - u_l), D(u) = d0 + d2 u^2
    d0::Float64
    d2::Float64
end
function (t::NonlinearDiffusionFlux)(f, u, edge, data)
    ū = (u[1, 1] + u[1, 2]) / 2
    f[1] = (t.d2 * ū^2 + t.d0) * (u[1, 1] - u[1, 2])
    return nothing
end

struct DriftDiffusionFlux <: RegistryTerm  # psi, n, p; Example160_UnipolarDriftDiffusion1D.jl:43-50
    λ2::Float64
    zn::Float64
    zp::Float64
end
function (t::DriftDiffusionFlux)(f, u, edge, data)
    dψ = u[1, 1] - u[1, 2]
    f[1] = t.λ2 * dψ
    bp, bm = VoronoiFVM.fbernoulli_pm(t.zn * dψ)
    f[2] = bm * u[2, 1] - bp * u[2, 2]
    bp, bm = VoronoiFVM.fbernoulli_pm(t.zp * dψ)
    f[3] = bm * u[3, 1] - bp * u[3, 2]
    return nothing
end

struct SinhReaction <: RegistryTerm  # Example105:56-59
    k::Float64
end
(t::SinhReaction)(f, u, node, data) = (f[1] = t.k * (exp(u[1]) - exp(-u[1])); nothing)

struct PowerReaction <: RegistryTerm  # Example207_NonlinearPoisson2D.jl:34-36
    m::Int
    k::Float64
end
(t::PowerReaction)(f, u, node, data) = (f[1] = t.k * u[1]^t.m; nothing)

struct ExchangeReaction <: RegistryTerm  # Example210:25-30
    k::Float64
end
(t::ExchangeReaction)(f, u, node, data) = (f[1] = t.k * (u[1] - u[2]); f[2] = t.k * (u[2] - u[1]); nothing)

struct DriftDiffusionReaction <: RegistryTerm  # n - p; (np - ni^2) / (n + p + 2 ni)
    ni::Float64
end
function (t::DriftDiffusionReaction)(f, u, node, data)
    f[1] = u[2] - u[3]
    R = (u[2] * u[3] - t.ni^2) / (u[2] + u[3] + 2 * t.ni)
    f[2] = R
    f[3] = R
    return nothing
end

struct IdentityStorage <: RegistryTerm  # Example106:63-66
end
(t::IdentityStorage)(f, u, node, data) = (f .= u; nothing)

# bcondition entries: (:dirichlet | :robin, species, region, value, factor); value(λ) =
# value + value_lambda * embedparam
struct BCEntry
    kind::Symbol
    species::Int
    region::Int
    value::Float64
    value_lambda::Float64
    factor::Float64
end
struct BoundaryConditions <: RegistryTerm
    entries::Vector{BCEntry}
end
function (t::BoundaryConditions)(f, u, bnode, data)
    for e in t.entries
        v = e.value + e.value_lambda * VoronoiFVM.embedparam(bnode)
        if e.kind == :dirichlet
            VoronoiFVM.boundary_dirichlet!(f, u, bnode; species = e.species, region = e.region, value = v)
        else
            VoronoiFVM.boundary_robin!(f, u, bnode; species = e.species, region = e.region, factor = e.factor,
                                       value = v)
        end
    end
    return nothing
end

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

# registry ids (include/vfvm_hip.h)
const FLUX_IDS = Dict(LinearFlux => 1, PowerFlux => 2, NonlinearDiffusionFlux => 3, DriftDiffusionFlux => 4)
const REACT_IDS = Dict(SinhReaction => 1, PowerReaction => 2, ExchangeReaction => 3, DriftDiffusionReaction => 4)
isnofunc(fn) = fn === VoronoiFVM.nofunc

"""
    hip_physics(physics::VoronoiFVM.Physics, nspec) -> Union{PhysicsT, Nothing}

The registry entry of a Physics whose flux, reaction and storage are registry terms (or
absent) and whose other callbacks are absent or a `BoundaryConditions`; `nothing`
otherwise: that system stays on the reference's CPU path.
"""
function hip_physics(physics, nspec)
    p = zeros(16)
    flux, react, stor = 0, 0, 0
    fl, re, st = physics.flux, physics.reaction, physics.storage
    if fl isa LinearFlux
        length(fl.D) == nspec || return nothing
        p[1:nspec] .= fl.D
    elseif fl isa PowerFlux
        p[1] = fl.D; p[5] = fl.m
    elseif fl isa NonlinearDiffusionFlux
        p[1] = fl.d0; p[4] = fl.d2
    elseif fl isa DriftDiffusionFlux
        nspec == 3 || return nothing
        p[8] = fl.λ2; p[9] = fl.zn; p[10] = fl.zp
    elseif !isnofunc(fl)
        return nothing
    end
    isnofunc(fl) || (flux = FLUX_IDS[typeof(fl)])
    if re isa SinhReaction
        p[6] = re.k
    elseif re isa PowerReaction
        p[6] = re.k; p[7] = re.m
    elseif re isa ExchangeReaction
        p[6] = re.k
    elseif re isa DriftDiffusionReaction
        p[11] = re.ni^2; p[12] = re.ni
    elseif !isnofunc(re)
        return nothing
    end
    isnofunc(re) || (react = REACT_IDS[typeof(re)])
    if st isa IdentityStorage
        stor = 1
    elseif !isnofunc(st) && st !== VoronoiFVM.default_storage
        return nothing
    end
    bc = physics.breaction
    (isnofunc(bc) || bc isa BoundaryConditions) || return nothing
    for fn in (physics.bflux, physics.bstorage, physics.bsource, physics.edgereaction, physics.generic)
        isnofunc(fn) || return nothing
    end
    src = physics.source
    isnofunc(src) || return nothing  # per-node sources go through vfvm_system_set_source
    return PhysicsT(flux, react, stor, 0, Tuple(p))
end

"""
    hip_bc_table(sys::VoronoiFVM.System, nbreg) -> Vector{BCT}

The (boundary region, species) table of a system: the entries of a BoundaryConditions
breaction, then the legacy conditions of boundary_dirichlet! / boundary_robin!
(sys.boundary_factors / boundary_values, src/vfvm_system.jl:854-933).
"""
function hip_bc_table(sys, nbreg)
    nspec = VoronoiFVM.num_species(sys)
    tab = [BCT(0, 0, 0.0, 0.0, 0.0) for _ in 1:(nbreg * nspec)]
    bc = sys.physics.breaction
    if bc isa BoundaryConditions
        for e in bc.entries
            k = (e.region - 1) * nspec + e.species
            tab[k] = BCT(e.kind == :dirichlet ? 1 : 3, 0, e.value, e.value_lambda, e.factor)
        end
    end
    for ireg in 1:nbreg, ispec in 1:nspec
        fac, val = sys.boundary_factors[ispec, ireg], sys.boundary_values[ispec, ireg]
        (fac == 0 && val == 0) && continue
        k = (ireg - 1) * nspec + ispec
        tab[k] = fac ≈ VoronoiFVM.Dirichlet(Float64) ? BCT(2, 0, val, 0.0, 0.0) : BCT(3, 0, val, 0.0, fac)
    end
    return tab
end

struct SpeciesT
    kind::Int32
    bregions::UInt32
end

"""
    hip_species(sys) -> Union{Vector{SpeciesT}, Nothing}

Species activity: BULK (enabled in every cell region), BOUNDARY (with its boundary
regions) or NONE; `nothing` for a bulk species restricted to some of several cell regions.
"""
function hip_species(sys)
    nspec = VoronoiFVM.num_species(sys)
    ncreg = size(sys.region_species, 2)
    out = SpeciesT[]
    for ispec in 1:nspec
        if VoronoiFVM.is_boundary_species(sys, ispec)
            bits = UInt32(0)
            for r in 1:size(sys.bregion_species, 2)
                sys.bregion_species[ispec, r] > 0 && (bits |= UInt32(1) << (r - 1))
            end
            push!(out, SpeciesT(1, bits))
        elseif VoronoiFVM.is_bulk_species(sys, ispec)
            all(r -> sys.region_species[ispec, r] > 0, 1:ncreg) || return nothing
            push!(out, SpeciesT(0, 0))
        else
            push!(out, SpeciesT(2, 0))
        end
    end
    return out
end

"""
    hip_enabled(sys) -> Bool

Whether a VoronoiFVM.System maps onto the device registry; if not, the reference's own
CPU path runs it.
"""
hip_enabled(sys) = hip_physics(sys.physics, VoronoiFVM.num_species(sys)) !== nothing && hip_species(sys) !== nothing &&
                   coordsys_code(sys.grid) !== nothing

"""
    hip_system(ctx, sys::VoronoiFVM.System) -> handle

Uploads the system's ExtendableGrid (1-based Int64 CellNodes / BFaceNodes, as stored)
and builds the device system: what System(grid; ...) + _complete! do for the path, in
the grid's coordinate system, with the system's species activity and assembly variant.
Throws for a system outside the registry (check `hip_enabled` first).
"""
function hip_system(ctx, sys)
    grid = sys.grid
    nspec = VoronoiFVM.num_species(sys)
    physics = hip_physics(sys.physics, nspec)
    species = hip_species(sys)
    cs = coordsys_code(grid)
    (physics === nothing || species === nothing || cs === nothing) &&
        error("system outside the device registry: run it on the reference CPU path")
    coord = grid[Coordinates]
    cn = Int64.(grid[CellNodes]); cr = Int64.(grid[CellRegions])
    bn = Int64.(grid[BFaceNodes]); br = Int64.(grid[BFaceRegions])
    nbreg = maximum(br)
    bc = hip_bc_table(sys, nbreg)
    assembly = sys.assembly_type == :cellwise ? Cint(1) : Cint(0)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    rc = ccall((:vfvm_system_create_ex, libvfvm), Cint,
               (Ptr{Cvoid}, Cint, Int64, Ptr{Float64}, Int64, Ptr{Int64}, Ptr{Int64}, Int64, Ptr{Int64}, Ptr{Int64},
                Cint, Ref{PhysicsT}, Cint, Ptr{BCT}, Cint, Ptr{SpeciesT}, Ptr{Cvoid}, Cint, Ptr{Ptr{Cvoid}}),
               ctx, size(coord, 1), size(coord, 2), coord, size(cn, 2), cn, cr, size(bn, 2), bn, br,
               nspec, physics, nbreg, bc, cs, species, C_NULL, assembly, h)
    check(rc, ctx)
    return h[]
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
    coordsys_code(grid) -> Union{Cint, Nothing}

The VFVM_COORD_* code of grid[CoordinateSystem] (circular_symmetric! /
spherical_symmetric!) for the grid's dimension; `nothing` for a coordinate system the
device form factors do not cover (that system stays on the reference CPU path).
"""
function coordsys_code(grid)
    cs = grid[CoordinateSystem]
    dim = size(grid[Coordinates], 1)
    cs <: Cartesian1D && dim == 1 && return Cint(0)
    cs <: Cartesian2D && dim == 2 && return Cint(0)
    cs <: Cartesian3D && dim == 3 && return Cint(0)
    cs <: Polar1D && dim == 1 && return Cint(1)
    cs <: Spherical1D && dim == 1 && return Cint(2)
    cs <: Cylindrical2D && dim == 2 && return Cint(3)
    return nothing
end

struct StepControlT
    d::Float64; d_min::Float64; d_max::Float64; d_grow::Float64; d_decrease::Float64
    du_opt::Float64; du_max_factor::Float64
    num_final_steps::Int32; force_first_step::Int32
    handle_exceptions::Int32; pad::Int32
end

"""
    step_control(control::VoronoiFVM.SolverControl, transient) -> StepControlT

The Δt (time stepping) or Δp (embedding) fields of a SolverControl
(src/vfvm_solvercontrol.jl:138-211).
"""
function step_control(c, transient)
    transient ?
        StepControlT(c.Δt, c.Δt_min, c.Δt_max, c.Δt_grow, c.Δt_decrease, c.Δu_opt, c.Δu_max_factor,
                     c.num_final_steps, c.force_first_step, c.handle_exceptions, 0) :
        StepControlT(c.Δp, c.Δp_min, c.Δp_max, c.Δp_grow, c.Δp_decrease, c.Δu_opt, c.Δu_max_factor,
                     c.num_final_steps, c.force_first_step, c.handle_exceptions, 0)
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
    if hist.converged == 0  # handle_exceptions: returned before the end (src/vfvm_solver.jl:413-440)
        @warn unsafe_string(ccall((:vfvm_last_error, libvfvm), Cstring, (Ptr{Cvoid},), ctx))
    end
    n = min(nsteps[] + 1, max_out)
    return ts[1:n], [us[:, k] for k in 1:n], hist
end

"""
    hip_set_generic_matrix!(ctx, sys, G::SparseMatrixCSC)

Registry form of `Physics(; generic = f)` for a linear f (src/vfvm_assembly.jl:648-677,
examples/Example405_GenericOperator.jl): every assembly adds G*U to F and G to the
Jacobian. G's entries must lie in the system's pattern (hip_pattern).
"""
function hip_set_generic_matrix!(ctx, sys, G::SparseMatrixCSC)
    I, J, V = findnz(G)
    rc = ccall((:vfvm_system_set_generic_matrix, libvfvm), Cint,
               (Ptr{Cvoid}, Int64, Ptr{Int64}, Ptr{Int64}, Ptr{Float64}),
               sys, length(V), Vector{Int64}(I), Vector{Int64}(J), Vector{Float64}(V))
    check(rc, ctx)
    return sys
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
