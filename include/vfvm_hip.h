/*
 * vfvm_hip.h -- C ABI of the MI355X (gfx950) assemble-and-solve library.
 *
 * The library replaces VoronoiFVM.jl's hot path: eval_and_assemble's node,
 * edge and boundary-node loops, the damped Newton loop of solve_step! and the
 * linear solve of _solve_linear!.  A Julia host keeps the VoronoiFVM.System /
 * Physics / solve(sys; inival, control) surface and reaches these entry points
 * with `ccall` (see the Julia shim in voronoifvm.jl_amd/julia).  Everything here
 * is plain C: pointers, sizes and status codes; no torch types.
 *
 * Conventions
 *  - Arrays coming from Julia keep Julia's layout: coordinates dim x nnodes
 *    column-major, cellnodes (dim+1) x ncells and bfacenodes dim x nbfaces as
 *    1-based int64 (grid[Coordinates], grid[CellNodes], grid[BFaceNodes]).
 *  - Solution-like arrays (U, Uold, F, source) are nspec x nnodes column-major,
 *    i.e. the reference's dofs(U) order: dof(K, i) = (K-1)*nspec + i
 *    (src/vfvm_system.jl:808-811, src/vfvm_densesolution.jl:49,68).
 *  - Matrices cross the boundary as CSC with 1-based int64 indices and rows
 *    sorted within columns, what SparseMatrixCSC(state.matrix) gives for the
 *    system's assembly variant (src/vfvm_linsolve.jl:3-4; VFVM_ASSEMBLY_EDGEWISE
 *    or _CELLWISE).  Inside, the device keeps CSR + a slot map.
 *  - Host pointers are copied in and out; no pointer is retained after a call.
 *    Device buffers belong to the handle and are freed by its destroy call.
 *  - Every call returns a status: 0 OK, or one of the VFVM_ERR_* codes that
 *    mirror the reference's exceptions (src/vfvm_logging_exceptions.jl:6-29).
 *  - Handles are independent and each owns one HIP stream: concurrent calls on
 *    different handles are allowed, calls on one handle must be serialised
 *    (the reference serialises through gridaccesslock, src/vfvm_system.jl:118).
 */
#ifndef VFVM_HIP_H
#define VFVM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (src/vfvm_logging_exceptions.jl:6-29) ---- */
#define VFVM_OK 0
#define VFVM_ERR_CONVERGENCE 1 /* ConvergenceError, src/vfvm_solver.jl:196-198 */
#define VFVM_ERR_ASSEMBLY 2    /* AssemblyError / "trying to assemble NaN", src/vfvm_assembly.jl:10-12 */
#define VFVM_ERR_LINEAR 3      /* LinearSolverError, src/vfvm_linsolve.jl:50-58 */
#define VFVM_ERR_ARG 4         /* invalid argument (shape, unsupported physics) */
#define VFVM_ERR_HIP 5         /* HIP runtime failure or no device */

/* ---- physics registry: the fixed set of device callbacks ----
 * Flux (src/vfvm_physics.jl:67-182 `flux`), reaction, storage.  Parameters in p[]:
 *   p[0..2] D_i (LINEAR per species; POWER and NLDIFF use p[0])
 *   p[3] d2 (NLDIFF: D(u) = p[0] + p[3]*u^2 at the edge mean), p[4] m (POWER exponent)
 *   p[5] k (SINH/POW/EXCHANGE factor), p[6] POW exponent
 *   p[7] lambda^2, p[8] z_n, p[9] z_p, p[10] n_i^2, p[11] n_i (DD)
 *   p[12] != 0: the node source is scaled by the embedding parameter lambda
 *         (a source callback returning embedparam(node) * f, src/vfvm_geometryitems.jl:13) */
#define VFVM_FLUX_NONE 0
#define VFVM_FLUX_LINEAR 1 /* D_i (u_k - u_l)          Example105_NonlinearPoisson1D.jl:41-44 */
#define VFVM_FLUX_POWER 2  /* D (u_k^m - u_l^m)        Example106_NonlinearDiffusion1D.jl:57-60 */
#define VFVM_FLUX_NLDIFF 3 /* (d0 + d2 ubar^2)(u_k-u_l) src/vfvm_physics.jl:478-480 */
#define VFVM_FLUX_DD 4     /* psi: lambda^2 dpsi; n,p: Scharfetter-Gummel, Example160:43-50 */
#define VFVM_REACT_NONE 0
#define VFVM_REACT_SINH 1     /* k (e^u - e^-u)      Example105:56-59 */
#define VFVM_REACT_POW 2      /* k u^m               Example207:34-36 */
#define VFVM_REACT_EXCHANGE 3 /* k(u1-u2), k(u2-u1)  Example210:25-30 */
#define VFVM_REACT_DD 4       /* n - p ; R = (np - ni^2)/(n + p + 2 ni) */
#define VFVM_STOR_NONE 0
#define VFVM_STOR_IDENTITY 1 /* storage u_i          Example106:63-66 */

typedef struct {
  int32_t flux, reaction, storage, pad;
  double p[16];
} vfvm_physics_t;

/* ---- boundary conditions per (boundary region, species) ----
 * DIRICHLET: breaction with boundary_dirichlet!, penalty 1e30 * fac (src/vfvm_physics.jl:487-494)
 * DIRICHLET_LEGACY: boundary_dirichlet!(sys, ...), penalty 1e30 without fac (src/vfvm_assembly.jl:370-379)
 * ROBIN: boundary_robin!/boundary_neumann! (factor*u - value), src/vfvm_physics.jl:533,552
 * value(lambda) = value + value_lambda * lambda (embedding / bias continuation). */
#define VFVM_BC_NONE 0
#define VFVM_BC_DIRICHLET 1
#define VFVM_BC_DIRICHLET_LEGACY 2
#define VFVM_BC_ROBIN 3

typedef struct {
  int32_t kind, pad;
  double value, value_lambda, factor;
} vfvm_bc_t;

/* ---- linear solver selection (SolverControl.method_linear, src/vfvm_solvercontrol.jl:106) ---- */
#define VFVM_KRYLOV_BICGSTAB 0
#define VFVM_KRYLOV_GMRES 1
#define VFVM_PRECON_JACOBI 0       /* point Jacobi (row scaling) */
#define VFVM_PRECON_BLOCKJACOBI 1  /* nspec x nspec node blocks */
/* ILU(0) of the rows in a multicolour order: rows of one colour are independent, so a
 * triangular solve takes one launch per colour (a handful); a reordering, hence another
 * preconditioner than ILUZero's */
#define VFVM_PRECON_ILU0 2
/* smoothed aggregation AMG, one V(1,1) cycle per application (scalar systems):
 * KrylovJL_BICGSTAB(precs = SmoothedAggregationPreconBuilder()),
 * examples/DevEx003_Solvers.jl:149-156 */
#define VFVM_PRECON_AMG 3
/* ILU(0) in the natural dof order, ILUZero.jl's ilu0 (Project.toml:67; used as
 * KrylovJL_BICGSTAB(precs = ILUZeroPreconBuilder()), examples/DevEx003_Solvers.jl:90-120):
 * factorisation and triangular solves level-scheduled over the dependency levels of the
 * natural order (one launch per level: about 2N levels on N x N grids, 3N on N^3) */
#define VFVM_PRECON_ILU0_NATURAL 4
#define VFVM_PREC_FP64 0
#define VFVM_PREC_FP32_REFINE 1    /* FP32 Krylov + preconditioner inside FP64 iterative refinement */

/* SolverControl subset on the path (src/vfvm_solvercontrol.jl:30-131) plus the
 * device solver choice.  vfvm_default_control() fills the reference defaults. */
typedef struct {
  double abstol, reltol, tol_round, tol_mono, damp_initial, damp_growth;
  int32_t maxiters, max_round, updatecontrol, is_linear;
  double reltol_linear, abstol_linear;
  int32_t maxiters_linear, krylov, precon, precision;
  /* reuse_precon = !factorize_every_newtonstep (src/vfvm_solvercontrol.jl:123-131): the
   * ILU(0) factor of a Newton solve's first step serves its later steps (reuse_precs,
   * src/vfvm_solver.jl:99); default 1 */
  int32_t gmres_restart, max_refinements, reuse_precon, check_every;
  double inner_reltol; /* FP32 inner solve tolerance per refinement sweep */
} vfvm_control_t;

/* NewtonSolverHistory subset (src/vfvm_history.jl:10-37) */
typedef struct {
  int32_t niter, nlin, nlu, converged;
  double time, tasm, tlinsolve;
  double last_update_norm;
  double residual_norm; /* ||F||_inf over non-penalty rows at the returned solution */
  double updatenorm[64];
  /* linear solves run or finished in FP64 although FP32 refinement was requested: an
   * FP32 refinement that fails hands over to FP64, and the rest of the call (Newton
   * solve, or time / parameter sweep) then solves in FP64 directly */
  int32_t fp64_fallbacks, pad;
} vfvm_history_t;

typedef struct {
  int32_t iters, refinements, converged;
  int32_t fp64_fallback; /* 1 when an FP32 refinement step failed and the solve finished in FP64 */
  double resnorm0, resnorm;
} vfvm_linstats_t;

typedef struct vfvm_ctx vfvm_ctx;
typedef struct vfvm_sys vfvm_sys;

void vfvm_default_control(vfvm_control_t *c);

/* Context: one device, one stream.  No counterpart in the reference (it has no
 * device); it plays the role of the process-wide state SystemState relies on. */
int vfvm_ctx_create(int device, vfvm_ctx **out);
int vfvm_ctx_destroy(vfvm_ctx *ctx);
/* last error message of this context (or of the last failed create when ctx is NULL) */
const char *vfvm_last_error(vfvm_ctx *ctx);
/* Device memory on the context's device, for callers that keep inputs and results in HBM
 * between calls (the *_device entry points).  No counterpart in the reference.  copy:
 * kind 0 host to device, 1 device to host, 2 device to device; returns when done. */
int vfvm_device_alloc(vfvm_ctx *ctx, int64_t bytes, void **ptr);
int vfvm_device_free(vfvm_ctx *ctx, void *ptr);
int vfvm_device_copy(vfvm_ctx *ctx, void *dst, const void *src, int64_t bytes, int kind);

/* System creation: grid upload + discretisation setup (dof map, Voronoi
 * factors, edgewise aggregation, boundary node factors, CSR pattern and slot
 * map).  Replaces System(grid; physics...) + _complete!/update_grid_edgewise!
 * (src/vfvm_system.jl:215-261, 530-596, 690-755; src/vfvm_state.jl:99-157).
 * cellregions/bfaceregions are 1-based region numbers (int64). */
int vfvm_system_create(vfvm_ctx *ctx, int dim, int64_t nnodes, const double *coord, int64_t ncells,
                       const int64_t *cellnodes, const int64_t *cellregions, int64_t nbfaces,
                       const int64_t *bfacenodes, const int64_t *bfaceregions, int nspec,
                       const vfvm_physics_t *physics, int nbregions, const vfvm_bc_t *bc,
                       vfvm_sys **out);
/* The same for a grid in a symmetric coordinate system (ExtendableGrids'
 * circular_symmetric! / spherical_symmetric!, used by
 * examples/Example203_CoordinateSystems.jl): the form factors of
 * src/vfvm_formfactors.jl:25-61 (Polar1D, Spherical1D: dim 1, the coordinate is the
 * radius) and 102-161, 276-291 (Cylindrical2D: dim 2, coordinates r, z). */
#define VFVM_COORD_CARTESIAN 0
#define VFVM_COORD_POLAR1D 1
#define VFVM_COORD_SPHERICAL1D 2
#define VFVM_COORD_CYLINDRICAL2D 3
int vfvm_system_create_cs(vfvm_ctx *ctx, int dim, int64_t nnodes, const double *coord, int64_t ncells,
                          const int64_t *cellnodes, const int64_t *cellregions, int64_t nbfaces,
                          const int64_t *bfacenodes, const int64_t *bfaceregions, int nspec,
                          const vfvm_physics_t *physics, int nbregions, const vfvm_bc_t *bc, int coordsys,
                          vfvm_sys **out);
/* Species activity (enable_species! / enable_boundary_species!, src/vfvm_system.jl:433-515):
 * BULK species are enabled in every cell region; a BOUNDARY species lives on the nodes of
 * the boundary faces of the regions whose bits (region r: bit r-1) are set in bregions; a
 * NONE species is enabled nowhere.  Dofs of a species not active at a node get identity
 * rows, F = U - Uold, and are set to 0 at the start of each Newton solve
 * (_eval_and_assemble_inactive_species / _initialize_inactive_dof!, src/vfvm_system.jl:
 * 1010-1042, dense unknown storage).  A bulk species enabled in only some of several cell
 * regions is not expressible here. */
#define VFVM_SPECIES_BULK 0
#define VFVM_SPECIES_BOUNDARY 1
#define VFVM_SPECIES_NONE 2
typedef struct {
  int32_t kind;
  uint32_t bregions;
} vfvm_species_t;

/* Boundary physics of one boundary region: the registry's breaction, bstorage, bsource and
 * bflux callbacks (src/vfvm_physics.jl:67-182; assembled by assemble_bnodes /
 * assemble_bedges, src/vfvm_assembly.jl:318-510):
 *   breaction  f_i = sum_j react[i][j] u_j (+ bcoef[i,K] u_i where `fields` is set)
 *              (examples/Example220_NonlinearPoisson2D_BoundarySpecies.jl:25-33,
 *              Example311_HeatEquation_BoundaryDiffusion.jl:66-75)
 *   bsource    bsrc[i,K] where `fields` is set; the item adds fac (breaction - bsource)
 *   bstorage   f_i = u_i for the species bits i of `storage` (Example220:33-37)
 *   bflux      f_i = bflux[i] (u_i,K - u_i,L) on the region's boundary edges, with the
 *              boundary-face edge factors (examples/Example230_BoundaryFlux.jl:125-130)
 * Terms act on the species active at the node (the edge: at both nodes).  These add to the
 * boundary-condition table of vfvm_bc_t. */
typedef struct {
  double react[3][3];
  double bflux[3];
  int32_t storage, fields;
} vfvm_bphysics_t;

/* Assembly variant of System(grid; assembly), src/vfvm_system.jl:215-227: EDGEWISE sums the
 * cell factors per edge first and stores no Jacobian entry for an edge whose summed factor
 * is zero (update_grid_edgewise!, lines 690-755); CELLWISE (the reference's default)
 * stores an entry for every cell edge (_addnz tests the flux derivative, not the factor,
 * src/vfvm_assembly.jl:20-28), so its pattern is that of SparseMatrixCSC(state.matrix) of
 * a cellwise system; its values equal the cellwise sums to rounding. */
#define VFVM_ASSEMBLY_EDGEWISE 0
#define VFVM_ASSEMBLY_CELLWISE 1

/* System creation with species activity (species: nspec entries, NULL = all BULK), boundary
 * physics (bphys: nbregions entries, NULL = none) and the assembly variant; the rest as
 * vfvm_system_create_cs.  Systems with boundary species, inactive dofs or boundary physics
 * run an extra boundary pass after the bulk assembly; they cannot join a batch. */
int vfvm_system_create_ex(vfvm_ctx *ctx, int dim, int64_t nnodes, const double *coord, int64_t ncells,
                          const int64_t *cellnodes, const int64_t *cellregions, int64_t nbfaces,
                          const int64_t *bfacenodes, const int64_t *bfaceregions, int nspec,
                          const vfvm_physics_t *physics, int nbregions, const vfvm_bc_t *bc, int coordsys,
                          const vfvm_species_t *species, const vfvm_bphysics_t *bphys, int assembly,
                          vfvm_sys **out);
/* Per-node fields, nspec x nnodes each (NULL clears): rcoef adds rcoef[i,K] u_i to the bulk
 * reaction (a reaction callback reading per-node data, examples/Example230_BoundaryFlux.jl:
 * 63-66); bsrc and bcoef are the boundary source and boundary reaction coefficient of the
 * regions whose vfvm_bphysics_t.fields is set. */
int vfvm_system_set_fields(vfvm_sys *sys, const double *rcoef, const double *bsrc, const double *bcoef);
/* node_dof activity: bit i of mask[K] is set if species i is active at node K (nnodes bytes) */
int vfvm_node_species(vfvm_sys *sys, uint8_t *mask);
int vfvm_system_destroy(vfvm_sys *sys);
/* per-node source term, nspec x nnodes (Physics(source = array), src/vfvm_system.jl:312-319); NULL clears */
int vfvm_system_set_source(vfvm_sys *sys, const double *src);
/* replace the boundary-condition table (nbregions x nspec entries, region-major) */
int vfvm_system_set_bc(vfvm_sys *sys, const vfvm_bc_t *bc);
/* sizes: ndof, nnz of the Jacobian pattern, number of stored edges */
int vfvm_system_sizes(vfvm_sys *sys, int64_t *ndof, int64_t *nnz, int64_t *nedges);
/* node volumes (sum of node factors), for integrate() and tests (src/vfvm_postprocess.jl:397-412) */
int vfvm_nodevolumes(vfvm_sys *sys, double *vol);
/* update_grid!(system; grid) for a grid with new coordinates and the same cells and
 * boundary faces (src/vfvm_system.jl:607-612, update_grid_edgewise! 690-755): coord is
 * dim x nnodes (as at creation).  The form factors, node volumes and boundary face
 * factors are recomputed on the device, bit for bit those of a system created with
 * these coordinates.  VFVM_ERR_ARG (geometry left unchanged) if an edge whose factor was
 * exactly zero at creation, so has no Jacobian entry, gets a nonzero factor.  The
 * compact factor table is rebuilt on the host only by vfvm_set_compact_graph(sys, 1);
 * until then the assembly reads the FP64 factors. */
int vfvm_update_grid(vfvm_sys *sys, const double *coord);
/* The same with the coordinates in device memory (a moving mesh computed on the GPU):
 * dcoord is a device pointer, dim x nnodes. */
int vfvm_update_grid_device(vfvm_sys *sys, const double *dcoord);
/* Where vfvm_system_create builds the node graph, edge factors and node volumes
 * (the cells part of _complete! / update_grid_edgewise!, src/vfvm_system.jl:690-755):
 * 0 on the host, 1 on the GPU, 2 (default) on the GPU for grids of at least 2^18
 * cell-node incidences and on the host below.  All give bit for bit the same system;
 * the boundary items, the dof pattern and the solver structures are built on the host
 * either way.  No counterpart in the reference (a choice of where setup runs). */
int vfvm_ctx_set_setup(vfvm_ctx *ctx, int mode);

/* Jacobian pattern as CSC, 1-based, rows sorted: what SparseMatrixCSC(state.matrix)
 * exposes for the system's assembly variant (src/vfvm_linsolve.jl:3-4,
 * src/vfvm_state.jl:136; vfvm_system_create and _cs give the edgewise pattern,
 * vfvm_system_create_ex either).  colptr has ndof+1 entries. */
int vfvm_pattern(vfvm_sys *sys, int64_t *colptr, int64_t *rowval);

/* Dirichlet initialisation of U (in place): _initialize_dirichlet!, src/vfvm_system.jl:947-1003 */
int vfvm_initialize_dirichlet(vfvm_sys *sys, double *U, double time, double lambda);

/* One assembly: eval_and_assemble(system, U, UOld, F, matrix, ...), src/vfvm_assembly.jl:520-643.
 * tstep = +Inf means stationary.  F is nspec x nnodes; nzval (may be NULL) receives the
 * Jacobian values in CSC order of vfvm_pattern. */
int vfvm_assemble(vfvm_sys *sys, const double *U, const double *Uold, double time, double tstep,
                  double lambda, double *F, double *nzval);

/* ODE/DAE form of the discrete system (src/vfvm_diffeq_interface.jl): with
 * M du/dt = rhs(u), rhs(u) = -F(u) and jacobian = -dF/du, F assembled at U = Uold = u,
 * tstep = Inf, embedding 0 (_eval_res_jac!, lines 10-23).
 *   vfvm_eval_rhs       eval_rhs!(du, u, state, t)       lines 31-36   (u, du: ndof)
 *   vfvm_eval_jacobian  eval_jacobian!(J, u, state, t)   lines 44-50   (values in the
 *                       order of vfvm_pattern, i.e. nonzeros of the CSC matrix)
 *   vfvm_mass_matrix    mass_matrix(state)               lines 58-96   (the diagonal, ndof
 *                       entries: storage derivative times node volume; the registry's
 *                       storage terms are local per species, so the matrix is diagonal)
 * The _device forms take and return device memory: ddu has ndof entries, dnzval_csr
 * the Jacobian values in the row-major order of vfvm_pattern_csr. */
int vfvm_eval_rhs(vfvm_sys *sys, const double *u, double time, double *du);
int vfvm_eval_jacobian(vfvm_sys *sys, const double *u, double time, double *nzval);
int vfvm_mass_matrix(vfvm_sys *sys, double *diag);
int vfvm_eval_rhs_device(vfvm_sys *sys, const double *du, double time, double *ddu);
int vfvm_eval_jacobian_device(vfvm_sys *sys, const double *du, double time, double *dnzval_csr);

/* Small-signal impedance (src/vfvm_impedance.jl).
 * vfvm_impedance_create: ImpedanceSystem(system, U0, excited_spec, excited_bc), lines
 *   159-186 with 49-152: J(U0) assembled with tstep = 1e30 and Uold = U0, the storage
 *   derivative, and the excitation of species excited_spec (1-based) on boundary region
 *   excited_bc: 1e30 per boundary item for a legacy Dirichlet condition
 *   (boundary_dirichlet!(sys, ...)), factor times the item's measure for a Robin one,
 *   nothing otherwise (as the reference, which reads only the legacy boundary factors).
 * vfvm_measurement_derivatives: measurement_derivative (lines 235-252) of
 *   U -> -integrate_stdy(sys, T, U)[spec] and U -> -integrate_tran(sys, T, U)[spec]
 *   (src/vfvm_testfunctions.jl:398-420) at U0 for a test function T (nnodes values):
 *   dstdy, dtran have ndof entries.
 * vfvm_impedance_sweep: solve!(UZ, isys, w) and impedance(isys, w, U0, dstdy, dtran)
 *   (lines 205-232, 268-288) for nomega frequencies at once on the GPU (batched BiCGStab,
 *   relative residual of the row-scaled system <= reltol).  UZ (nullable): nomega x ndof
 *   complex values, interleaved (re, im); Z (nullable, needs dstdy and dtran): nomega
 *   complex impedances 1 / (dstdy.UZ + i w dtran.UZ); iters (nullable): iterations per
 *   frequency.  VFVM_ERR_LINEAR if a frequency does not converge. */
/* integrate(system, T, U) for a stationary solution (src/vfvm_testfunctions.jl:110-130,
 * integrate_stdy 398-410; Uold NULL, tstep = Inf) and integrate(system, T, U, Uold,
 * tstep) for an implicit-Euler step (lines 194-210): per species sum_K T_K F_K of the
 * flux, reaction, source and storage terms (no boundary conditions), i.e.
 * I_flux + I_react - I_src + (I_stor - I_oldstor) / tstep; nspec values.  With T from
 * testfunction(factory, bc0, bc1) this is the flux through the bc1 boundary (the
 * currents of examples/Example161_BipolarDriftDiffusionCurrent.jl). */
int vfvm_integrate_testfunction(vfvm_sys *sys, const double *U, const double *Uold, double tstep, const double *T,
                                double *result);
/* The same at the system parameter lambda (integrate(system, T, U; params), with the
 * registry's one parameter standing for the params; the form above evaluates at 1, the
 * reference's embedding parameter of integrate). */
int vfvm_integrate_testfunction_at(vfvm_sys *sys, const double *U, const double *Uold, double tstep,
                                   const double *T, double lambda, double *result);
typedef struct vfvm_impedance vfvm_impedance;
int vfvm_impedance_create(vfvm_sys *sys, const double *U0, int excited_spec, int excited_bc,
                          vfvm_impedance **out);
/* ImpedanceSystem(system, U0; lambda0) for a system whose parameter is lambda (lines
 * 49-112): the excitation is F = -dF/dlambda (vfvm_assemble_dudp). */
int vfvm_impedance_create_param(vfvm_sys *sys, const double *U0, double lambda0, vfvm_impedance **out);
int vfvm_impedance_destroy(vfvm_impedance *imp);
int vfvm_measurement_derivatives(vfvm_sys *sys, const double *U0, const double *T, int spec,
                                 double *dstdy, double *dtran);
/* measurement_derivative at the system parameter lambda (the form above: lambda = 0), for
 * systems whose coefficients depend on lambda */
int vfvm_measurement_derivatives_at(vfvm_sys *sys, const double *U0, const double *T, int spec, double lambda,
                                    double *dstdy, double *dtran);
int vfvm_impedance_sweep(vfvm_impedance *imp, int64_t nomega, const double *omega, const double *dstdy,
                         const double *dtran, double reltol, int32_t maxiters, double *UZ, double *Z,
                         int32_t *iters);

/* Parameter derivatives.  The registry's system parameter is lambda: it enters the
 * boundary values value + value_lambda * lambda and, with p[12] set, the source
 * lambda * f -- the registry's form of `nparams = 1, params = [lambda]`.
 * vfvm_assemble_dudp: state.dudp[1] = dF/dlambda (src/vfvm_assembly.jl:116-120,
 *   194-198, 398-405), ndof values, of the boundary values and the embedded source (exact;
 *   independent of U and lambda); with lambda-dependent coefficients use
 *   vfvm_assemble_dudp_at.
 * vfvm_parameter_derivative: the sensitivity du/dlambda = -J(U)^-1 dF/dlambda of a
 *   stationary solution U at lambda (what Example430_ParameterDerivativesStationary.jl
 *   obtains by dual numbers through the solve), with the linear solver of `control`;
 *   dF/dlambda includes the lambda-dependent coefficients. */
int vfvm_assemble_dudp(vfvm_sys *sys, double *dudp);
/* Registry coefficients that depend on the parameter lambda (the reference's params reaching
 * the flux and reaction callbacks, src/vfvm_assembly.jl:116-120, 194-198;
 * examples/Example430_ParameterDerivativesStationary.jl:142-160): slot j of the physics
 * block is p[j] + dp[j] lambda in every assembly at lambda (dp: 16 values, NULL clears;
 * the integer exponents p[4], p[6] cannot depend on lambda). */
int vfvm_system_set_param_dependence(vfvm_sys *sys, const double *dp);
/* state.dudp[1] = dF/dlambda at U and lambda: the boundary values and embedded source as
 * vfvm_assemble_dudp, plus the flux and reaction terms through the lambda-dependent
 * coefficients, evaluated with dual-number parameters.  with_bc = 0 leaves out the
 * boundary conditions (the derivative of integrate(system, T, U; params), the
 * test-function integral, src/vfvm_testfunctions.jl:110-130). */
int vfvm_assemble_dudp_at(vfvm_sys *sys, const double *U, double lambda, int with_bc, double *dudp);
int vfvm_parameter_derivative(vfvm_sys *sys, const double *U, double lambda, const vfvm_control_t *control,
                              double *dudl);

/* Linear solve A x = b on the currently assembled Jacobian (after vfvm_assemble or inside
 * vfvm_newton): _solve_linear!, src/vfvm_linsolve.jl:6-61. */
int vfvm_linsolve(vfvm_sys *sys, const double *b, double *x, const vfvm_control_t *control,
                  vfvm_linstats_t *stats);

/* ---- smoothed aggregation AMG (VFVM_PRECON_AMG) ----
 * The hierarchy is AlgebraicMultigrid.jl's MultiLevel (levels with A, P, R = P^T), the
 * preconditioner of examples/DevEx003_Solvers.jl:149-156; the reference reads none of it
 * back, these calls let tests check it.
 * Standard aggregation of a CSR pattern (0-based, diagonal ignored), host only: agg[i] in
 * [0, *naggregates). */
int vfvm_amg_aggregate(int64_t n, const int32_t *rowptr, const int32_t *colind, int32_t *agg,
                       int64_t *naggregates);
/* Levels of the system's current hierarchy (built by a solve or vfvm_linsolve with
 * VFVM_PRECON_AMG) and the rows of each (rows: up to 16 entries, may be NULL). */
int vfvm_amg_levels(vfvm_sys *sys, int *nlevels, int64_t *rows);
/* Level l: sizes[0..3] = rows, nonzeros of A, coarse rows (0 on the coarsest level),
 * nonzeros of P.  A (0-based CSR; level 0: diag(rowscale) J) and P (rows x coarse rows),
 * the aggregates and omega = (4/3) / rho; any output pointer may be NULL. */
int vfvm_amg_level(vfvm_sys *sys, int level, int64_t *sizes, int32_t *arp, int32_t *aci, double *aval,
                   int32_t *prp, int32_t *pci, double *pval, int32_t *agg, double *omega);

/* Generic operator (Physics(; generic = f), _eval_and_assemble_generic_operator,
 * src/vfvm_assembly.jl:648-677; examples/Example405_GenericOperator.jl) in the registry
 * form of a constant matrix G over the dofs: every assembly adds G U to F and G to the
 * Jacobian, after the bulk and boundary terms.  COO triplets with 1-based int64 dof
 * indices (repeated entries add up); each entry must lie in the system's sparsity pattern
 * (VFVM_ERR_ARG otherwise, the operator unchanged): the edge couplings are in it for the
 * species a flux couples, so a system whose only coupling is G takes VFVM_FLUX_LINEAR
 * with zero coefficients.  nnz = 0 removes the operator.  A
 * nonlinear generic operator has no registry form.  Systems with a generic operator
 * cannot join a batch. */
int vfvm_system_set_generic_matrix(vfvm_sys *sys, int64_t nnz, const int64_t *rows, const int64_t *cols,
                                   const double *vals);

/* The ILU(0) factor of the last linear solve with an ILU(0) preconditioner, in the
 * system's CSR order (lu[k] for the k-th stored entry of the dof pattern): the strictly
 * lower part holds L (unit diagonal not stored), the rest U, of the Jacobian without the
 * row scaling the solver applies; lower/upper refer to the factor's ordering (natural for
 * VFVM_PRECON_ILU0_NATURAL).  Inspection for tests, as ILUZero.ILU0Precon's l/u fields
 * (VFVM_ERR_ARG if no factor exists). */
int vfvm_ilu0_factor(vfvm_sys *sys, double *lu);

/* Damped Newton: solve_step!, src/vfvm_solver.jl:13-222.  On entry U is ignored and Uold is
 * the initial value / old time step; on return U holds the solution.  Assembly, linear
 * solves, damping and the norms stay on the device; hist mirrors NewtonSolverHistory. */
int vfvm_newton(vfvm_sys *sys, double *U, const double *Uold, double time, double tstep,
                double lambda, const vfvm_control_t *control, vfvm_history_t *hist);

/* Implicit-Euler time stepping with fixed steps or embedding with fixed parameter steps:
 * solve_transient!, src/vfvm_solver.jl:267-537 with Δt_min = Δt_max (fixed_timesteps!,
 * src/vfvm_solvercontrol.jl:320-327).  transient = 1: lambdas are times and each step is
 * solve_step!(tstep = Δt); transient = 0: lambdas are embedding values (tstep = Inf).
 * Solutions stay on the device between steps; U_final receives the last one.  hist (may be
 * NULL) accumulates over the steps: niter and nlin are totals, residual_norm is the largest
 * per-step ||F||_inf over non-penalty rows, converged is 1 only if every step converged. */
int vfvm_solve_fixed_steps(vfvm_sys *sys, const double *inival, int64_t nlambdas,
                           const double *lambdas, int transient, const vfvm_control_t *control,
                           double *U_final, vfvm_history_t *hist);

/* Step-size control of solve_transient! (src/vfvm_solver.jl:267-537); for embedding
 * the Δp fields of SolverControl, for time stepping the Δt fields
 * (src/vfvm_solvercontrol.jl:138-211).  vfvm_default_stepcontrol gives the reference's
 * defaults (transient != 0: Δt = 0.1, Δt_min = 1e-3, Δt_max = 1, grow 1.2; embedding:
 * Δp = 1, Δp_min = 1e-3, Δp_max = 1, grow 1; decrease 0.5, Δu_opt = 0.1,
 * Δu_max_factor = 1.2, num_final_steps = 5).  fixed_timesteps! corresponds to
 * d = d_min = d_max and du_opt = DBL_MAX. */
typedef struct {
  double d, d_min, d_max, d_grow, d_decrease;
  double du_opt, du_max_factor;
  int32_t num_final_steps, force_first_step;
  /* SolverControl.handle_exceptions (src/vfvm_solvercontrol.jl:205-216): 1 retries a
   * failing Newton step with a smaller step, and returns the steps done so far, instead of
   * an error, when the smallest step fails (src/vfvm_solver.jl:391-452); default 0 */
  int32_t handle_exceptions, pad;
} vfvm_stepcontrol_t;
void vfvm_default_stepcontrol(int transient, vfvm_stepcontrol_t *sc);

/* solve_transient! with its step-size control, on the device Newton loop: Δu =
 * ||U - Uold||_inf; a step with Δu > du_max_factor * du_opt is retried with the step
 * times d_decrease down to d_min, where the call fails (VFVM_ERR_CONVERGENCE) unless the
 * first step is forced; a failing Newton solve fails the call.  With handle_exceptions a
 * failing Newton solve is retried like a rejected step, and where d_min fails too (or a
 * first step that is forced fails its Newton solve) the call stops and returns VFVM_OK
 * with the steps done so far and hist->converged = 0.  Accepted steps grow as
 * min(d_max, d d_grow, d du_opt / Δu, ...) with the
 * num_final_steps rule.  Outputs: U_final; up to max_out entries (lambda, solution) in
 * out_lambdas / out_U (max_out x ndof; both nullable): entry 0 is the start (the initial
 * value for time stepping, the solution at lambdas[0] for embedding), then one entry per
 * accepted step, as the reference's TransientSolution; *nsteps = accepted steps. */
int vfvm_solve_transient(vfvm_sys *sys, const double *inival, int64_t nlambdas, const double *lambdas, int transient,
                         const vfvm_control_t *control, const vfvm_stepcontrol_t *steps, double *U_final,
                         int64_t max_out, double *out_lambdas, double *out_U, int64_t *nsteps,
                         vfvm_history_t *hist);

/* ||F||_inf over non-penalty rows for U (residual-only assembly) */
int vfvm_residual_norm(vfvm_sys *sys, const double *U, const double *Uold, double time,
                       double tstep, double lambda, double *norm);

/* The same norm with U and Uold in device memory (dUold NULL: dU), written to dnorm (one
 * double in device memory), and NO allocation: no hipMalloc, hipFree or allocating library
 * call.  No counterpart in the reference: it is for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid it).
 * The residual it reduces lives in a static device array of VFVM_NOALLOC_MAX_DOF doubles
 * (32 MiB per GPU, held until the process ends); a system with more dofs gets VFVM_ERR_ARG.
 * The call is enqueued on the handle's stream and returns without synchronising.  Every
 * call on one GPU shares that array, so calls that overlap on one GPU race. */
#define VFVM_NOALLOC_MAX_DOF ((int64_t)1 << 22)
int vfvm_residual_norm_device_noalloc(vfvm_sys *sys, const double *dU, const double *dUold, double time,
                                      double tstep, double lambda, double *dnorm);

/* Autotune the lanes-per-node / lanes-per-row group size of the assembly and SpMV
 * kernels for this system (project-written: times each legal size on the device).
 * Assembly groups: 2..32 lanes per node; 1 = one node per lane with the block's
 * adjacency and Jacobian rows staged in LDS; 0 = one node per lane with only the
 * Jacobian rows staged. */
int vfvm_autotune(vfvm_sys *sys, int32_t *asm_group, int32_t *spmv_group);

/* Set the group sizes directly (assembly 0, 1, 2, 4, 8, 16 or 32; CSR residual 2..32). */
int vfvm_set_groups(vfvm_sys *sys, int32_t asm_group, int32_t spmv_group);
/* Column indices the SELL-64 SpMV reads: 16 = offsets from the row in 16 bits (the
 * default wherever every row's columns lie within 32767 rows of it), 32 = absolute
 * 32-bit columns.  Same arithmetic either way; 16 bits move 2 bytes less per entry. */
int vfvm_set_sell_index(vfvm_sys *sys, int bits);
/* Node graph the assembly reads: 1 (default) = compact where it fits (neighbours as
 * 16-bit offsets from the node, edge factors as 16-bit indices into a table of the
 * graph's distinct values, bit-exact), 0 = int32 neighbours and FP64 factors per entry.
 * After vfvm_update_grid, enabling it rebuilds the factor table (a host pass). */
int vfvm_set_compact_graph(vfvm_sys *sys, int enable);

/* Benchmark helper (no counterpart in the reference; it has no device): average device
 * time in milliseconds of `reps` back-to-back launches of one hot kernel on the handle's
 * stream, measured with HIP events recorded on that stream, at the current device state
 * (U, Uold, Jacobian).  kernel: 0 Jacobian assembly (transient form, tstep = 1),
 * 1 residual-only assembly, 2 FP64 SpMV with the row-scaled Jacobian, 3 FP32 SpMV. */
#define VFVM_KERNEL_ASSEMBLE 0
#define VFVM_KERNEL_RESIDUAL 1
#define VFVM_KERNEL_SPMV64 2
#define VFVM_KERNEL_SPMV32 3
/* probe classes only: the batched second operator application fused with the BiCGStab
 * half step (s = r - alpha v gathered per column, t = Ahat s), FP64 and FP32 */
#define VFVM_KERNEL_HALFSTEP64 4
#define VFVM_KERNEL_HALFSTEP32 5
/* probe classes only: the AMG V-cycle's finest-level post-smoothing sweep
 * (out = x + omega (b - Ahat x)), FP64 and FP32 */
#define VFVM_KERNEL_AMG64 6
#define VFVM_KERNEL_AMG32 7
/* probe classes only: the batched one-block-per-member BiCGStab solve (k_msolve: one
 * launch per linear solve of the batch), FP64 and FP32 */
#define VFVM_KERNEL_MSOLVE64 8
#define VFVM_KERNEL_MSOLVE32 9
int vfvm_time_kernel(vfvm_sys *sys, int kernel, int reps, double *avg_ms);

/* Benchmark helper (no counterpart in the reference): live timing of kernel classes
 * inside real calls on the context.  After vfvm_probe_arm(kernel), every launch of that
 * class (VFVM_KERNEL_ASSEMBLE: Newton-step Jacobian assembly; _SPMV64 / _SPMV32: Krylov
 * operator applications; _HALFSTEP64 / _HALFSTEP32: the batched operator applications
 * fused with the half step; _AMG64 / _AMG32: the AMG cycle's finest-level sweep),
 * single-system or batched, is bracketed by HIP events on the
 * launching stream, up to max_launches launches.  Several classes may be armed at once.
 * vfvm_probe_read returns one class's summed duration, the number of timed launches and
 * (work, 3 entries, nullable) the work those launches did, then disarms that class:
 * SpMV: rows, nonzeros, 0; assembly: nodes, adjacency entries without the node itself
 * (2 x edges), Jacobian slots; AMG sweep: rows, nonzeros, 0; member solve: the sums over
 * members of iterations x rows, iterations x operator nonzeros and iterations x the
 * matrix entries one AMG V-cycle reads (A of every level twice, P and P^T once; 0 without
 * AMG).  In batches
 * only the members still active count. */
int vfvm_probe_arm(vfvm_ctx *ctx, int kernel, int64_t max_launches);
int vfvm_probe_read(vfvm_ctx *ctx, int kernel, double *total_ms, int64_t *launches, int64_t *work);

/* ---- device-resident interface (pointers are device pointers on the handle's device) ----
 * For hosts that keep U on the GPU (e.g. a Julia AMDGPU.jl array): same contracts as the host
 * versions, nzval_csr in the device CSR order of vfvm_pattern_csr. */
int vfvm_pattern_csr(vfvm_sys *sys, int32_t *rowptr, int32_t *colind);
int vfvm_assemble_device(vfvm_sys *sys, const double *dU, const double *dUold, double time,
                         double tstep, double lambda, double *dF, double *dnzval_csr);

/* ---- batched solves (the independent-SystemState pattern of
 * examples/Example440_ParallelState.jl:36-45, packed into single launches) ----
 * A batch concatenates systems created on one context into a block-diagonal device system.
 * Members share species count and physics kinds; each keeps its own parameters, source,
 * boundary table, damping and convergence.  Member data are copied at creation (later
 * set_source / set_bc calls on a member do not reach the batch). */
typedef struct vfvm_batch vfvm_batch;
int vfvm_batch_create(vfvm_ctx *ctx, int64_t nsys, vfvm_sys *const *systems, vfvm_batch **out);
int vfvm_batch_destroy(vfvm_batch *b);

/* Batched Krylov iterations between status reads run as captured HIP graphs (default on);
 * enable = 0 launches the same kernels one by one.  No counterpart in the reference. */
int vfvm_batch_set_graphs(vfvm_batch *b, int enable);
/* The batched Krylov solves run as one block per member (block = 256, 512 or 1024
 * threads): each member's whole BiCGStab (and AMG V-cycles) in one block, one launch per
 * linear solve, no host read between iterations; block = 0 selects the kernel-per-step
 * form (one launch per BiCGStab step for all members, iteration blocks as graphs).  Same
 * iteration and stopping rules either way.  No counterpart in the reference. */
int vfvm_batch_set_member_solver(vfvm_batch *b, int block);
/* the batch's SpMV column width, as vfvm_set_sell_index */
int vfvm_batch_set_sell_index(vfvm_batch *b, int bits);
/* the batch's assembly graph, as vfvm_set_compact_graph */
int vfvm_batch_set_compact_graph(vfvm_batch *b, int enable);
/* replace member m's source (nspec x nnodes; NULL keeps it) and boundary table (NULL keeps
 * it) inside the batch, for repeated solves of one grid with new data */
int vfvm_batch_update_member(vfvm_batch *b, int64_t m, const double *src, const vfvm_bc_t *bc);
/* benchmark helper (no counterpart in the reference): average device time in ms of `reps`
 * launches of one batched kernel (VFVM_KERNEL_ASSEMBLE, _SPMV64, _SPMV32) over all members,
 * HIP events on the batch's stream */
int vfvm_batch_time_kernel(vfvm_batch *b, int kernel, int reps, double *avg_ms);
/* project-written autotuner for the batch's assembly and SpMV group sizes */
int vfvm_batch_autotune(vfvm_batch *b, int32_t *asm_group, int32_t *spmv_group);
/* solve_step! for every member (src/vfvm_solver.jl:13-222): U, Uold are nsys host pointers to
 * nspec x nnodes arrays; hist (may be NULL) receives nsys histories.  Batched solves use
 * BiCGStab with row-scaling Jacobi, FP64 or FP32 inside FP64 refinement.  Returns
 * VFVM_ERR_CONVERGENCE if any member failed (the others' solutions are still returned). */
int vfvm_batch_newton(vfvm_batch *b, double *const *U, const double *const *Uold, double time,
                      double tstep, double lambda, const vfvm_control_t *control,
                      vfvm_history_t *hist);
/* The same with start values and solutions in device memory, packed in member order
 * (member m's dofs at offsets dof_offsets[m] .. dof_offsets[m+1] of vfvm_batch_sizes):
 * no host transfer of the unknowns.  No counterpart in the reference. */
int vfvm_batch_newton_device(vfvm_batch *b, double *U_dev, const double *Uold_dev, double time, double tstep,
                             double lambda, const vfvm_control_t *control, vfvm_history_t *hist);
/* total dofs of the batch and (if not NULL) the nsys + 1 member offsets */
int vfvm_batch_sizes(vfvm_batch *b, int64_t *ndof, int64_t *dof_offsets);

/* ---- learned Newton initial guess (no counterpart in the reference) ----
 * A message-passing network on the Voronoi edge graph of `sys` (edge weights sigma/h,
 * normalised per node) predicts the Newton initial iterate from per-node features.
 * Dense layers run on the matrix cores in bfloat16 with float32 accumulation; training and
 * inference use the library's own forward and backward kernels.  Newton always finishes the
 * solve, so the plain method stays the fallback and the judge of the prediction.
 * features: nsamples x nnodes x 8 float32 (sample-major, then node, then feature);
 * targets: nsamples x nnodes float32.  width must be 64. */
#define VFVM_MPNN_NFEAT 8
typedef struct vfvm_mpnn vfvm_mpnn;
int vfvm_mpnn_create(vfvm_sys *sys, int nlayers, int width, uint64_t seed, vfvm_mpnn **out);
int vfvm_mpnn_destroy(vfvm_mpnn *m);
int vfvm_mpnn_params(vfvm_mpnn *m, int64_t *nparams, float *params /* may be NULL */);
int vfvm_mpnn_set_params(vfvm_mpnn *m, const float *params);
/* Adam on the mean squared error, mini-batches of `batch` samples, `epochs` passes */
int vfvm_mpnn_train(vfvm_mpnn *m, int64_t nsamples, const float *features, const float *targets,
                    int64_t epochs, int64_t batch, float lr, float *final_loss);
/* out: nsamples x nnodes predictions */
int vfvm_mpnn_infer(vfvm_mpnn *m, int64_t nsamples, const float *features, double *out);

/* Inference with the 8 features of a 2D family with one source field and a Dirichlet
 * value per side of the grid's bounding box, computed on the device: f (nsamples x nnodes
 * float64 sources), sides (nsamples x 4: south, east, north, west), eps the diffusion
 * coefficient (side decay length sqrt(eps / 2)).  Features: f / 5, asinh(f / 2),
 * side value k * exp(-distance to side k / sqrt(eps / 2)), x, y. */
int vfvm_mpnn_infer_sides(vfvm_mpnn *m, int64_t nsamples, const double *f, const double *sides, double eps,
                          double *out);
/* the same with f, sides and out in device memory (out: nsamples x nnodes float64) */
int vfvm_mpnn_infer_sides_device(vfvm_mpnn *m, int64_t nsamples, const double *f_dev, const double *sides_dev,
                                 double eps, double *out_dev);
/* benchmark helper: average device time in ms of one inference hidden layer (graph
 * aggregation fused with the 128 x 64 GEMM, residual add; the kernel inference runs)
 * over `rows` rows (whole samples of the bound grid: rows = samples x nodes) */
int vfvm_mpnn_time_layer(vfvm_mpnn *m, int64_t rows, int reps, double *avg_ms);
/* loss and gradient for one batch (checked against the CPU oracle) */
int vfvm_mpnn_loss_grad(vfvm_mpnn *m, int64_t nsamples, const float *features, const float *targets,
                        float *loss, float *grad);

#ifdef __cplusplus
}
#endif
#endif
