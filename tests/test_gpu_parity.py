"""Parity of the HIP path (through the C ABI) with the CPU oracle.

Bars: Jacobian pattern identical; residual and Jacobian values within 1e-12
relative (normwise, penalty and non-penalty rows separately); converged
solutions within 1e-9 in the max norm; ||F||_inf <= 1e-10 over non-penalty
rows at the returned solution.
"""
import math

import numpy as np
import pytest

import cases
import oracle as O
import vfvm
from vfvm.grids import linspace, simplexgrid

pytestmark = pytest.mark.gpu

ASM_RTOL = 1e-12
SOL_ATOL = 1e-9


def _print(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("n", [5, 33])
def test_assembly_nonlinear_poisson_2d(n):
    _print("case assembly poisson2d", n)
    sys, osys = cases.nonlinear_poisson_2d(n)
    rng = np.random.default_rng(n)
    U = rng.uniform(-1, 1, (1, sys.grid.num_nodes))
    cases.compare_assembly(sys, osys, U)


def test_assembly_nonlinear_diffusion_3d_transient():
    _print("case assembly nldiff3d")
    sys, osys = cases.nonlinear_diffusion_3d(9)
    rng = np.random.default_rng(1)
    U = rng.uniform(0, 1, (1, sys.grid.num_nodes))
    cases.compare_assembly(sys, osys, U, Uold=0.7 * U, tstep=1e-2)


def test_assembly_drift_diffusion_2d():
    _print("case assembly dd2d")
    sys, osys = cases.drift_diffusion_2d(17, lambda2=1e-2)
    rng = np.random.default_rng(2)
    U = cases.dd_initial(sys.grid)
    U[0] += rng.uniform(-2, 2, sys.grid.num_nodes)  # Bernoulli arguments across all branches
    U[1:] *= rng.uniform(0.5, 1.5, (2, sys.grid.num_nodes))
    cases.compare_assembly(sys, osys, U, embed=0.3)


@pytest.mark.parametrize("group", [0, 1, 2, 4, 8, 16, 32])
def test_assembly_kernel_variants(group):
    """Every assembly kernel variant the autotuner may pick (lanes per node; 1 and 0: one
    node per lane with LDS staging) against the oracle, for 1 and 3 species."""
    _print("case assembly variant", group)
    sys, osys = cases.nonlinear_diffusion_3d(9)
    sys.set_groups(group)
    u0 = cases.initial_sine_3d(sys.grid).reshape(1, -1)
    cases.compare_assembly(sys, osys, 1.3 * u0, Uold=u0, tstep=1e-2)
    sys, osys = cases.drift_diffusion_2d(17, lambda2=1e-2)
    sys.set_groups(group)
    rng = np.random.default_rng(5)
    U = cases.dd_initial(sys.grid)
    U[0] += rng.uniform(-2, 2, sys.grid.num_nodes)
    cases.compare_assembly(sys, osys, U, embed=0.3)


def test_assembly_two_species_exchange_legacy_bc():
    _print("case assembly exchange")
    X = linspace(0, 1, 12)
    g = simplexgrid(X, X)
    rng = np.random.default_rng(3)
    src = rng.random((2, g.num_nodes))
    sys, osys = cases.pair(g, 2, dict(flux=vfvm.LinearFlux([1e-2, 2e-2]), reaction=vfvm.ExchangeReaction(1.0),
                                      storage=vfvm.IdentityStorage()),
                           legacy=[(1, 2, 0.1), (2, 4, 0.3)], source=np.asfortranarray(src))
    U = rng.random((2, g.num_nodes))
    cases.compare_assembly(sys, osys, np.asfortranarray(U), Uold=np.asfortranarray(0.5 * U), tstep=0.01)


def test_assembly_1d_power_robin():
    _print("case assembly power1d")
    g = simplexgrid(linspace(0, 1, 41))
    sys, osys = cases.pair(g, 1, dict(flux=vfvm.PowerFlux(2, 1.0), reaction=vfvm.PowerReaction(2, 1.0)),
                           [vfvm.Dirichlet(1, 1, 0.1), vfvm.Neumann(1, 2, 3.0)])
    U = np.linspace(0.1, 2.0, g.num_nodes).reshape(1, -1)
    cases.compare_assembly(sys, osys, U)


def test_newton_nonlinear_poisson_matches_oracle():
    """cfg0: 101 x 101 nonlinear Poisson, FP64 device Newton vs oracle (direct solve)."""
    _print("case newton poisson2d 101")
    sys, osys = cases.nonlinear_poisson_2d(101)
    Uo, ho = osys.solve_step(np.zeros(sys.grid.num_nodes))
    c = vfvm.SolverControl(reltol_linear=1e-12, abstol_linear=1e-14, maxiters_linear=2000, precon="ilu0")
    U = vfvm.solve(sys, inival=0.0, control=c)
    err = np.abs(vfvm.dofs(U) - Uo).max()
    _print("newton iters gpu", sys.history.niter, "oracle", ho.niter, "err", err, "res", sys.history.residual_norm)
    assert err <= SOL_ATOL
    assert sys.history.residual_norm <= 1e-10


@pytest.mark.parametrize("precon,precision", [("jacobi", "fp64"), ("ilu0", "fp64"), ("jacobi", "fp32"),
                                              ("ilu0", "fp32")])
def test_newton_solver_variants(precon, precision):
    _print("case newton variants", precon, precision)
    sys, osys = cases.nonlinear_poisson_2d(41, eps=1e-2, f=2.0, values=(-0.5, 1.0, 0.2, 0.0))
    Uo, _ = osys.solve_step(np.zeros(sys.grid.num_nodes))
    c = vfvm.SolverControl(precon=precon, precision=precision, reltol_linear=1e-10, abstol_linear=1e-14,
                           maxiters_linear=3000)
    U = vfvm.solve(sys, inival=0.0, control=c)
    assert np.abs(vfvm.dofs(U) - Uo).max() <= SOL_ATOL
    assert sys.history.residual_norm <= 1e-10


@pytest.mark.parametrize("krylov,precon,precision", [("bicgstab", "jacobi", "fp64"), ("bicgstab", "ilu0", "fp64"),
                                                     ("gmres", "jacobi", "fp64"), ("gmres", "ilu0", "fp32"),
                                                     ("bicgstab", "ilu0", "fp32")])
def test_linear_solvers_agree_with_direct(krylov, precon, precision):
    """examples/DevEx003_Solvers.jl:85-177 rule: every solver agrees with the direct one to 1e-7."""
    _print("case linsolve", krylov, precon, precision)
    import scipy.sparse.linalg as sla
    sys, osys = cases.nonlinear_poisson_2d(51)
    rng = np.random.default_rng(4)
    U = rng.uniform(-1, 1, (1, sys.grid.num_nodes))
    F, A = vfvm.evaluate_residual_and_jacobian(sys, U)
    b = vfvm.dofs(F)
    xd = sla.spsolve(A.tocsc(), b)
    from vfvm import _capi
    import ctypes as C
    c = vfvm.SolverControl(krylov=krylov, precon=precon, precision=precision, reltol_linear=1e-11,
                           abstol_linear=0.0, maxiters_linear=5000).to_c()
    x = np.zeros_like(b)
    st = _capi.LinStatsT()
    _capi.check(_capi.lib().vfvm_linsolve(sys.handle(), _capi.dptr(np.ascontiguousarray(b)), _capi.dptr(x),
                                          C.byref(c), C.byref(st)))
    _print("iters", st.iters, "refinements", st.refinements)
    assert np.abs(x - xd).max() <= 1e-7 * max(1.0, np.abs(xd).max())


@pytest.mark.parametrize("case", ["poisson2d", "nldiff3d", "dd2d"])
def test_natural_ilu0_factor_matches_oracle(case):
    """precon "ilu0_natural" is ILUZero's ilu0 (natural order, level-scheduled on the
    device): its L and U equal the oracle's natural-order factor of the same Jacobian, and
    its BiCGStab solve agrees with the direct one.  The multicolour "ilu0" factors a
    reordered matrix, so its factor differs."""
    _print("case natural ilu0", case)
    import scipy.sparse.linalg as sla
    rng = np.random.default_rng(11)
    if case == "poisson2d":
        sys, _ = cases.nonlinear_poisson_2d(17)
        U, tstep = rng.uniform(-1, 1, (1, sys.grid.num_nodes)), math.inf
    elif case == "nldiff3d":
        sys, _ = cases.nonlinear_diffusion_3d(6)
        U, tstep = rng.uniform(0, 1, (1, sys.grid.num_nodes)), 0.05
    else:
        sys, _ = cases.drift_diffusion_2d(9, lambda2=1e-2)
        U, tstep = vfvm.unknowns(sys, inival=cases.dd_initial(sys.grid)), math.inf
    F, A = vfvm.evaluate_residual_and_jacobian(sys, U, tstep=tstep)
    b = vfvm.dofs(F)
    rp, ci = sys.pattern_csr()
    Ar = A.tocsr()
    Ar.sort_indices()
    assert np.array_equal(Ar.indptr, rp) and np.array_equal(Ar.indices, ci)
    ref = O.ilu0(rp, ci, Ar.data)
    rowmax = np.array([np.abs(ref[rp[i]:rp[i + 1]]).max() for i in range(rp.size - 1)])
    tol = 1e-10 * np.abs(ref) + 1e-13 * np.repeat(rowmax, np.diff(rp))
    kw = dict(reltol_linear=1e-11, abstol_linear=0.0, maxiters_linear=5000)
    x, st = vfvm.linsolve(sys, b, precon="ilu0_natural", **kw)
    lu = vfvm.ilu0_factor(sys)
    worst = np.max(np.abs(lu - ref) / tol)
    xd = sla.spsolve(A.tocsc(), b)
    _print(case, "natural ilu0: iters", st["iters"], "factor err / tol", worst)
    assert worst <= 1.0
    assert np.abs(vfvm.dofs(x) - xd).max() <= 1e-7 * max(1.0, np.abs(xd).max())
    try:  # the multicolour factor is built whether or not its solve converges (on dd2d it does not)
        vfvm.linsolve(sys, b, precon="ilu0", reltol_linear=1e-2, abstol_linear=0.0, maxiters_linear=50)
    except vfvm.LinearSolverError:
        pass
    assert np.max(np.abs(vfvm.ilu0_factor(sys) - ref) / tol) > 1e3
    # back to the natural structure: the factor is rebuilt in that order
    vfvm.linsolve(sys, b, precon="ilu0_natural", **kw)
    assert np.max(np.abs(vfvm.ilu0_factor(sys) - ref) / tol) <= 1.0


def test_newton_with_natural_ilu0_matches_oracle():
    _print("case newton natural ilu0")
    sys, osys = cases.nonlinear_poisson_2d(33)
    for precision in ("fp64", "fp32"):
        ctrl = vfvm.SolverControl(precon="ilu0_natural", precision=precision, reltol_linear=1e-10,
                                  abstol_linear=0.0, maxiters_linear=2000)
        U = vfvm.solve(sys, inival=0.0, control=ctrl)
        Uo, _ = osys.solve_step(np.zeros(sys.grid.num_nodes))
        assert np.abs(vfvm.dofs(U) - Uo).max() <= SOL_ATOL
        assert sys.history.residual_norm <= 1e-10


def test_block_jacobi_gmres_drift_diffusion_continuation():
    """cfg2 at small size: bias continuation with GMRES + 3x3 block-Jacobi vs oracle."""
    _print("case dd continuation")
    sys, osys = cases.drift_diffusion_2d(33, lambda2=1e-2)
    U0 = cases.dd_initial(sys.grid)
    lam = np.linspace(0.0, 1.0, 6)
    c = O.control(method_linear="direct", dp=0.2, dp_min=0.2, dp_max=0.2, du_opt=10.0)
    Uo, _, _ = osys.solve_transient(vfvm.dofs(U0), lam, c, transient=False)
    for precision in ("fp64", "fp32"):
        ctrl = vfvm.SolverControl(krylov="gmres", precon="blockjacobi", precision=precision, reltol_linear=1e-10,
                                  abstol_linear=0.0, maxiters_linear=3000, gmres_restart=30, Δp=0.2, Δp_min=0.2,
                                  Δp_max=0.2, Δu_opt=10.0)
        U = vfvm.solve(sys, inival=U0, control=ctrl, embed=lam).u[-1]
        err = np.abs(vfvm.dofs(U) - Uo).max()
        _print("dd", precision, "err", err)
        assert err <= SOL_ATOL


def test_cfg2_configured_solver_at_65():
    """The cfg2 solver as configured: GMRES(30) with 3x3 block Jacobi, FP32 inside FP64
    refinement, Debye length lambda = 1e-2 (lambda^2 = 1e-4), a bias sweep in steps of 0.05 V,
    on 65 x 65 nodes against the oracle's direct solves.  The FP32 solves that cannot reach
    the tolerance are finished in FP64 (counted in the history), so the sweep's answer is
    the oracle's whatever their number."""
    _print("case cfg2 solver 65")
    sys, osys = cases.drift_diffusion_2d(65, lambda2=1e-4)
    U0 = cases.dd_initial(sys.grid)
    lam = np.linspace(0.0, 0.5, 11)
    c = O.control(method_linear="direct", dp=0.05, dp_min=0.05, dp_max=0.05, du_opt=10.0)
    Uo, _, _ = osys.solve_transient(vfvm.dofs(U0), lam, c, transient=False)
    ctrl = vfvm.SolverControl(krylov="gmres", precon="blockjacobi", precision="fp32", gmres_restart=30,
                              reltol_linear=1e-10, abstol_linear=0.0, maxiters_linear=6000, Δp=0.05, Δp_min=0.05,
                              Δp_max=0.05, Δu_opt=10.0)
    U = vfvm.solve(sys, inival=U0, control=ctrl, embed=lam).u[-1]
    err = np.abs(vfvm.dofs(U) - Uo).max()
    _print("cfg2 65: err", err, "fp64 fallbacks of the last step", sys.history.fp64_fallbacks,
           "krylov", sys.history.nlin)
    assert err <= SOL_ATOL
    assert sys.history.residual_norm <= 1e-10


def test_ilu0_bicgstab_drift_diffusion_three_species():
    """cfg2 physics (lambda = 1e-2) on 65 x 65 nodes: BiCGStab with multicolour ILU(0)
    of the 3-species dof graph, FP64 and FP32 inside FP64 refinement, against the
    oracle's Newton with direct solves, bias steps 0, 0.25, 0.5."""
    _print("case dd ilu0")
    sys, osys = cases.drift_diffusion_2d(65, lambda2=1e-4)
    U0 = vfvm.dofs(cases.dd_initial(sys.grid)).copy()
    lam = [0.0, 0.25, 0.5]
    Uo = U0
    ref = []
    for l in lam:
        Uo, _ = osys.solve_step(Uo, embed=l)
        ref.append(Uo)
    for precision in ("fp64", "fp32"):
        ctrl = vfvm.SolverControl(krylov="bicgstab", precon="ilu0", precision=precision, reltol_linear=1e-8,
                                  abstol_linear=0.0, maxiters_linear=3000, max_refinements=20)
        U = U0.reshape(3, -1, order="F")
        for l, Ur in zip(lam, ref):
            U = vfvm.solve(sys, inival=U, control=ctrl, embed=[l]).u[-1]
            err = np.abs(vfvm.dofs(U) - Ur).max()
            _print("dd ilu0", precision, l, "err", err, "newton", sys.history.niter, "krylov", sys.history.nlin)
            assert err <= SOL_ATOL
            assert sys.history.residual_norm <= 1e-10


def test_transient_nonlinear_diffusion_3d():
    """cfg1 at small size: implicit Euler, 10 steps of 1e-2, BiCGStab + ILU(0) FP32 vs oracle."""
    _print("case transient nldiff3d")
    sys, osys = cases.nonlinear_diffusion_3d(9)
    u0 = cases.initial_sine_3d(sys.grid).reshape(1, -1)
    # the same (time, tstep) sequence on both paths: t_k = t_{k-1} + 1e-2 in floating point
    times = np.cumsum([0.0] + [1e-2] * 10)
    Uo = u0.ravel().copy()
    for k in range(1, times.size):
        Uo, _ = osys.solve_step(Uo, time=times[k], tstep=times[k] - times[k - 1])
    for precision in ("fp64", "fp32"):
        ctrl = vfvm.fixed_timesteps(vfvm.SolverControl(precon="ilu0", precision=precision, reltol_linear=1e-10,
                                                       abstol_linear=0.0, maxiters_linear=2000), 1e-2)
        U = vfvm.solve(sys, inival=np.asfortranarray(u0), control=ctrl, times=times).u[-1]
        assert np.abs(vfvm.dofs(U) - Uo).max() <= SOL_ATOL


def test_examples_known_answers_on_gpu():
    """The GPU path reproduces the reference's own testvals (Example105, Example301)."""
    _print("case known answers")
    X = np.arange(11) / 10.0
    g = simplexgrid(X)
    sys = vfvm.System(g, vfvm.Physics(flux=vfvm.LinearFlux(1e-3), reaction=vfvm.SinhReaction(),
                                      source=np.where(X <= 0.5, 1.0, -1.0)), species=[1])
    sys.boundary_dirichlet(1, 1, 0.0)
    sys.boundary_dirichlet(1, 2, 1.0)
    c = vfvm.SolverControl(reltol_linear=1e-13, abstol_linear=0.0, maxiters_linear=200)
    U = vfvm.solve(sys, inival=0.5, control=c)
    assert U.sum() == pytest.approx(1.5247901344230088, rel=1.4901161193847656e-08)
    X = np.arange(6) / 5.0
    g = simplexgrid(X, X, X)
    x, y, z = g.coord
    sys = vfvm.System(g, vfvm.Physics(flux=vfvm.LinearFlux(1.0), source=x * np.sin(5.0 * y) * np.exp(z)),
                      species=[1])
    sys.boundary_dirichlet(1, 5, 0.0)
    sys.boundary_dirichlet(1, 6, 0.0)
    U = vfvm.solve(sys, control=c)
    assert vfvm.dofs(U)[42] == pytest.approx(0.012234524449380824, rel=1.4901161193847656e-08)


def test_nan_raises_assembly_error():
    _print("case nan")
    sys, _ = cases.nonlinear_poisson_2d(9)
    U = np.zeros((1, sys.grid.num_nodes), order="F")
    U[0, 40] = np.nan
    with pytest.raises(vfvm.AssemblyError):
        vfvm.evaluate_residual_and_jacobian(sys, U)


def test_maxiters_raises_convergence_error():
    _print("case maxiters")
    sys, _ = cases.nonlinear_poisson_2d(21)
    with pytest.raises(vfvm.ConvergenceError):
        vfvm.solve(sys, inival=0.0, control=vfvm.SolverControl(maxiters=1))


def test_adaptive_time_stepping_matches_oracle():
    """solve_transient! step control on the device Newton loop against the oracle: the
    same accepted times (rejections and growth included) and the same solutions."""
    _print("case adaptive steps decay")
    X = linspace(0.0, 1.0, 6)
    g = simplexgrid(X)
    sys, osys = cases.pair(g, 1, dict(flux=vfvm.LinearFlux(0.0), reaction=vfvm.PowerReaction(1, 5.0),
                                      storage=vfvm.IdentityStorage()))
    ctrl = vfvm.SolverControl(reltol_linear=1e-12, abstol_linear=0.0)
    tsol = vfvm.solve(sys, inival=1.0, control=ctrl, times=[0.0, 1.0])
    Uo, n, _, steps = osys.solve_transient(np.ones(g.num_nodes), [0.0, 1.0], O.control(), return_steps=True)
    assert len(tsol.t) == n + 1 and tsol.t[0] == 0.0
    assert np.abs(np.asarray(tsol.t[1:]) - steps).max() <= 1e-14
    assert np.abs(vfvm.dofs(tsol.u[-1]) - Uo).max() <= SOL_ATOL
    _print("case adaptive steps nldiff 2d")
    X = linspace(0.0, 1.0, 17)
    g = simplexgrid(X, X)
    x, y = g.coord
    u0 = np.exp(-30.0 * ((x - 0.5) ** 2 + (y - 0.5) ** 2))
    bcs = [vfvm.Dirichlet(species=1, region=r, value=0.0) for r in (1, 2, 3, 4)]
    sys, osys = cases.pair(g, 1, dict(flux=vfvm.NonlinearDiffusionFlux(1e-2, 1.0), storage=vfvm.IdentityStorage()),
                           bcs)
    for precision in ("fp64", "fp32"):
        ctrl = vfvm.SolverControl(precon="ilu0", precision=precision, reltol_linear=1e-12, abstol_linear=0.0,
                                  maxiters_linear=2000, Δt=1e-3, Δt_min=1e-5, Δt_max=0.5, Δu_opt=0.05)
        tsol = vfvm.solve(sys, inival=u0.reshape(1, -1), control=ctrl, times=[0.0, 0.2, 1.0])
        oc = O.control(dt=1e-3, dt_min=1e-5, dt_max=0.5, du_opt=0.05)
        Uo, n, _, steps = osys.solve_transient(u0, [0.0, 0.2, 1.0], oc, return_steps=True)
        _print("adaptive", precision, "steps", n, "gpu", len(tsol.t) - 1)
        assert len(tsol.t) == n + 1
        assert np.abs(np.asarray(tsol.t[1:]) - steps).max() <= 1e-12
        assert np.abs(vfvm.dofs(tsol.u[-1]) - Uo).max() <= SOL_ATOL


def test_handle_exceptions_matches_oracle():
    """SolverControl.handle_exceptions (src/vfvm_solver.jl:391-452) on the device driver:
    with five Newton iterations the large steps of u_t = div((1e-2 + u^2) grad u) fail and
    are retried with half the step; the accepted times and solutions equal the oracle's.
    Without it the same run raises; a run whose smallest step fails returns its start
    with a warning."""
    _print("case handle_exceptions")
    X = linspace(0.0, 1.0, 17)
    g = simplexgrid(X, X)
    x, y = g.coord
    u0 = np.exp(-30.0 * ((x - 0.5) ** 2 + (y - 0.5) ** 2))
    bcs = [vfvm.Dirichlet(species=1, region=r, value=0.0) for r in (1, 2, 3, 4)]
    sys, osys = cases.pair(g, 1, dict(flux=vfvm.NonlinearDiffusionFlux(1e-2, 1.0), storage=vfvm.IdentityStorage()),
                           bcs)
    kw = dict(precon="ilu0", reltol_linear=1e-12, abstol_linear=0.0, maxiters_linear=2000, maxiters=5,
              Δt=0.5, Δt_min=1e-4, Δt_max=0.5, Δu_opt=10.0)
    with pytest.raises(vfvm.ConvergenceError):
        vfvm.solve(sys, inival=u0.reshape(1, -1), control=vfvm.SolverControl(**kw), times=[0.0, 1.0])
    tsol = vfvm.solve(sys, inival=u0.reshape(1, -1), control=vfvm.SolverControl(handle_exceptions=True, **kw),
                      times=[0.0, 1.0])
    oc = O.control(dt=0.5, dt_min=1e-4, dt_max=0.5, du_opt=10.0, maxiters=5, handle_exceptions=1)
    Uo, n, _, steps = osys.solve_transient(u0, [0.0, 1.0], oc, return_steps=True)
    _print("handle_exceptions steps", n, "gpu", len(tsol.t) - 1, "first", tsol.t[1])
    assert len(tsol.t) == n + 1
    assert np.abs(np.asarray(tsol.t[1:]) - steps).max() <= 1e-12
    assert np.abs(vfvm.dofs(tsol.u[-1]) - Uo).max() <= SOL_ATOL
    kw.update(maxiters=3, Δt_min=0.1)
    with pytest.warns(UserWarning, match="smallest step"):
        tsol = vfvm.solve(sys, inival=u0.reshape(1, -1), control=vfvm.SolverControl(handle_exceptions=True, **kw),
                          times=[0.0, 1.0])
    assert len(tsol.t) == 1 and tsol.t[0] == 0.0
    assert np.array_equal(vfvm.dofs(tsol.u[0]), u0)  # the start entry: the initial value as given


def test_continuation_with_embedded_source_matches_oracle():
    """Continuation (solve(...; embed)) on the nonlinear Poisson problem with source and
    Dirichlet values ramped by the embedding parameter: assembly at lambda = 0.37 to
    1e-12, the accepted parameter sequence of the step control and the final solution
    against the oracle, and the end point against the direct solve of the unramped
    system."""
    _print("case continuation embedded source")
    X = linspace(0.0, 1.0, 33)
    g = simplexgrid(X, X)
    rng = np.random.default_rng(11)
    f = rng.uniform(-5.0, 5.0, g.num_nodes)
    vals = (0.3, -0.7, 0.9, -0.2)
    bcs = [vfvm.Dirichlet(species=1, region=r + 1, value=0.0, value_lambda=v) for r, v in enumerate(vals)]
    sys, osys = cases.pair(g, 1, dict(flux=vfvm.LinearFlux(1e-2), reaction=vfvm.SinhReaction(), embed_source=True),
                           bcs, source=f.reshape(1, -1))
    U = rng.uniform(-1.0, 1.0, (1, g.num_nodes))
    cases.compare_assembly(sys, osys, U, embed=0.37)
    ctrl = vfvm.SolverControl(precon="ilu0", reltol_linear=1e-12, abstol_linear=0.0, maxiters_linear=2000,
                              Δp=0.25, Δp_min=1e-3, Δp_max=0.5, Δp_grow=1.5)
    sol = vfvm.solve(sys, inival=0.0, control=ctrl, embed=[0.0, 1.0])
    Uo, n, _, steps = osys.solve_transient(np.zeros(g.num_nodes), [0.0, 1.0],
                                           O.control(dp=0.25, dp_min=1e-3, dp_max=0.5, dp_grow=1.5),
                                           transient=False, return_steps=True)
    _print("continuation steps", n, "gpu", len(sol.t) - 1)
    assert len(sol.t) == n + 1 and sol.t[0] == 0.0
    assert np.abs(np.asarray(sol.t[1:]) - steps).max() <= 1e-12
    assert np.abs(vfvm.dofs(sol.u[-1]) - Uo).max() <= SOL_ATOL
    plain, _ = cases.pair(g, 1, dict(flux=vfvm.LinearFlux(1e-2), reaction=vfvm.SinhReaction()),
                          [vfvm.Dirichlet(species=1, region=r + 1, value=v) for r, v in enumerate(vals)],
                          source=f.reshape(1, -1), oracle=False)
    Ud = vfvm.solve(plain, inival=0.0, control=ctrl)
    assert np.abs(vfvm.dofs(sol.u[-1]) - vfvm.dofs(Ud)).max() <= SOL_ATOL


def test_fp32_refinement_hands_over_to_fp64():
    """cfg2 physics (lambda = 1e-2) on 257 x 257 nodes, bias 0..0.25 V: the FP32 inner
    solves of the block-Jacobi BiCGStab stagnate at this conditioning.  The failed
    refinement restores x and finishes in FP64, the rest of the sweep solves in FP64
    directly, and the result equals the plain FP64 sweep."""
    _print("case fp32 hand-over dd 257")
    sys, _ = cases.drift_diffusion_2d(257, lambda2=1e-4, oracle=False)
    U0 = cases.dd_initial(sys.grid)
    out = {}
    for precision in ("fp64", "fp32"):
        ctrl = vfvm.SolverControl(krylov="bicgstab", precon="blockjacobi", precision=precision, reltol_linear=1e-6,
                                  abstol_linear=0.0, maxiters_linear=6000, inner_reltol=1e-3, max_refinements=12,
                                  Δp=0.05, Δp_min=0.05, Δp_max=0.05)
        sol = vfvm.solve(sys, inival=U0, control=ctrl, embed=np.linspace(0.0, 0.25, 6), store_all=False)
        h = sys.history
        _print(precision, "newton", h.niter, "krylov", h.nlin, "fp64 fallbacks", h.fp64_fallbacks)
        assert h.residual_norm <= 1e-10
        out[precision] = (vfvm.dofs(sol.u[-1]).copy(), h.fp64_fallbacks)
    assert out["fp64"][1] == 0 and out["fp32"][1] >= 1
    assert np.abs(out["fp32"][0] - out["fp64"][0]).max() <= SOL_ATOL


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_sell_index_widths_agree(precision):
    """The SpMV with 16-bit column offsets and with 32-bit columns: the same arithmetic,
    so the same Newton iterates bit for bit (2D and 3D patterns, single system)."""
    _print("case sell index widths", precision)
    for sys, osys in (cases.nonlinear_poisson_2d(41, eps=1e-2, f=2.0, values=(-0.5, 1.0, 0.2, 0.0)),
                      cases.nonlinear_diffusion_3d(9)):
        c = vfvm.SolverControl(precision=precision, reltol_linear=1e-10, abstol_linear=0.0, maxiters_linear=3000)
        out = []
        for bits in (16, 32):
            sys.set_sell_index(bits)
            U = vfvm.solve(sys, inival=0.5, control=c)
            out.append((vfvm.dofs(U).copy(), sys.history.nlin))
            assert sys.history.residual_norm <= 1e-10
        assert np.array_equal(out[0][0], out[1][0]) and out[0][1] == out[1][1]


def test_compact_graph_is_bitwise_identical():
    """The assembly reading the compact node graph (16-bit neighbour offsets, factor
    table) gives bit for bit the residual and Jacobian of the plain graph, for every
    kernel family and 1 and 3 species."""
    _print("case compact graph")
    for sys, osys in (cases.nonlinear_diffusion_3d(9), cases.drift_diffusion_2d(17, lambda2=1e-2),
                      cases.nonlinear_poisson_2d(33)):
        rng = np.random.default_rng(9)
        U = rng.uniform(0.2, 1.0, (sys.num_species, sys.grid.num_nodes))
        for group in (0, 1, 8):
            sys.set_groups(group)
            out = []
            for compact in (True, False):
                sys.set_compact_graph(compact)
                F, A = vfvm.evaluate_residual_and_jacobian(sys, np.asfortranarray(U), embed=0.3)
                out.append((vfvm.dofs(F).copy(), A.tocsc().data.copy()))
            assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1]), group
        sys.set_compact_graph(True)
        cases.compare_assembly(sys, osys, U, embed=0.3)


@pytest.mark.parametrize("name", ["rz2d.sg", "cyl_unstruct.sg"])
def test_unstructured_sg_grid_matches_oracle(name):
    """Unstructured triangulations read from the reference's .sg assets: assembly at a
    random state and the device Newton solution against the oracle."""
    _print("case unstructured", name)
    sys, osys = cases.nonlinear_poisson_sg(name)
    rng = np.random.default_rng(5)
    U = rng.uniform(-1, 1, (1, sys.grid.num_nodes))
    cases.compare_assembly(sys, osys, U)
    Uo, ho = osys.solve_step(np.zeros(sys.grid.num_nodes))
    c = vfvm.SolverControl(reltol_linear=1e-12, abstol_linear=1e-14, maxiters_linear=2000, precon="ilu0")
    Ug = vfvm.solve(sys, inival=0.0, control=c)
    err = np.abs(vfvm.dofs(Ug) - Uo).max()
    _print("newton iters gpu", sys.history.niter, "oracle", ho.niter, "err", err, "res", sys.history.residual_norm)
    assert err <= SOL_ATOL
    assert sys.history.residual_norm <= 1e-10


def test_generic_operator_matches_oracle():
    """Physics(generic = ...) as a constant matrix G on the dof pattern: Example405's known
    answer on the GPU, F and J against the oracle with G added to a nonlinear Poisson
    system on 17^2 nodes, and an entry outside the pattern refused."""
    _print("case generic operator")
    from test_oracle_known_answers import _example405_matrix
    X = np.linspace(0.0, 1.0, 11)
    g = simplexgrid(X)
    ph = vfvm.Physics(generic=vfvm.GenericMatrix(_example405_matrix(X)))
    sys = vfvm.System(g, ph, species=[1])
    sys.boundary_dirichlet(1, 1, 0.0)
    sys.boundary_dirichlet(1, 2, 1.0)
    U = vfvm.solve(sys, inival=0.5, control=vfvm.SolverControl(reltol_linear=1e-12, abstol_linear=0.0))
    assert float(np.sum(U)) == pytest.approx(1.099999999614456, rel=1e-10)
    # a random G on the pattern of a 2D system
    import scipy.sparse as sp
    X = linspace(0.0, 1.0, 17)
    g2 = simplexgrid(X, X)
    base, _ = cases.pair(g2, 1, dict(flux=vfvm.LinearFlux(1e-2), reaction=vfvm.SinhReaction()))
    rp, ci = base.pattern_csr()
    rng = np.random.default_rng(8)
    n = g2.num_nodes
    G = sp.csr_matrix((rng.uniform(-1, 1, ci.size), ci, rp), shape=(n, n))
    bcs = [vfvm.Dirichlet(species=1, region=1, value=0.0), vfvm.Dirichlet(species=1, region=3, value=1.0)]
    sys2, osys2 = cases.pair(g2, 1, dict(flux=vfvm.LinearFlux(1e-2), reaction=vfvm.SinhReaction(),
                                         generic=vfvm.GenericMatrix(G)), bcs)
    osys2.set_generic(G)
    Ut = rng.uniform(-1, 1, (1, n))
    errs = cases.compare_assembly(sys2, osys2, Ut)
    _print("generic operator assembly errors", errs)
    assert max(errs.values()) <= 1e-12
    # outside the pattern: node 0 and the last node are not neighbours
    far = sp.coo_matrix(([1.0], ([0], [n - 1])), shape=(n, n))
    bad = vfvm.System(g2, vfvm.Physics(flux=vfvm.LinearFlux(1e-2), generic=vfvm.GenericMatrix(far)), species=[1])
    with pytest.raises(vfvm.VFVMError):
        bad.handle()
