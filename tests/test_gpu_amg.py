"""The smoothed aggregation AMG preconditioner (SolverControl(precon="amg")) on the GPU.

The reference tests BiCGStab preconditioned by AlgebraicMultigrid.jl's smoothed
aggregation (examples/DevEx003_Solvers.jl:149-156) by the rule of that file: the solution
agrees with the direct solve to 1e-7.  Beyond that rule the device hierarchy is checked
against its definition with scipy's sparse products, level by level:
  A_0 = diag(J)^-1 J,  omega = (4/3) / max_i sum_j |A_ij| / |A_ii|,
  P = (I - omega D^-1 A) T with T from the aggregates and the candidate vector,
  A_next = P^T A P,
a system small enough for one level is inverted exactly (one Krylov iteration), and Newton
solves and batches with the preconditioner match the oracle's direct solves."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sla

import cases
import vfvm
from vfvm import amg

pytestmark = pytest.mark.gpu

SOL_ATOL = 1e-9


def _print(*a):
    print(*a, flush=True)


def _rel(X, Y):
    d = abs(sp.csr_matrix(X) - sp.csr_matrix(Y))
    return (d.max() if d.nnz else 0.0) / abs(sp.csr_matrix(Y)).max()


def _tentative(agg, B):
    n, na = agg.size, int(agg.max()) + 1
    nrm = np.sqrt(np.bincount(agg, weights=B * B, minlength=na))
    return sp.csr_matrix((B / nrm[agg], (np.arange(n), agg)), shape=(n, na)), nrm


def _system(case):
    if case == "poisson2d":
        sys, _ = cases.nonlinear_poisson_2d(41, oracle=False)
        U = np.random.default_rng(1).uniform(-1, 1, (1, sys.grid.num_nodes))
        return sys, U, np.inf
    sys, _ = cases.nonlinear_diffusion_3d(13, oracle=False)
    U = np.random.default_rng(2).uniform(0, 1, (1, sys.grid.num_nodes))
    return sys, U, 1e-2


@pytest.mark.parametrize("case", ["poisson2d", "nldiff3d"])
def test_hierarchy_is_smoothed_aggregation(case):
    _print("case amg hierarchy", case)
    sys, U, tstep = _system(case)
    F, J = vfvm.evaluate_residual_and_jacobian(sys, U, tstep=tstep)
    b = vfvm.dofs(F)
    x, st = vfvm.linsolve(sys, b, precon="amg", reltol_linear=1e-10, abstol_linear=0.0, maxiters_linear=500)
    H = amg.hierarchy(sys)
    _print("levels", [L["A"].shape[0] for L in H], "iters", st["iters"])
    assert len(H) >= 2 and H[-1]["A"].shape[0] <= 64
    J = J.tocsr()
    assert _rel(H[0]["A"], sp.diags(1.0 / J.diagonal()) @ J) <= 1e-15
    B = np.ones(J.shape[0])
    for lvl, L in enumerate(H[:-1]):
        A = L["A"]
        D = A.diagonal()
        rho = (np.asarray(abs(A).sum(axis=1)).ravel() / np.abs(D)).max()
        assert L["omega"] == pytest.approx((4.0 / 3.0) / rho, rel=1e-14)
        T, nrm = _tentative(L["agg"], B)
        P = T - L["omega"] * (sp.diags(1.0 / D) @ A @ T)
        errP = _rel(L["P"], P)
        Ac = (L["P"].T @ A @ L["P"]).tocsr()
        errA = _rel(H[lvl + 1]["A"], Ac)
        _print("level", lvl, "P err", errP, "Galerkin err", errA)
        assert errP <= 1e-13
        assert errA <= 1e-12
        B = nrm
    xd = sla.spsolve(J.tocsc(), b)
    assert np.abs(x - xd).max() <= 1e-7 * max(1.0, np.abs(xd).max())


def test_one_level_system_is_solved_exactly():
    """49 rows (at most 64): the hierarchy is its coarsest level, inverted by Gauss-Jordan, and
    right-preconditioned BiCGStab stops at the first half step."""
    _print("case amg exact")
    sys, _ = cases.nonlinear_poisson_2d(7, oracle=False)
    U = np.random.default_rng(5).uniform(-1, 1, (1, sys.grid.num_nodes))
    F, J = vfvm.evaluate_residual_and_jacobian(sys, U)
    b = vfvm.dofs(F)
    x, st = vfvm.linsolve(sys, b, precon="amg", reltol_linear=1e-12, abstol_linear=0.0, maxiters_linear=10)
    assert len(amg.hierarchy(sys)) == 1
    assert st["iters"] == 1
    xd = sla.spsolve(J.tocsc(), b)
    assert np.abs(x - xd).max() <= 1e-12 * np.abs(xd).max()


@pytest.mark.parametrize("krylov,precision", [("bicgstab", "fp64"), ("bicgstab", "fp32"), ("gmres", "fp64"),
                                              ("gmres", "fp32")])
def test_linsolve_agrees_with_direct_in_few_iterations(krylov, precision):
    """the DevEx003 rule (1e-7 against the direct solve), in a fraction of point Jacobi's
    iterations"""
    _print("case amg linsolve", krylov, precision)
    sys, _ = cases.nonlinear_poisson_2d(51, oracle=False)
    U = np.random.default_rng(4).uniform(-1, 1, (1, sys.grid.num_nodes))
    F, J = vfvm.evaluate_residual_and_jacobian(sys, U)
    b = vfvm.dofs(F)
    xd = sla.spsolve(J.tocsc(), b)
    kw = dict(krylov=krylov, precision=precision, reltol_linear=1e-11, abstol_linear=0.0, maxiters_linear=5000,
              gmres_restart=30)
    x, st = vfvm.linsolve(sys, b, precon="amg", **kw)
    _print("iterations amg", st["iters"], "refinements", st["refinements"])
    assert np.abs(x - xd).max() <= 1e-7 * max(1.0, np.abs(xd).max())
    if precision == "fp64":
        _, sj = vfvm.linsolve(sys, b, precon="jacobi", **kw)
        _print("iterations jacobi", sj["iters"])
        assert 4 * st["iters"] <= sj["iters"]


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("every_step", [False, True])
def test_newton_with_amg_matches_oracle(precision, every_step):
    """Newton with the hierarchy of the first step reused (the default, as an ILU(0) factor)
    or rebuilt every step (factorize_every_newtonstep)"""
    _print("case amg newton", precision, every_step)
    sys, osys = cases.nonlinear_poisson_2d(41, eps=1e-2, f=2.0, values=(-0.5, 1.0, 0.2, 0.0))
    Uo, _ = osys.solve_step(np.zeros(sys.grid.num_nodes))
    c = vfvm.SolverControl(precon="amg", precision=precision, reltol_linear=1e-10, abstol_linear=1e-14,
                           maxiters_linear=500, factorize_every_newtonstep=every_step)
    U = vfvm.solve(sys, inival=0.0, control=c)
    err = np.abs(vfvm.dofs(U) - Uo).max()
    _print("err", err, "newton", sys.history.niter, "krylov", sys.history.nlin)
    assert err <= SOL_ATOL
    assert sys.history.residual_norm <= 1e-10


def test_transient_3d_with_amg_matches_oracle():
    _print("case amg transient 3d")
    sys, osys = cases.nonlinear_diffusion_3d(9)
    u0 = cases.initial_sine_3d(sys.grid).reshape(1, -1)
    times = np.cumsum([0.0] + [1e-2] * 5)
    Uo = u0.ravel().copy()
    for k in range(1, times.size):
        Uo, _ = osys.solve_step(Uo, time=times[k], tstep=times[k] - times[k - 1])
    ctrl = vfvm.fixed_timesteps(vfvm.SolverControl(precon="amg", precision="fp32", reltol_linear=1e-10,
                                                   abstol_linear=0.0, maxiters_linear=500), 1e-2)
    U = vfvm.solve(sys, inival=np.asfortranarray(u0), control=ctrl, times=times).u[-1]
    assert np.abs(vfvm.dofs(U) - Uo).max() <= SOL_ATOL


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_batch_with_amg_matches_oracle(precision):
    """uneven members (9^2, 17^2, 33^2 nodes: hierarchies of two and three levels side by
    side), against the oracle, single-system AMG solves and the Jacobi batch's iterations"""
    _print("case amg batch", precision)
    members = [cases.batch_member(seed, sizes=(9, 17, 33)) for seed in range(12)]
    pairs = [cases.nonlinear_poisson_2d(n, eps, f, tuple(v)) for (n, eps, v, f) in members]
    systems = [p[0] for p in pairs]
    kw = dict(precision=precision, reltol_linear=1e-10, abstol_linear=0.0, maxiters_linear=2000)
    batch = vfvm.Batch(systems)
    Ub = batch.solve(0.0, vfvm.SolverControl(precon="amg", **kw))
    nlin_amg = sum(h.nlin for h in batch.histories)
    for k, ((s, o), U) in enumerate(zip(pairs, Ub)):
        Uo, _ = o.solve_step(np.zeros(s.grid.num_nodes))
        err = np.abs(vfvm.dofs(U) - Uo).max()
        assert err <= SOL_ATOL, (k, err)
        assert batch.histories[k].residual_norm <= 1e-10
    for k in (0, 5, 11):
        Us = vfvm.solve(systems[k], inival=0.0, control=vfvm.SolverControl(precon="amg", **kw))
        assert np.abs(vfvm.dofs(Us) - vfvm.dofs(Ub[k])).max() <= SOL_ATOL
    batch.solve(0.0, vfvm.SolverControl(precon="jacobi", **kw))
    nlin_jac = sum(h.nlin for h in batch.histories)
    _print("batch krylov iterations amg", nlin_amg, "jacobi", nlin_jac)
    assert 3 * nlin_amg <= nlin_jac


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_batch_amg_graph_replays_match_eager(precision):
    """the AMG V-cycles inside the captured Krylov blocks give the eager launches' results"""
    _print("case amg batch graphs", precision)
    members = [cases.batch_member(seed, sizes=(17, 33)) for seed in range(8)]
    systems = [cases.nonlinear_poisson_2d(n, eps, f, tuple(v), oracle=False)[0] for (n, eps, v, f) in members]
    ctrl = vfvm.SolverControl(precon="amg", precision=precision, reltol_linear=1e-10, abstol_linear=0.0,
                              maxiters_linear=500, check_every=2)
    batch = vfvm.Batch(systems)
    batch.set_graphs(False)
    Ue = [vfvm.dofs(u).copy() for u in batch.solve(0.0, ctrl)]
    it_e = [h.nlin for h in batch.histories]
    batch.set_graphs(True)
    for replay in range(3):
        Ug = batch.solve(0.0, ctrl)
        assert [h.nlin for h in batch.histories] == it_e
        for a, b in zip(Ue, Ug):
            assert np.array_equal(a, vfvm.dofs(b)), replay
        vfvm.solve(systems[0], inival=0.0, control=ctrl)
