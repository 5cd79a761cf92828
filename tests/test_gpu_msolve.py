"""The one-block-per-member batched Krylov solver (Batch.set_member_solver): each
member's whole BiCGStab, with its AMG V-cycles, in one block of one launch.  Against the
oracle and against the kernel-per-step form on the same batch: the same converged
solutions and about the same iteration counts (the dots are summed in another order)."""
import numpy as np
import pytest

import cases
import vfvm

pytestmark = pytest.mark.gpu

SOL_ATOL = 1e-9


def _print(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("block", [256, 1024])
@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("precon", ["jacobi", "amg"])
def test_member_solver_matches_oracle_and_step_form(precon, precision, block):
    """uneven members (9^2 nodes: a dense coarsest level 0; 17^2 and 33^2: two and three
    levels), Newton from zero, so members drop out at different Newton steps"""
    _print("case member solver", precon, precision, block)
    members = [cases.batch_member(seed, sizes=(9, 17, 33)) for seed in range(12)]
    pairs = [cases.nonlinear_poisson_2d(n, eps, f, tuple(v)) for (n, eps, v, f) in members]
    systems = [p[0] for p in pairs]
    ctrl = vfvm.SolverControl(precon=precon, precision=precision, reltol_linear=1e-10, abstol_linear=0.0,
                              maxiters_linear=2000)
    batch = vfvm.Batch(systems)
    Us = [vfvm.dofs(u).copy() for u in batch.solve(0.0, ctrl)]
    it_step = sum(h.nlin for h in batch.histories)
    batch.set_member_solver(block)
    for rep in range(2):
        Um = batch.solve(0.0, ctrl)
        it_mem = sum(h.nlin for h in batch.histories)
        for k, ((s, o), U) in enumerate(zip(pairs, Um)):
            assert batch.histories[k].residual_norm <= 1e-10, (k, batch.histories[k].residual_norm)
            assert np.abs(vfvm.dofs(U) - Us[k]).max() <= SOL_ATOL, k
            if rep == 0 and k % 4 == 0:
                Uo, _ = o.solve_step(np.zeros(s.grid.num_nodes))
                assert np.abs(vfvm.dofs(U) - Uo).max() <= SOL_ATOL, k
        _print("krylov iterations: step form", it_step, "member solver", it_mem)
        assert abs(it_mem - it_step) <= 0.1 * it_step + 12


def test_member_solver_device_resident_and_back():
    """the device-resident entry point with the member solver, then the step form again on
    the same batch: equal solutions"""
    _print("case member solver device")
    members = [cases.batch_member(seed, sizes=(17, 33)) for seed in range(8)]
    systems = [cases.nonlinear_poisson_2d(n, eps, f, tuple(v), oracle=False)[0] for (n, eps, v, f) in members]
    ctrl = vfvm.SolverControl(precon="amg", precision="fp32", reltol_linear=1e-4, abstol_linear=0.0,
                              maxiters_linear=500, inner_reltol=1e-4)
    batch = vfvm.Batch(systems)
    ctx = systems[0].ctx
    ndof, _ = batch.sizes()
    Z = vfvm.DeviceArray.from_numpy(ctx, np.zeros(ndof))
    U = vfvm.DeviceArray(ctx, ndof)
    batch.set_member_solver(512)
    batch.solve_device(U, Z, ctrl)
    a = U.numpy().copy()
    assert max(h.residual_norm for h in batch.histories) <= 1e-10
    batch.set_member_solver(0)
    batch.solve_device(U, Z, ctrl)
    assert np.abs(U.numpy() - a).max() <= SOL_ATOL
