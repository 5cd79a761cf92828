"""Batched solves (many systems in one launch sequence) against the oracle and
against one-at-a-time GPU solves."""
import numpy as np
import pytest

import cases
import vfvm

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_batch_matches_oracle_and_single_solves(precision):
    print("case batch", precision, flush=True)
    members = [cases.batch_member(seed, sizes=(9, 17, 33)) for seed in range(12)]
    pairs = [cases.nonlinear_poisson_2d(n, eps, f, tuple(v)) for (n, eps, v, f) in members]
    systems = [p[0] for p in pairs]
    ctrl = vfvm.SolverControl(precision=precision, reltol_linear=1e-10, abstol_linear=0.0, maxiters_linear=2000)
    batch = vfvm.Batch(systems)
    Ub = batch.solve(0.0, ctrl)
    for k, ((s, o), U) in enumerate(zip(pairs, Ub)):
        Uo, _ = o.solve_step(np.zeros(s.grid.num_nodes))
        err = np.abs(vfvm.dofs(U) - Uo).max()
        assert err <= 1e-9, (k, err)
        assert batch.histories[k].residual_norm <= 1e-10
    # the same members one at a time through the single-system path
    for k in (0, 5, 11):
        Us = vfvm.solve(systems[k], inival=0.0, control=ctrl)
        assert np.abs(vfvm.dofs(Us) - vfvm.dofs(Ub[k])).max() <= 1e-9


def test_batch_member_failure_is_reported():
    print("case batch failure", flush=True)
    pairs = [cases.nonlinear_poisson_2d(9, 1e-2, 1.0, (0, 1, 1, 0)) for _ in range(3)]
    batch = vfvm.Batch([p[0] for p in pairs])
    with pytest.raises(vfvm.ConvergenceError):
        batch.solve(0.0, vfvm.SolverControl(maxiters=1))


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_batch_graph_replays_match_eager(precision):
    """The captured Krylov iteration blocks give the eager launches' results, on every
    replay, with other device work and synchronisations between the replays."""
    print("case batch graphs", precision, flush=True)
    members = [cases.batch_member(seed, sizes=(17, 33)) for seed in range(8)]
    pairs = [cases.nonlinear_poisson_2d(n, eps, f, tuple(v)) for (n, eps, v, f) in members]
    systems = [p[0] for p in pairs]
    ctrl = vfvm.SolverControl(precision=precision, reltol_linear=1e-10, abstol_linear=0.0, maxiters_linear=2000,
                              check_every=4)
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
        # eager work in between: a single-system solve on the same context
        vfvm.solve(systems[0], inival=0.0, control=ctrl)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_batch_device_resident_entry_point(precision):
    """vfvm_batch_newton_device (start values and solutions packed in device memory, member
    order) gives the host-buffer entry point's solutions and histories, on uneven members."""
    print("case batch device-resident", precision, flush=True)
    members = [cases.batch_member(seed, sizes=(17, 33)) for seed in range(6)]
    systems = [cases.nonlinear_poisson_2d(n, eps, f, tuple(v), oracle=False)[0] for (n, eps, v, f) in members]
    ctrl = vfvm.SolverControl(precision=precision, reltol_linear=1e-10, abstol_linear=0.0, maxiters_linear=2000)
    batch = vfvm.Batch(systems)
    rng = np.random.default_rng(3)
    U0 = [rng.uniform(-0.5, 0.5, (1, s.grid.num_nodes)) for s in systems]
    Uh = [vfvm.dofs(u).copy() for u in batch.solve(U0, ctrl)]
    hh = [(h.niter, h.nlin) for h in batch.histories]
    ndof, off = batch.sizes()
    assert ndof == off[-1] and list(np.diff(off)) == [s.grid.num_nodes for s in systems]
    ctx = systems[0].ctx
    Uold_dev = vfvm.DeviceArray.from_numpy(ctx, np.concatenate([u.ravel() for u in U0]))
    U_dev = vfvm.DeviceArray(ctx, ndof)
    batch.solve_device(U_dev, Uold_dev, ctrl)
    Ud = U_dev.numpy()
    assert [(h.niter, h.nlin) for h in batch.histories] == hh
    for m in range(len(systems)):
        assert np.array_equal(Ud[off[m]:off[m + 1]], Uh[m])
        assert batch.histories[m].residual_norm <= 1e-10


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_batch_sell_index_widths_agree(precision):
    """Batched SpMV with 16-bit column offsets and with 32-bit columns: bitwise the same
    solutions and iteration counts on uneven members (graphs recaptured on the switch)."""
    print("case batch sell index widths", precision, flush=True)
    members = [cases.batch_member(seed, sizes=(17, 33, 65)) for seed in range(6)]
    systems = [cases.nonlinear_poisson_2d(n, eps, f, tuple(v), oracle=False)[0] for (n, eps, v, f) in members]
    ctrl = vfvm.SolverControl(precision=precision, reltol_linear=1e-10, abstol_linear=0.0, maxiters_linear=2000)
    batch = vfvm.Batch(systems)
    out = []
    for bits in (16, 32, 16):
        batch.set_sell_index(bits)
        U = [vfvm.dofs(u).copy() for u in batch.solve(0.0, ctrl)]
        out.append((U, [h.nlin for h in batch.histories]))
    for U, it in out[1:]:
        assert it == out[0][1]
        for a, b in zip(U, out[0][0]):
            assert np.array_equal(a, b)
