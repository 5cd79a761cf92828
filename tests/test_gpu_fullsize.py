"""The benchmark configurations at their full sizes, checked through size-independent
properties (the oracle finishes only the smaller members in seconds):

* cfg1, 129^3 nonlinear diffusion: two implicit-Euler steps of 1e-2, every Newton solve at
  ||F||_inf <= 1e-10, the FP32-refinement form equal to the FP64 form to 1e-9, the
  solution inside [0, max u0] (a maximum principle of the diffusion equation) and
  decaying;
* cfg3, 256 seeded nonlinear-Poisson instances of N in {33, 65, 129, 257}: one batched
  solve, every member at ||F||_inf <= 1e-10, eight members of N <= 129 against the oracle
  within 1e-9, and the members of one size against their single-system solves;
* cfg2's bias sweep as configured, 0 -> 1 V in 20 steps, GMRES(30) + 3x3 block Jacobi with
  FP32 inside FP64 refinement, on 65 x 65 nodes against the oracle."""
import numpy as np
import pytest

import cases
import oracle as O
import vfvm
from vfvm import problems

pytestmark = pytest.mark.gpu


def _print(*a):
    print(*a, flush=True)


def test_cfg1_full_size_two_steps():
    _print("case cfg1 129^3")
    sys = problems.nonlinear_diffusion_3d(129)
    u0 = problems.initial_sine_3d(sys.grid).reshape(1, -1)
    times = np.array([0.0, 1e-2, 2e-2])
    out = {}
    for precision in ("fp64", "fp32"):
        ctrl = vfvm.fixed_timesteps(vfvm.SolverControl(precon="ilu0", precision=precision, reltol_linear=1e-10,
                                                       abstol_linear=0.0, maxiters_linear=2000), 1e-2)
        sol = vfvm.solve(sys, inival=np.asfortranarray(u0), control=ctrl, times=times)
        assert sys.history.residual_norm <= 1e-10
        out[precision] = [vfvm.dofs(u).copy() for u in sol.u]
        _print(precision, "newton", sys.history.niter, "krylov", sys.history.nlin)
    for a, b in zip(out["fp64"], out["fp32"]):
        assert np.abs(a - b).max() <= 1e-9
    u1, u2 = out["fp64"][1], out["fp64"][2]
    assert u2.min() >= -1e-12 and u2.max() <= u0.max() + 1e-12
    assert u2.max() < u1.max() < u0.max()


def test_cfg3_full_size_batch():
    _print("case cfg3 256 members")
    members = [problems.batch_member(seed) for seed in range(256)]
    systems = [problems.nonlinear_poisson_2d(n, eps, f, tuple(v)) for (n, eps, v, f) in members]
    ndof = sum(s.grid.num_nodes for s in systems)
    _print("members", len(systems), "unknowns", ndof)
    ctrl = vfvm.SolverControl(precision="fp32", reltol_linear=1e-6, abstol_linear=0.0, maxiters_linear=6000,
                              inner_reltol=1e-2, max_refinements=12, check_every=8)
    batch = vfvm.Batch(systems)
    U = batch.solve(0.0, ctrl)
    worst = max(h.residual_norm for h in batch.histories)
    _print("worst residual", worst)
    assert worst <= 1e-10
    picked = [k for k, m in enumerate(members) if m[0] <= 129][:8]
    for k in picked:
        n, eps, v, f = members[k]
        _, o = cases.nonlinear_poisson_2d(n, eps, f, tuple(v))
        Uo, _ = o.solve_step(np.zeros(systems[k].grid.num_nodes))
        err = np.abs(vfvm.dofs(U[k]) - Uo).max()
        _print("member", k, "N", n, "err vs oracle", err)
        assert err <= 1e-9
    big = [k for k, m in enumerate(members) if m[0] == 257][:2]
    for k in big:
        Us = vfvm.solve(systems[k], inival=0.0, control=ctrl)
        assert np.abs(vfvm.dofs(Us) - vfvm.dofs(U[k])).max() <= 1e-9


def test_cfg2_configured_sweep_65():
    """0 -> 1 V in 20 continuation steps with the configured solver"""
    _print("case cfg2 sweep 65")
    sys, osys = cases.drift_diffusion_2d(65, lambda2=1e-4)
    U0 = cases.dd_initial(sys.grid)
    lam = np.linspace(0.0, 1.0, 21)
    c = O.control(method_linear="direct", dp=0.05, dp_min=0.05, dp_max=0.05, du_opt=10.0)
    Uo, _, _ = osys.solve_transient(vfvm.dofs(U0), lam, c, transient=False)
    ctrl = vfvm.SolverControl(krylov="gmres", precon="blockjacobi", precision="fp32", gmres_restart=30,
                              reltol_linear=1e-10, abstol_linear=0.0, maxiters_linear=6000, Δp=0.05, Δp_min=0.05,
                              Δp_max=0.05, Δu_opt=10.0)
    U = vfvm.solve(sys, inival=U0, control=ctrl, embed=lam).u[-1]
    err = np.abs(vfvm.dofs(U) - Uo).max()
    _print("cfg2 sweep 65: err", err, "fp64 fallbacks (last step)", sys.history.fp64_fallbacks)
    assert err <= 1e-9
    assert sys.history.residual_norm <= 1e-10
