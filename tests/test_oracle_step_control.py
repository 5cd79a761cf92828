"""Pins the oracle's handle_exceptions step control (src/vfvm_solver.jl:359-455) by
brute force over single implicit-Euler steps (orc_solve_step, itself pinned by the
known-answer tests): a Newton solve that fails within maxiters is retried with half the
step, the accepted times are exactly those a step-by-step replay accepts, and a run that
cannot take even the smallest step returns its start instead of raising."""
import numpy as np
import pytest

import cases
import oracle as O
import vfvm
from vfvm.grids import linspace, simplexgrid


def _gaussian_case():
    """u_t = div(D(u) grad u), D(u) = 1e-2 + u^2, a Gaussian bump on 17 x 17 nodes,
    u = 0 on the boundary; large steps need more than five Newton iterations."""
    X = linspace(0.0, 1.0, 17)
    g = simplexgrid(X, X)
    x, y = g.coord
    u0 = np.exp(-30.0 * ((x - 0.5) ** 2 + (y - 0.5) ** 2))
    ph = vfvm.Physics(flux=vfvm.NonlinearDiffusionFlux(1e-2, 1.0), storage=vfvm.IdentityStorage())
    obcs = {(r, 0): ("dirichlet", 0.0, 0.0, 0.0) for r in (1, 2, 3, 4)}
    osys = O.OracleSystem(g, 1, cases.oracle_physics(ph, 1), obcs, O.EDGEWISE)
    return g, u0, osys


def _step_ok(osys, U, t, h, c):
    try:
        V, _ = osys.solve_step(U, c, time=t + h, tstep=h)
        return V
    except RuntimeError:
        return None


def test_newton_failure_raises_without_handle_exceptions():
    _, u0, osys = _gaussian_case()
    c = O.control(dt=0.5, dt_min=1e-4, dt_max=0.5, du_opt=10.0, maxiters=5)
    with pytest.raises(RuntimeError, match="error 1"):
        osys.solve_transient(u0, [0.0, 1.0], c)


def test_handle_exceptions_retries_failed_newton_steps():
    g, u0, osys = _gaussian_case()
    c = O.control(dt=0.5, dt_min=1e-4, dt_max=0.5, du_opt=10.0, maxiters=5, handle_exceptions=1)
    U, n, _, steps = osys.solve_transient(u0, [0.0, 1.0], c, return_steps=True)
    assert steps[-1] == pytest.approx(1.0, abs=1e-14)
    # the first accepted step is the first of 0.5, 0.25, ... whose Newton solve converges
    # within maxiters; every larger one fails
    h, tried = 0.5, []
    while _step_ok(osys, u0, 0.0, h, c) is None:
        tried.append(h)
        h *= 0.5
    assert len(tried) >= 2
    assert steps[0] == pytest.approx(h, rel=1e-14)
    # a replay along the accepted times: each step converges within maxiters and the
    # chain ends at the returned solution
    U_r, t = u0.copy(), 0.0
    for tk in steps:
        V = _step_ok(osys, U_r, t, tk - t, c)
        assert V is not None
        U_r, t = V, tk
    assert np.abs(U_r - U).max() <= 1e-13
    # accepted steps grow by at most dt_grow = 1.2
    h = np.diff(np.concatenate([[0.0], steps]))
    assert np.all(h[1:] <= 1.2 * h[:-1] * (1 + 1e-12))
    # with handle_exceptions on a run without failures the result is that of the plain run
    c0 = O.control(dt=1e-3, dt_min=1e-5, dt_max=0.5, du_opt=0.05)
    c1 = O.control(dt=1e-3, dt_min=1e-5, dt_max=0.5, du_opt=0.05, handle_exceptions=1)
    A, na, _, sa = osys.solve_transient(u0, [0.0, 0.2], c0, return_steps=True)
    B, nb, _, sb = osys.solve_transient(u0, [0.0, 0.2], c1, return_steps=True)
    assert na == nb and np.array_equal(sa, sb) and np.array_equal(A, B)


def test_handle_exceptions_gives_up_at_the_smallest_step():
    """D(u) = 1e-2 + u^2 with three Newton iterations: not even the smallest step
    converges; the run returns no steps and its start, Dirichlet values set."""
    g, u0, osys = _gaussian_case()
    c = O.control(dt=0.5, dt_min=0.1, dt_max=0.5, du_opt=10.0, maxiters=3, handle_exceptions=1)
    assert _step_ok(osys, u0, 0.0, 0.1, c) is None
    U, n, _, steps = osys.solve_transient(u0, [0.0, 1.0], c, return_steps=True)
    assert n == 0 and steps.size == 0
    expect = u0.copy()
    x, y = g.coord
    on_boundary = (x == 0.0) | (x == 1.0) | (y == 0.0) | (y == 1.0)
    expect[on_boundary] = 0.0
    assert np.array_equal(U, expect)
