"""Pins the CPU oracle to the reference's own known answers.

Each case restates one of the reference's tests/examples and asserts its
`testval` with the reference's tolerance (Julia `≈`, rtol = sqrt(eps)).
"""
import numpy as np
import pytest

import oracle as O
from vfvm.grids import linspace, simplexgrid

RTOL = 1.4901161193847656e-08  # Julia isapprox default: sqrt(eps(Float64))


@pytest.mark.parametrize("assembly", [O.CELLWISE, O.EDGEWISE])
def test_example105_nonlinear_poisson_1d(assembly):
    # examples/Example105_NonlinearPoisson1D.jl:29-101
    X = np.arange(11) / 10.0
    g = simplexgrid(X)
    ph = O.physics("linear", "sinh", D0=1e-3, k=1.0)
    src = np.where(X <= 0.5, 1.0, -1.0)
    s = O.OracleSystem(g, 1, ph, {(1, 0): ("dirichlet_legacy", 0.0), (2, 0): ("dirichlet_legacy", 1.0)},
                       assembly, src)
    U, _ = s.solve_step(np.full(11, 0.5))
    assert U.sum() == pytest.approx(1.5247901344230088, rel=RTOL)


def _barenblatt(x, t, m):
    # examples/Example106_NonlinearDiffusion1D.jl:30-38
    tx = t ** (-1.0 / (m + 1.0))
    xx = (x * tx) ** 2
    xx = 1 - xx * (m - 1) / (2.0 * m * (m + 1))
    return tx * max(xx, 0.0) ** (1.0 / (m - 1.0))


@pytest.mark.parametrize("assembly", [O.CELLWISE, O.EDGEWISE])
def test_example106_nonlinear_diffusion_1d(assembly):
    # examples/Example106_NonlinearDiffusion1D.jl:40-118 (implicit Euler, fixed step)
    X = np.array([(-10 + k) / 10.0 for k in range(21)])
    g = simplexgrid(X)
    s = O.OracleSystem(g, 1, O.physics("power", storage="identity", D0=1.0, m=2), {}, assembly)
    inival = np.array([_barenblatt(x, 0.001, 2.0) for x in X])
    c = O.control(dt=1e-4, dt_min=1e-4, dt_max=1e-4, du_opt=1.0)
    U, nsteps, _ = s.solve_transient(inival, [0.001, 0.01], c)
    assert nsteps == 90
    assert U.sum() == pytest.approx(46.66666666647518, rel=RTOL)


@pytest.mark.parametrize("assembly", [O.CELLWISE, O.EDGEWISE])
def test_example207_nonlinear_poisson_2d(assembly):
    # examples/Example207_NonlinearPoisson2D.jl:18-77 (100 implicit Euler steps via tstep)
    X = np.arange(11) / 10.0
    g = simplexgrid(X, X)
    x, y = g.coord
    src = np.exp(-20.0 * ((x - 0.5) ** 2 + (y - 0.5) ** 2))
    s = O.OracleSystem(g, 1, O.physics("power", "pow", "identity", D0=1e-2, m=2, k=1.0, rm=2),
                       {(2, 0): ("dirichlet_legacy", 0.1), (4, 0): ("dirichlet_legacy", 0.1)}, assembly, src)
    U = np.full(g.num_nodes, 0.5)
    time, tstep = 0.0, 0.01
    while time < 1.0:
        time = time + tstep
        U, _ = s.solve_step(U, tstep=tstep, time=time)
    assert U[14] == pytest.approx(0.3554284760906605, rel=RTOL)


@pytest.mark.parametrize("assembly", [O.CELLWISE, O.EDGEWISE])
def test_example210_two_species_2d(assembly):
    # examples/Example210_NonlinearPoisson2D_Reaction.jl:11-84
    X = np.arange(11) / 10.0
    g = simplexgrid(X, X)
    x, y = g.coord
    src = np.zeros((g.num_nodes, 2))
    src[:, 0] = np.exp(-20 * ((x - 0.5) ** 2 + (y - 0.5) ** 2))
    s = O.OracleSystem(g, 2, O.physics("linear", "exchange", "identity", D0=1e-2, D1=1e-2, k=1.0), {},
                       assembly, src)
    c = O.control(dt=0.01, dt_min=0.01, dt_max=0.01, du_opt=1e5)
    U, nsteps, _ = s.solve_transient(np.zeros(2 * g.num_nodes), [0.0, 1.0], c)
    assert nsteps == 100
    assert U.sum() == pytest.approx(16.01812472041518, rel=RTOL)


@pytest.mark.parametrize("assembly", [O.CELLWISE, O.EDGEWISE])
def test_example301_laplace_3d(assembly):
    # examples/Example301_Laplace3D.jl:13-50: pins tet factors and 3D boundary regions 5/6
    X = np.arange(6) / 5.0
    g = simplexgrid(X, X, X)
    x, y, z = g.coord
    s = O.OracleSystem(g, 1, O.physics("linear", D0=1.0),
                       {(5, 0): ("dirichlet_legacy", 0.0), (6, 0): ("dirichlet_legacy", 0.0)}, assembly,
                       x * np.sin(5.0 * y) * np.exp(z))
    U, _ = s.solve_step(np.zeros(g.num_nodes))
    assert U[42] == pytest.approx(0.012234524449380824, rel=RTOL)


def test_example440_parallel_state():
    # examples/Example440_ParallelState.jl:26-49: 100 solves with different influx
    X = linspace(0.0, 1.0, 10 * 2 ** 5 + 1)
    g = simplexgrid(X)
    total = 0.0
    for influx in linspace(0.0, 10.0, 100):
        s = O.OracleSystem(g, 1, O.physics("power", D0=1.0, m=2),
                           {(1, 0): ("dirichlet", 0.1), (2, 0): ("robin", influx, 0.0)}, O.EDGEWISE)
        U, _ = s.solve_step(np.full(g.num_nodes, 0.1), O.control(abstol=1e-15, reltol=1e-20))
        total += (s.nodevolumes() * U).sum()  # integrate(sys, sol)[1, 1]
    assert total == pytest.approx(140.79872772042577, rel=RTOL)


def test_test030_linear_exactness():
    # test/test030_cachesol.jl:21-47: A x = -F(0) reproduces the linear solution x.
    # The reference's bound 3e-16 is UMFPACK's rounding; the banded LU lands within 1e-15.
    X = linspace(0.0, 1.0, 10)
    g = simplexgrid(X)
    s = O.OracleSystem(g, 1, O.physics("linear", D0=1.0), {(1, 0): ("dirichlet", 0.0), (2, 0): ("dirichlet", 1.0)})
    F = s.assemble(np.zeros(10))
    x, _ = s.linsolve(-F)
    assert np.abs(x - X).max() <= 1e-15


def test_bernoulli_goldens():
    # test/test010_bernoulli.jl:5-22 against mpmath goldens, max abs error < 1e-14
    import os
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "bernoulli.npz"))
    x, b, db = d["x"], d["b"], d["db"]
    err = max(abs(O.bernoulli(v) - w) for v, w in zip(x, b))
    assert err < 1e-14
    pm = [O.bernoulli_pm(v) for v in x]
    assert max(abs(p[0] - w) for p, w in zip(pm, b)) < 1e-14
    bneg = [O.bernoulli(-v) for v in x]
    assert max(abs(p[1] - w) for p, w in zip(pm, bneg)) < 1e-14
    # dual derivative vs the BigFloat derivative of pluto-examples/bernoulli.jl:45-51
    mask = np.abs(x) <= 50
    rel = [abs(O.bernoulli_dual(v)[1] - w) / abs(w) for v, w in zip(x[mask], db[mask]) if w != 0]
    assert max(rel) < 1e-12


def _example405_matrix(X, D=1.0e-2, v=1.0):
    """Example405's generic operator as its matrix: on each interval i, du = D (u_i -
    u_{i+1}) / h_i + v u_i (upwind, v > 0) is added to row i and subtracted from row i + 1
    (examples/Example405_GenericOperator.jl:21-32); the operator is linear in u."""
    n = X.size
    G = np.zeros((n, n))
    for i in range(n - 1):
        h = X[i + 1] - X[i]
        G[i, i] += D / h + v
        G[i, i + 1] -= D / h
        G[i + 1, i] -= D / h + v
        G[i + 1, i + 1] += D / h
    return G


def test_example405_generic_operator():
    # examples/Example405_GenericOperator.jl:57-62: testval = sum of the solution
    X = np.linspace(0.0, 1.0, 11)
    g = simplexgrid(X)
    s = O.OracleSystem(g, 1, O.physics(), {(1, 0): ("dirichlet_legacy", 0.0), (2, 0): ("dirichlet_legacy", 1.0)})
    s.set_generic(_example405_matrix(X))
    U, _ = s.solve_step(np.full(g.num_nodes, 0.5), O.control())
    assert U.sum() == pytest.approx(1.099999999614456, rel=1e-12)
    # F = G U at a state (the boundary rows carry the penalty terms on top)
    rng = np.random.default_rng(5)
    u = rng.random(g.num_nodes)
    F = s.assemble(u, u)
    A = s.matrix_scipy().toarray()
    G = _example405_matrix(X)
    inner = slice(1, g.num_nodes - 1)
    assert np.abs(F[inner] - (G @ u)[inner]).max() <= 1e-14
    assert np.abs(A[inner] - G[inner]).max() <= 1e-14
