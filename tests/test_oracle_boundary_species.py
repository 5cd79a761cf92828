"""Pins the oracle's boundary species, boundary-edge fluxes (bflux), boundary reaction /
storage / source and inactive-species rows:

* examples/Example220_NonlinearPoisson2D_BoundarySpecies.jl:118 known answer
  u5 = 0.0020781361856598 (boundary species coupled to two bulk species, bstorage);
* examples/Example230_BoundaryFlux.jl:218-223: a boundary species with bflux, breaction
  and bsource on one side of a square equals the same problem solved as a 1D bulk
  problem (< 1e-14);
* the Jacobian of every new term against finite differences of the residual, with
  the pattern of inactive rows (identity, src/vfvm_system.jl:1010-1026) checked.
"""
import numpy as np
import pytest

import oracle as O
from vfvm.grids import linspace, simplexgrid

JULIA_RTOL = 1.4901161193847656e-08  # isapprox's default, sqrt(eps)


def example220_oracle(n=10):
    """Example220's system on the oracle: species 1, 2 bulk (flux eps (u_k - u_l), storage
    u), species 3 a boundary species on region 2 with the breaction of lines 25-33 and
    bstorage u3 there; source exp(-20 |x - (0.5, 0.5)|^2) for species 1."""
    X = np.arange(n + 1) / n
    g = simplexgrid(X, X)
    k, eps = 1.0, 1e-2
    ph = O.physics("linear", "none", "identity", D0=eps, D1=eps, D2sp=0.0)
    x, y = g.coord
    src = np.zeros((3, g.num_nodes))
    src[0] = np.exp(-20.0 * ((x - 0.5) ** 2 + (y - 0.5) ** 2))
    react = [[k, 0.0, -k], [0.0, k, -k], [-k, -k, 2 * k]]
    s = O.OracleSystem(g, 3, ph, {}, O.EDGEWISE, src.T.ravel(),
                       species=[("bulk",), ("bulk",), ("boundary", [2])],
                       bphysics={2: dict(react=react, storage=[2])})
    return g, s


def test_example220_known_answer():
    g, s = example220_oracle()
    n = 10
    U = np.zeros(3 * g.num_nodes)
    c = O.control(reltol=1e-5, reltol_linear=1e-5)
    t = 0.0
    steps = 0
    while t < 1.0:
        t += 0.01
        U, h = s.solve_step(U, c, time=t, tstep=0.01)
        steps += 1
    assert steps == 100
    # the 5th node of the boundary subgrid of region 2 (x = 1, y = 0.4)
    u5 = U.reshape(-1, 3)[n + (n + 1) * 4, 2]
    assert u5 == pytest.approx(0.0020781361856598, rel=JULIA_RTOL)


def example230_pair(nn=20, d1=5.0, db=5.0, kmax=2.0, cmax=3.0):
    """The two problems of Example230 on the oracle: (1D system, its rcoef and source) and
    (2D system with the boundary species on region 2)."""
    X = linspace(0.0, 1.0, nn)
    half = nn // 2
    g1 = simplexgrid(X)
    k1 = np.where(np.arange(nn) < half, kmax, 0.0)
    c1 = np.where(np.arange(nn) < half, 0.0, cmax)
    s1 = O.OracleSystem(g1, 1, O.physics("linear", D0=d1), {}, O.EDGEWISE, c1)
    s1.set_rcoef(k1)
    g2 = simplexgrid(X, X)
    x, y = g2.coord
    src = np.zeros((2, g2.num_nodes))
    src[0] = 1.0  # c2
    rc = np.zeros((2, g2.num_nodes))
    rc[0] = 1.0  # k2
    bsrc = np.zeros((2, g2.num_nodes))
    bsrc[1] = np.where(y <= 0.5, 0.0, cmax)
    bcoef = np.zeros((2, g2.num_nodes))
    bcoef[1] = np.where(y <= 0.5, kmax, 0.0)
    s2 = O.OracleSystem(g2, 2, O.physics("linear", D0=1.0, D1=0.0), {}, O.EDGEWISE, src.T.ravel(),
                        species=[("bulk",), ("boundary", [2])],
                        bphysics={2: dict(bflux=[0.0, db], fields=True)},
                        bfields=(bsrc.T.ravel(), bcoef.T.ravel()))
    s2.set_rcoef(rc.T.ravel())
    east = np.nonzero(np.abs(x - 1.0) < 1e-12)[0]
    east = east[np.argsort(y[east])]
    return (g1, s1), (g2, s2), east


def test_example230_boundary_flux_equals_1d_problem():
    (g1, s1), (g2, s2), east = example230_pair()
    c = O.control(reltol_linear=1e-12)
    u1, _ = s1.solve_step(np.zeros(g1.num_nodes), c)
    u2, _ = s2.solve_step(np.zeros(2 * g2.num_nodes), c)
    ub = u2.reshape(-1, 2)[east, 1]
    assert np.abs(ub - u1).max() <= 1e-14 * np.abs(u1).max()
    # the boundary species lives on the east side only; elsewhere it keeps the start value 0
    mask = s2.node_species()
    assert np.array_equal(np.nonzero(mask & 2)[0], np.sort(east))
    assert np.all(u2.reshape(-1, 2)[(mask & 2) == 0, 1] == 0.0)


def _fd(s, U, Uold, tstep):
    n = U.size
    J = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n)
        h = 1e-6 * max(1.0, abs(U[j]))
        e[j] = h
        J[:, j] = (s.assemble(U + e, Uold, tstep=tstep) - s.assemble(U - e, Uold, tstep=tstep)) / (2 * h)
    return J


@pytest.mark.parametrize("dim", [2, 3])
def test_boundary_terms_jacobian_matches_finite_differences(dim):
    """Two bulk species and a boundary species on two boundary regions: breaction
    matrix, per-node coefficient and source, bstorage, bflux (3D: the boundary
    triangles' edge factors).  J = dF/dU by central differences; inactive rows are
    identity rows with F = U - Uold."""
    axes = [linspace(0, 1, 4), linspace(0, 1, 3)] + ([linspace(0, 1, 3)] if dim == 3 else [])
    g = simplexgrid(*axes)
    rng = np.random.default_rng(5)
    ph = O.physics("nldiff", "none", "identity", D0=0.5, d2=1.0)
    react = [[0.7, 0.0, -0.4], [0.0, 0.0, 0.0], [-0.3, 0.2, 1.1]]
    nb = g.num_bfaceregions
    bfields = (rng.random(3 * g.num_nodes), rng.random(3 * g.num_nodes))
    s = O.OracleSystem(g, 3, ph, {(1, 0): ("dirichlet", 0.2)}, O.EDGEWISE, rng.random(3 * g.num_nodes),
                       species=[("bulk",), ("bulk",), ("boundary", [2, nb])],
                       bphysics={2: dict(react=react, bflux=[0.0, 0.0, 1.3], storage=[2], fields=True),
                                 nb: dict(bflux=[0.0, 0.0, 0.6])},
                       bfields=bfields)
    U = 0.5 + rng.random(3 * g.num_nodes)
    Uold = 0.8 * U
    tstep = 0.25
    s.assemble(U, Uold, tstep=tstep)
    J = s.matrix_scipy().toarray()
    Jfd = _fd(s, U, Uold, tstep)
    pen = np.abs(np.diag(J)) > 1e20
    ok = ~pen
    assert np.abs(J[ok] - Jfd[ok]).max() <= 1e-6 * np.abs(J[ok]).max()
    mask = s.node_species()
    inactive = [K * 3 + 2 for K in range(g.num_nodes) if not (mask[K] >> 2) & 1]
    assert len(inactive) > 0
    F = s.assemble(U, Uold, tstep=tstep)
    for r in inactive:
        assert F[r] == U[r] - Uold[r]
        row = J[r]
        assert row[r] == 1.0 and np.count_nonzero(row) == 1
