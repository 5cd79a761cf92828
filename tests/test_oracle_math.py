"""Pins the oracle's factors and assembly by mathematics: closed forms,
invariants, finite-difference Jacobians and cellwise/edgewise agreement.
Each check is chosen so a dropped term, wrong sign or transposed index fails."""
import numpy as np
import pytest

import oracle as O
from vfvm.grids import linspace, simplexgrid


def test_node_volumes_sum_to_domain_measure():
    # test/test120_norms.jl:103-108,129
    for axes in ([linspace(0, 1, 7)], [linspace(0, 1, 5), linspace(0, 2, 4)],
                 [linspace(0, 1, 4), linspace(0, 1, 3), linspace(0, 3, 5)]):
        g = simplexgrid(*axes)
        meas = np.prod([a[-1] - a[0] for a in axes])
        for asm in (O.CELLWISE, O.EDGEWISE):
            s = O.OracleSystem(g, 1, O.physics("linear", D0=1.0), {}, asm)
            assert s.nodevolumes().sum() == pytest.approx(meas, rel=1e-14)


def test_triangle_factors_closed_form_and_embedded():
    # 2D factor (b^2+c^2-a^2)/(8A) vs the 3D boundary-triangle formula (test/test020_formfactors.jl:8-27)
    rng = np.random.default_rng(1)
    for _ in range(20):
        P = rng.random((3, 2))
        coord = P.ravel().copy()
        npar, epar = np.zeros(3), np.zeros(3)
        O.lib().orc_cellfactors(2, coord, np.array([0, 1, 2], np.int32), npar, epar)
        a2 = [np.sum((P[(i + 1) % 3] - P[(i + 2) % 3]) ** 2) for i in range(3)]
        u, v = P[1] - P[0], P[2] - P[0]
        area = 0.5 * abs(u[0] * v[1] - u[1] * v[0])
        for i in range(3):
            j, k = (i + 1) % 3, (i + 2) % 3
            assert epar[i] == pytest.approx((a2[j] + a2[k] - a2[i]) / (8 * area), rel=1e-12, abs=1e-12)
        assert npar.sum() == pytest.approx(area, rel=1e-12)
        P3 = np.hstack([P, np.zeros((3, 1))])
        bn = np.zeros(3)
        O.lib().orc_bfacefactors(3, P3.ravel().copy(), np.array([0, 1, 2], np.int32), bn)
        assert np.allclose(bn, npar, rtol=1e-7, atol=1e-12)


def test_kuhn_cube_factors():
    # uniform cube grid: interior axis edges sigma/h = h, diagonals ~ 0, interior node volume h^3
    n, h = 5, 0.25
    X = linspace(0, 1, n)
    g = simplexgrid(X, X, X)
    s = O.OracleSystem(g, 1, O.physics("linear", D0=1.0), {}, O.EDGEWISE)
    en, ef = s.edges()
    c = g.coord
    d = np.abs(c[:, en[:, 0]] - c[:, en[:, 1]])
    naxis = (d > 1e-12).sum(axis=0)
    interior = np.all((np.minimum(c[:, en[:, 0]], c[:, en[:, 1]]) > 1e-12) &
                      (np.maximum(c[:, en[:, 0]], c[:, en[:, 1]]) < 1 - 1e-12), axis=0)
    assert np.allclose(ef[(naxis == 1) & interior], h, rtol=1e-12)
    assert np.all(np.abs(ef[naxis > 1]) < 1e-14)
    vol = s.nodevolumes()
    inner = np.all((c > 1e-12) & (c < 1 - 1e-12), axis=0)
    assert np.allclose(vol[inner], h ** 3, rtol=1e-12)


def _fd_jacobian(s, U, **kw):
    n = U.size
    J = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n)
        hj = 1e-6 * max(1.0, abs(U[j]))
        e[j] = hj
        J[:, j] = (s.assemble(U + e, **kw) - s.assemble(U - e, **kw)) / (2 * hj)
    return J


CASES = [
    ("linear", dict(flux="linear", D0=0.7), 1, {}),
    ("power+pow", dict(flux="power", reaction="pow", D0=0.3, m=3, k=2.0, rm=2), 1, {}),
    ("nldiff+storage", dict(flux="nldiff", storage="identity", D0=1.0, d2=1.0), 1, {}),
    ("sinh", dict(flux="linear", reaction="sinh", D0=1e-2, k=1.0), 1, {}),
    ("exchange", dict(flux="linear", reaction="exchange", storage="identity", D0=0.5, D1=0.2, k=1.5), 2, {}),
    ("dd", dict(flux="dd", reaction="dd", lambda2=0.05, zn=-1.0, zp=1.0, ni2=1.0, ni=1.0), 3, {}),
]


@pytest.mark.parametrize("name,kw,ns,_", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("assembly", [O.CELLWISE, O.EDGEWISE])
def test_jacobian_matches_finite_differences(name, kw, ns, _, assembly):
    g = simplexgrid(linspace(0, 1, 4), linspace(0, 1, 3))
    rng = np.random.default_rng(7)
    s = O.OracleSystem(g, ns, O.physics(**kw), {(1, 0): ("dirichlet", 0.3), (2, 0): ("robin", 0.2, 1.5)},
                       assembly, rng.random(ns * g.num_nodes))
    U = 0.5 + rng.random(ns * g.num_nodes)
    if name == "dd":
        U.reshape(-1, 3)[:, 0] = rng.uniform(-3, 3, g.num_nodes)  # Bernoulli args across all branches
    tstep = 0.3 if kw.get("storage") else np.inf
    Uold = U * 0.9
    s.assemble(U, Uold, tstep=tstep)
    J = s.matrix_scipy().toarray()
    # FD on the penalty-free part: rows with a Dirichlet penalty carry 1e30
    pen = np.abs(np.diag(J)) > 1e20
    Jfd = _fd_jacobian(s, U, Uold=Uold, tstep=tstep)
    ok = ~pen
    scale = np.abs(J[ok]).max()
    assert np.abs(J[ok] - Jfd[ok]).max() <= 1e-6 * scale


def test_edge_conservation_and_symmetry():
    # each edge adds +g at K and -g at L: without sources, reactions or BCs sum(F) = 0
    g = simplexgrid(linspace(0, 1, 6), linspace(0, 1, 5))
    rng = np.random.default_rng(3)
    s = O.OracleSystem(g, 1, O.physics("nldiff", D0=1.0, d2=1.0), {}, O.EDGEWISE)
    F = s.assemble(rng.random(g.num_nodes))
    assert abs(F.sum()) < 1e-13
    # the linear diffusion Jacobian is symmetric with zero row sums
    s2 = O.OracleSystem(g, 1, O.physics("linear", D0=2.0), {}, O.EDGEWISE)
    s2.assemble(rng.random(g.num_nodes))
    A = s2.matrix_scipy().toarray()
    assert np.allclose(A, A.T, atol=1e-14)
    assert np.abs(A.sum(axis=1)).max() < 1e-13


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_cellwise_equals_edgewise(dim):
    # the reference's examples assert both assembly loops agree (e.g. Example105:96-101)
    axes = [linspace(0, 1, 5)] * dim
    g = simplexgrid(*axes)
    rng = np.random.default_rng(dim)
    kw = dict(flux="nldiff", reaction="sinh", storage="identity", D0=1.0, d2=0.5, k=1.0)
    U = rng.random(g.num_nodes)
    res = []
    for asm in (O.CELLWISE, O.EDGEWISE):
        s = O.OracleSystem(g, 1, O.physics(**kw), {(1, 0): ("dirichlet", 0.1)}, asm, rng.random(g.num_nodes))
        F = s.assemble(U, U * 0.5, tstep=0.1)
        res.append((F, s.matrix_scipy().toarray()))
    (F1, A1), (F2, A2) = res
    assert np.abs(F1 - F2).max() <= 1e-12 * np.abs(F1).max()
    assert np.abs(A1 - A2).max() <= 1e-12 * np.abs(A1).max()


def test_linear_problem_converges_in_one_step():
    g = simplexgrid(linspace(0, 1, 9), linspace(0, 1, 9))
    s = O.OracleSystem(g, 1, O.physics("linear", D0=1.0),
                       {(4, 0): ("dirichlet", 0.0), (2, 0): ("dirichlet", 1.0)}, O.EDGEWISE)
    U, h = s.solve_step(np.zeros(g.num_nodes), O.control(is_linear=1))
    assert h.niter == 1
    # u = x solves the Laplace problem with these BCs exactly (Neumann on y faces)
    assert np.abs(U - g.coord[0]).max() < 1e-12


def test_krylov_agrees_with_direct():
    # examples/DevEx003_Solvers.jl:85-177: every solver agrees with the direct one to 1e-7
    g = simplexgrid(linspace(0, 1, 21), linspace(0, 1, 21))
    src = np.ones(g.num_nodes)
    bcs = {(1, 0): ("dirichlet", 0.0), (4, 0): ("dirichlet", 0.0), (2, 0): ("dirichlet", 1.0), (3, 0): ("dirichlet", 1.0)}
    s = O.OracleSystem(g, 1, O.physics("linear", "sinh", D0=1e-2, k=1.0), bcs, O.EDGEWISE, src)
    ref, _ = s.solve_step(np.zeros(g.num_nodes))
    for m in ("bicgstab_jacobi", "bicgstab_ilu0", "gmres_blockjacobi"):
        U, _ = s.solve_step(np.zeros(g.num_nodes), O.control(method_linear=m, reltol_linear=1e-10,
                                                              abstol_linear=1e-13, maxiters_linear=2000))
        assert np.abs(U - ref).max() < 1e-7, m


def test_direct_solver_matches_scipy():
    import scipy.sparse.linalg as sla
    g = simplexgrid(linspace(0, 1, 9), linspace(0, 1, 7))
    rng = np.random.default_rng(5)
    s = O.OracleSystem(g, 3, O.physics("dd", "dd", lambda2=0.1, zn=-1.0, zp=1.0, ni2=1.0, ni=1.0),
                       {(4, 0): ("dirichlet", 0.2), (4, 1): ("dirichlet", 1.0), (4, 2): ("dirichlet", 1.0)})
    U = np.abs(rng.random(3 * g.num_nodes)) + 0.5
    F = s.assemble(U)
    x, _ = s.linsolve(F)
    y = sla.spsolve(s.matrix_scipy().tocsc(), F)
    assert np.abs(x - y).max() <= 1e-10 * np.abs(y).max()


def test_step_control_on_exact_decay():
    """solve_transient! step control (src/vfvm_solver.jl:359-500) on u' = -k u at every
    node (no flux, reaction k u, storage u): implicit Euler gives u_n = prod 1/(1 + k dt_i)
    over the accepted steps, so the accepted sequence the oracle reports must reproduce
    the final value; with the reference's defaults the first step (0.1) is rejected
    twice (du = 1/3, 1/5 > 1.2 * 0.1) and accepted at 0.025; accepted updates stay below
    du_max_factor * du_opt, growth stays below dt_grow, and the run ends at t = 1."""
    X = linspace(0.0, 1.0, 6)
    g = simplexgrid(X)
    k = 5.0
    s = O.OracleSystem(g, 1, O.physics("linear", "pow", "identity", D0=0.0, k=k, rm=1), {}, O.EDGEWISE)
    c = O.control()
    U, nsteps, _, steps = s.solve_transient(np.ones(g.num_nodes), [0.0, 1.0], c, return_steps=True)
    assert steps[-1] == 1.0
    dts = np.diff(np.concatenate([[0.0], steps]))
    assert dts[0] == pytest.approx(0.025, rel=1e-15)
    u = np.cumprod(1.0 / (1.0 + k * dts))
    assert np.abs(U - u[-1]).max() <= 1e-13
    du = np.abs(np.diff(np.concatenate([[1.0], u])))
    assert (du <= 1.2 * 0.1 + 1e-15).all()
    assert (dts[1:] <= 1.2 * dts[:-1] * (1 + 1e-12)).all()
    assert (dts >= 1e-3 * (1 - 1e-12)).all()


def test_fixed_timesteps_give_exact_sequence():
    """fixed_timesteps! (dt_min = dt_max = dt, du_opt = floatmax): exactly (t1 - t0) / dt
    steps of dt, whatever the update size."""
    X = linspace(0.0, 1.0, 6)
    g = simplexgrid(X)
    s = O.OracleSystem(g, 1, O.physics("linear", "pow", "identity", D0=0.0, k=50.0, rm=1), {}, O.EDGEWISE)
    c = O.control(dt=0.05, dt_min=0.05, dt_max=0.05, dt_grow=1.0, du_opt=np.finfo(float).max)
    U, nsteps, _, steps = s.solve_transient(np.ones(g.num_nodes), [0.0, 0.5], c, return_steps=True)
    assert nsteps == 10
    assert np.allclose(np.diff(np.concatenate([[0.0], steps])), 0.05, rtol=1e-12)
    assert np.abs(U - (1.0 / 3.5) ** 10).max() <= 1e-13


def test_embedded_source_scales_with_lambda():
    """Source embedding (p[12] != 0): the node source is lambda * f.  Only the source
    and the Dirichlet values (value_lambda) depend on lambda, so the residual is affine
    in lambda; at lambda = 1 it equals the system with the plain source and values; at
    lambda = 0 the solution is u = 0; continuation over [0, 1] ends at the direct
    solution of the lambda = 1 system."""
    g = simplexgrid(linspace(0, 1, 9), linspace(0, 1, 9))
    rng = np.random.default_rng(7)
    f = rng.uniform(-5.0, 5.0, g.num_nodes)
    vals = (0.3, -0.7, 0.9, -0.2)
    emb = O.OracleSystem(g, 1, O.physics("linear", "sinh", D0=1e-2, k=1.0, embed_source=1.0),
                         {(r + 1, 0): ("dirichlet", 0.0, 0.0, v) for r, v in enumerate(vals)}, O.EDGEWISE, f)
    plain = O.OracleSystem(g, 1, O.physics("linear", "sinh", D0=1e-2, k=1.0),
                           {(r + 1, 0): ("dirichlet", v) for r, v in enumerate(vals)}, O.EDGEWISE, f)
    U = rng.uniform(-1.0, 1.0, g.num_nodes)
    F0, F1, Fl = (emb.assemble(U, U, embed=l) for l in (0.0, 1.0, 0.37))
    free = np.abs(F1) < 1e20
    assert np.abs(Fl[free] - (F0[free] + 0.37 * (F1[free] - F0[free]))).max() <= 1e-14 * np.abs(F1[free]).max()
    assert np.array_equal(F1, plain.assemble(U, U, embed=0.0))
    Z, _ = emb.solve_step(np.zeros(g.num_nodes), embed=0.0)
    assert np.abs(Z).max() == 0.0
    ref, _ = plain.solve_step(np.zeros(g.num_nodes))
    c = O.control(dp=0.25, dp_min=1e-3, dp_max=0.5, dp_grow=1.5)
    Ue, nsteps, _ = emb.solve_transient(np.zeros(g.num_nodes), [0.0, 1.0], c, transient=False)
    assert nsteps >= 3
    assert np.abs(Ue - ref).max() <= 1e-9


def _sg(name):
    import os
    from vfvm.grids import read_sg
    return read_sg(os.path.join(os.path.dirname(__file__), "golden", name))


@pytest.mark.parametrize("name,counts", [("rz2d.sg", (335, 602, 66, 5)), ("cyl_unstruct.sg", (433, 761, 103, 2))])
def test_sg_grids_read_and_measure(name, counts):
    """The reference's unstructured grid assets (assets/rz2d.sg, assets/cyl_unstruct.sg):
    item counts as the files declare them, every cell non-degenerate, node volumes summing
    to the triangle areas (shoelace formula, computed here independently), every mesh
    edge owned by one cell listed as a boundary face, and the edge factors passing the
    Gauss identity sum_j (sigma_ij/h_ij)(x_i - x_j) = 0 at interior nodes (the discrete
    divergence of a linear function's flux vanishes; a wrong factor breaks it)."""
    g = _sg(name)
    assert (g.num_nodes, g.num_cells, g.num_bfaces, g.bfaceregions.max()) == counts
    assert g.cellnodes.min() == 0 and g.cellnodes.max() == g.num_nodes - 1
    P = g.coord[:, g.cellnodes]  # (2, 3, C)
    area = 0.5 * np.abs((P[0, 1] - P[0, 0]) * (P[1, 2] - P[1, 0]) - (P[0, 2] - P[0, 0]) * (P[1, 1] - P[1, 0]))
    assert area.min() > 0
    s = O.OracleSystem(g, 1, O.physics("linear", D0=1.0), {}, O.EDGEWISE)
    assert s.nodevolumes().sum() == pytest.approx(area.sum(), rel=1e-13)
    # every boundary node of the mesh (an edge in one cell only) is on some boundary face
    from collections import Counter
    cnt = Counter()
    for c in range(g.num_cells):
        n = sorted(g.cellnodes[:, c])
        for a, b in ((n[0], n[1]), (n[0], n[2]), (n[1], n[2])):
            cnt[(a, b)] += 1
    outer = {e for e, k in cnt.items() if k == 1}
    faces = {tuple(sorted(f)) for f in g.bfacenodes.T}
    assert outer <= faces
    # edge factors of a linear diffusion: the discrete operator annihilates constants and
    # reproduces a linear function's flux balance on interior nodes (consistency)
    en, ef = s.edges()
    bnodes = set(g.bfacenodes.ravel())
    x = g.coord[0]
    bal = np.zeros(g.num_nodes)
    np.add.at(bal, en[:, 0], ef * (x[en[:, 0]] - x[en[:, 1]]))
    np.add.at(bal, en[:, 1], ef * (x[en[:, 1]] - x[en[:, 0]]))
    inner = np.array([i for i in range(g.num_nodes) if i not in bnodes])
    assert np.abs(bal[inner]).max() <= 1e-11 * np.abs(ef).max()


def _dd_physics(zn=-1.0, zp=1.0, lambda2=1e-2):
    return O.physics("dd", "dd", lambda2=lambda2, zn=zn, zp=zp, ni2=1.0, ni=1.0)


@pytest.mark.parametrize("dim", [1, 2])
def test_dd_zero_bias_equilibrium(dim):
    """n = e^psi, p = e^-psi: every Scharfetter-Gummel carrier flux and the recombination
    vanish, so the carrier rows of F are zero to rounding for arbitrary psi.  The same
    state with the charge numbers flipped (equivalently B(x) and B(-x) swapped) is far
    from equilibrium, which shows the check sees such a mistake."""
    import dd_pins
    axes = [linspace(0, 1, 9)] * dim
    g = simplexgrid(*axes)
    rng = np.random.default_rng(11)
    U = dd_pins.equilibrium_state(g.num_nodes, rng)
    s = O.OracleSystem(g, 3, _dd_physics(), {}, O.EDGEWISE)
    _, ef = s.edges()
    scale = np.abs(U[1:]).max() * np.abs(ef).max()
    F = s.assemble(U.ravel(order="F")).reshape(3, -1, order="F")
    assert np.abs(F[1:]).max() <= 1e-13 * scale
    flipped = O.OracleSystem(g, 3, _dd_physics(zn=1.0, zp=-1.0), {}, O.EDGEWISE)
    Ff = flipped.assemble(U.ravel(order="F")).reshape(3, -1, order="F")
    assert np.abs(Ff[1:]).max() >= 1e-2 * scale


@pytest.mark.parametrize("psiK,psiL,nK,nL", [(0.3, -0.1, 1.2, 0.4), (-2.0, 1.5, 0.5, 3.0),
                                             (0.01, 0.0, 2.0, 1.0), (40.0, -30.0, 1e-3, 2.0)])
def test_dd_two_node_current_matches_ode(psiK, psiL, nK, nL):
    """A two-node 1D grid: the carrier rows of F are the edge currents of the exact
    1D boundary-value problem (integrated numerically), +j at K and -j at L; p = 1/n
    makes the recombination vanish, so only the flux is left."""
    import dd_pins
    h = 0.37
    g = simplexgrid(np.array([0.0, h]))
    U = np.array([[psiK, psiL], [nK, nL], [1.0 / nK, 1.0 / nL]], order="F")
    s = O.OracleSystem(g, 3, _dd_physics(), {}, O.EDGEWISE)
    F = s.assemble(U.ravel(order="F")).reshape(3, 2, order="F")
    for ispec, z in ((1, -1.0), (2, 1.0)):
        j = dd_pins.sg_current_ode(U[ispec, 0], U[ispec, 1], psiK, psiL, z, h)
        sc = max(abs(j), abs(U[ispec]).max() / h)
        assert abs(F[ispec, 0] - j) <= 1e-9 * sc, (ispec, F[ispec, 0], j)
        assert abs(F[ispec, 1] + j) <= 1e-9 * sc, (ispec, F[ispec, 1], -j)


def _five_point(n, rng):
    """Random diagonally dominant matrix with the 5-point pattern of an n x n grid (CSR)."""
    N = n * n
    rows = []
    for i in range(N):
        x, y = i % n, i // n
        cols = [c for c, ok in ((i - n, y > 0), (i - 1, x > 0), (i, True), (i + 1, x < n - 1), (i + n, y < n - 1))
                if ok]
        vals = [-rng.random() if c != i else 0.0 for c in cols]
        vals[cols.index(i)] = 4.5 + rng.random()
        rows.append((cols, vals))
    rp = np.cumsum([0] + [len(c) for c, _ in rows])
    ci = np.concatenate([c for c, _ in rows]).astype(np.int32)
    v = np.concatenate([v for _, v in rows])
    return rp, ci, v


def _dense_lu(rp, ci, lu):
    N = rp.size - 1
    L, U = np.eye(N), np.zeros((N, N))
    for i in range(N):
        for k in range(rp[i], rp[i + 1]):
            (L if ci[k] < i else U)[i, ci[k]] = lu[k]
    return L, U


def test_oracle_ilu0_is_exact_lu_on_tridiagonal():
    """ILU(0) drops no fill on a tridiagonal matrix, so it is the LU factorisation."""
    rng = np.random.default_rng(3)
    n = 12
    rp = np.array([0] + [min(i + 2, n) - max(i - 1, 0) for i in range(n)]).cumsum()
    ci = np.concatenate([np.arange(max(i - 1, 0), min(i + 2, n)) for i in range(n)]).astype(np.int32)
    v = rng.random(ci.size) + 3.0 * (ci == np.repeat(np.arange(n), np.diff(rp)))
    L, U = _dense_lu(rp, ci, O.ilu0(rp, ci, v))
    A = np.zeros((n, n))
    A[np.repeat(np.arange(n), np.diff(rp)), ci] = v
    assert np.abs(L @ U - A).max() <= 1e-14 * np.abs(A).max()


def test_oracle_ilu0_reproduces_the_matrix_on_its_pattern():
    """The defining property of ILU(0) (Saad, Iterative Methods, sec. 10.3.2): L unit
    lower and U upper on the pattern of A with (LU)_ij = a_ij for every stored (i, j);
    such factors are unique.  Off the pattern LU carries the dropped fill."""
    rng = np.random.default_rng(4)
    rp, ci, v = _five_point(7, rng)
    L, U = _dense_lu(rp, ci, O.ilu0(rp, ci, v))
    N = rp.size - 1
    A = np.zeros((N, N))
    mask = np.zeros((N, N), bool)
    r = np.repeat(np.arange(N), np.diff(rp))
    A[r, ci] = v
    mask[r, ci] = True
    P = L @ U
    assert np.abs(P[mask] - A[mask]).max() <= 1e-13 * np.abs(A).max()
    assert np.abs(P[~mask]).max() > 1e-3  # not the exact LU: fill was dropped
