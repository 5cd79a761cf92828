"""Device geometry updates (update_grid!, src/vfvm_system.jl:607-612, 690-755):
new coordinates on the same cells give, bit for bit, the factors, node volumes
and assembly of a system created with those coordinates, and the oracle's
assembly within 1e-12; a change that would alter the sparsity pattern is refused
and leaves the system as it was."""
import numpy as np
import pytest

import cases
import vfvm
from vfvm.grids import Grid, linspace, simplexgrid

pytestmark = pytest.mark.gpu


def _print(*a):
    print(*a, flush=True)


def _moved(g, scale, seed, boundary=False):
    """Interior nodes (on no boundary face; all nodes with boundary=True) moved by up to
    `scale` times the shortest cell edge in each coordinate."""
    rng = np.random.default_rng(seed)
    P = g.coord[:, g.cellnodes]
    h = min(np.sqrt(((P[:, a] - P[:, b]) ** 2).sum(axis=0)).min()
            for a in range(g.dim + 1) for b in range(a + 1, g.dim + 1))
    inner = np.ones(g.num_nodes, bool)
    if not boundary:
        inner[g.bfacenodes.ravel()] = False
    coord = g.coord.copy()
    coord[:, inner] += scale * h * rng.uniform(-1, 1, (g.dim, int(inner.sum())))
    return Grid(g.dim, coord, g.cellnodes, g.cellregions, g.bfacenodes, g.bfaceregions, g.shape, g.coordsys)


def _asm(sys, U):
    F, A = vfvm.evaluate_residual_and_jacobian(sys, np.asfortranarray(U))
    A = A.tocsc()
    return vfvm.dofs(F).copy(), A.data.copy(), A.indices.copy(), A.indptr.copy()


def _same(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a, b))


def _fan(m=70):
    """A centre node joined to m rim nodes (node graph row of m+1 entries: more than
    one 64-lane group); rim faces region 1 on the first half, region 2 on the rest."""
    t = 2 * np.pi * np.arange(m) / m
    coord = np.vstack([np.r_[0.0, np.cos(t)], np.r_[0.0, np.sin(t)]])
    k = np.arange(m)
    cells = np.vstack([np.zeros(m, np.int32), 1 + k, 1 + (k + 1) % m]).astype(np.int32)
    faces = np.vstack([1 + k, 1 + (k + 1) % m]).astype(np.int32)
    regs = np.where(k < m // 2, 1, 2).astype(np.int32)
    return Grid(2, coord, cells, np.ones(m, np.int32), faces, regs)


def _grids():
    X1 = np.sort(np.r_[0.0, 1.0, np.random.default_rng(2).uniform(0, 1, 30)])
    g1 = simplexgrid(X1)
    X = linspace(0.0, 1.0, 9)
    g3 = _moved(simplexgrid(X, X, X), 0.1, 7, boundary=True)  # generic: every edge factor nonzero
    return [("1d", g1, 0.2, False), ("cyl_unstruct", cases.sg_grid("cyl_unstruct.sg"), 0.15, False),
            ("kuhn9", g3, 0.1, True), ("fan70", _fan(), 0.3, False)]


@pytest.mark.parametrize("case", [0, 1, 2, 3])
def test_update_grid_equals_fresh_system(case):
    name, g0, scale, bnd = _grids()[case]
    _print("case update_grid", name)
    sys, _ = cases.nonlinear_poisson_on(g0)
    sys.set_compact_graph(False)
    rng = np.random.default_rng(3)
    U = rng.uniform(-1, 1, (1, g0.num_nodes))
    before = _asm(sys, U)
    vol0 = sys.nodevolumes()
    # the same coordinates: nothing changes
    sys.update_grid(g0)
    assert _same(_asm(sys, U), before)
    assert np.array_equal(sys.nodevolumes(), vol0)
    g1 = _moved(g0, scale, 11, bnd)
    sys.update_grid(g1)
    fresh, ofresh = cases.nonlinear_poisson_on(g1)
    fresh.set_compact_graph(False)
    got, want = _asm(sys, U), _asm(fresh, U)
    assert _same(got, want), "factors differ from a system created with the new coordinates"
    assert np.array_equal(sys.nodevolumes(), fresh.nodevolumes())
    assert not np.array_equal(got[0], before[0])
    cases.compare_assembly(sys, ofresh, U)
    # the compact factor table is rebuilt from the device factors on request
    sys.set_compact_graph(True)
    assert _same(_asm(sys, U), want)
    # and a Newton solve on the moved grid agrees with the oracle's
    Uo, _ = ofresh.solve_step(np.zeros(g1.num_nodes))
    c = vfvm.SolverControl(reltol_linear=1e-12, abstol_linear=1e-14, maxiters_linear=2000, precon="ilu0")
    Ug = vfvm.solve(sys, inival=0.0, control=c)
    err = np.abs(vfvm.dofs(Ug) - Uo).max()
    _print("newton on moved grid err", err, "res", sys.history.residual_norm)
    assert err <= 1e-9 and sys.history.residual_norm <= 1e-10


def test_update_grid_refuses_pattern_change():
    """A uniform 2D tensor grid has exactly-zero factors on its diagonals (no Jacobian
    entry); moving interior nodes makes them nonzero: refused, system unchanged.
    Scaling the grid by 2 keeps them exactly zero: accepted, equal to a fresh system."""
    _print("case update_grid pattern change")
    X = linspace(0.0, 1.0, 9)
    g0 = simplexgrid(X, X)
    sys, _ = cases.nonlinear_poisson_on(g0)
    sys.set_compact_graph(False)
    U = np.random.default_rng(4).uniform(-1, 1, (1, g0.num_nodes))
    before = _asm(sys, U)
    with pytest.raises(vfvm.VFVMError):
        sys.update_grid(_moved(g0, 0.1, 5))
    assert _same(_asm(sys, U), before)
    assert sys.grid is g0
    g2 = Grid(2, 2.0 * g0.coord, g0.cellnodes, g0.cellregions, g0.bfacenodes, g0.bfaceregions, g0.shape)
    sys.update_grid(g2)
    fresh, ofresh = cases.nonlinear_poisson_on(g2)
    fresh.set_compact_graph(False)
    assert _same(_asm(sys, U), _asm(fresh, U))
    cases.compare_assembly(sys, ofresh, U)
    # non-finite coordinates: refused, factors kept
    bad = g2.coord.copy()
    bad[0, 3] = np.nan
    with pytest.raises(vfvm.VFVMError):
        sys.update_grid(Grid(2, bad, g0.cellnodes, g0.cellregions, g0.bfacenodes, g0.bfaceregions, g0.shape))
    assert _same(_asm(sys, U), _asm(fresh, U))


def _setup_grids():
    X = linspace(0.0, 1.0, 9)
    X7 = linspace(0.0, 1.0, 7)
    return [("1d", _grids()[0][1]), ("tensor2d", simplexgrid(X, X)), ("kuhn7", simplexgrid(X7, X7, X7)),
            ("rz2d", cases.sg_grid("rz2d.sg")), ("cyl_unstruct", cases.sg_grid("cyl_unstruct.sg")),
            ("kuhn9_moved", _grids()[2][1]), ("fan70", _fan())]


@pytest.mark.parametrize("case", range(7))
def test_device_setup_equals_host_setup(case):
    """System creation with the node graph, form factors and node volumes built on the
    GPU gives bit for bit the system the host build gives: pattern, edge count, node
    volumes, residual and Jacobian; grids with exactly-zero factors (dropped edges)
    included.  A geometry update of the device-built system matches too."""
    name, g = _setup_grids()[case]
    _print("case device setup", name)
    ctx = vfvm.System(g, nspec=1).ctx
    out = []
    try:
        for on_device in (False, True):
            ctx.set_setup(on_device)
            sys, _ = cases.nonlinear_poisson_on(g)
            sys.handle()
            U = np.random.default_rng(6).uniform(-1, 1, (1, g.num_nodes))
            out.append((sys, sys.pattern(), sys.sizes(), sys.nodevolumes(), _asm(sys, U), U))
    finally:
        ctx.set_setup("auto")
    (sh, ph, zh, vh, ah, U), (sd, pd, zd, vd, ad, _) = out
    assert all(np.array_equal(a, b) for a, b in zip(ph, pd)) and zh == zd
    assert np.array_equal(vh, vd)
    assert _same(ah, ad)
    g1 = _moved(g, 0.1, 3, boundary=True) if name in ("kuhn9_moved", "fan70", "1d", "cyl_unstruct") else None
    if g1 is not None:
        sh.update_grid(g1)
        sd.update_grid(g1)
        assert _same(_asm(sh, U), _asm(sd, U))


def test_device_setup_rejects_bad_cells():
    _print("case device setup bad cells")
    X = linspace(0.0, 1.0, 5)
    g = simplexgrid(X, X)
    cn = g.cellnodes.copy()
    cn[1, 3] = g.num_nodes  # one past the last node
    bad = Grid(2, g.coord, cn, g.cellregions, g.bfacenodes, g.bfaceregions, g.shape)
    sys = vfvm.System(bad, flux=vfvm.LinearFlux(1.0), species=[1])
    with pytest.raises(vfvm.VFVMError):
        sys.handle()


def test_update_grid_from_device_memory():
    """The device-memory form gives the host form's factors bit for bit."""
    _print("case update_grid device memory")
    g0 = cases.sg_grid("cyl_unstruct.sg")
    g1 = _moved(g0, 0.15, 21)
    a, _ = cases.nonlinear_poisson_on(g0)
    b, _ = cases.nonlinear_poisson_on(g0)
    for s in (a, b):
        s.set_compact_graph(False)
    a.update_grid(g1)
    d = vfvm.DeviceArray.from_numpy(b.ctx, np.asfortranarray(g1.coord).ravel(order="F"))
    b.update_grid_device(d)
    U = np.random.default_rng(12).uniform(-1, 1, (1, g0.num_nodes))
    assert _same(_asm(a, U), _asm(b, U))
    assert np.array_equal(a.nodevolumes(), b.nodevolumes())
    # the system's grid holds the new coordinates: a rebuild of the device system keeps them
    assert np.array_equal(b.grid.coord, g1.coord)
    b._drop()
    b.set_compact_graph(False)
    assert _same(_asm(a, U), _asm(b, U))
    # arrays of another size or context are refused before any device work
    with pytest.raises(ValueError):
        b.update_grid_device(vfvm.DeviceArray(b.ctx, d.n - 1))
