"""Symmetric coordinate systems (polar, spherical, cylindrical form factors) against
the known answers of examples/Example203_CoordinateSystems.jl: the exact solutions of
-Laplace u = 1 on a disk, cylinder and sphere (the discretisation is exact there, the
reference asserts 1e-14), the quadratic convergence bounds on ring, cylinder shell and
sphere shell, and the unstructured cylinder mesh (assets/cyl_unstruct.sg, < 0.0012)."""
import numpy as np
import pytest

import cases
import vfvm
from vfvm.grids import circular_symmetric, simplexgrid, spherical_symmetric

pytestmark = pytest.mark.gpu

H = 0.1


def _print(*a):
    print(*a, flush=True)


def _solve(g, dirichlet, source):
    sys = vfvm.System(g, flux=vfvm.LinearFlux(1.0), species=[1],
                      source=np.full((1, g.num_nodes), source, order="F") if source else None)
    for reg, val in dirichlet:
        sys.boundary_dirichlet(1, reg, val)
    c = vfvm.SolverControl(reltol_linear=1e-15, abstol_linear=0.0, maxiters_linear=20000, precon="ilu0",
                           abstol=1e-15, reltol=1e-15, maxiters=20)
    return vfvm.dofs(vfvm.solve(sys, inival=0.0, control=c))


def _r(a, b):
    return np.array([a + H * k for k in range(int(round((b - a) / H)) + 1)])


def test_disk_and_sphere_are_exact():
    _print("case polar / spherical exact")
    r = _r(0.0, 5.0)
    u = _solve(circular_symmetric(simplexgrid(r)), [(2, 0.0)], 1.0)
    err = np.abs(u - 0.25 * (25.0 - r ** 2)).max()
    _print("disk", err)
    assert err < 1e-14
    u = _solve(spherical_symmetric(simplexgrid(r)), [(2, 0.0)], 1.0)
    err = np.abs(u - (25.0 - r ** 2) / 6.0).max()
    _print("sphere", err)
    assert err < 1e-14


def test_ring_and_sphere_shell_converge_quadratically():
    _print("case polar / spherical shells")
    r = _r(1.0, 5.0)
    u = _solve(circular_symmetric(simplexgrid(r)), [(1, 1.0), (2, 0.0)], 0.0)
    exact = (np.log(r) - np.log(5.0)) / (np.log(1.0) - np.log(5.0))
    e1 = np.abs(u - exact).max() / H ** 2
    u = _solve(spherical_symmetric(simplexgrid(r)), [(1, 1.0), (2, 0.0)], 0.0)
    exact = (5.0 * 1.0 / r - 1.0) / (5.0 - 1.0)
    e2 = np.abs(u - exact).max() / H ** 2
    _print("ring", e1, "sphere shell", e2)
    assert e1 < 0.01 and e2 < 0.04


def test_cylinder_exact_and_shell_quadratic():
    _print("case cylindrical")
    r, z = _r(0.0, 5.0), _r(0.0, 1.0)
    g = circular_symmetric(simplexgrid(r, z))
    u = _solve(g, [(2, 0.0)], 1.0)
    err = np.abs(u - 0.25 * (25.0 - g.coord[0] ** 2)).max()
    _print("cylinder", err)
    assert err < 1e-14
    g = circular_symmetric(simplexgrid(_r(1.0, 5.0), z))
    u = _solve(g, [(4, 1.0), (2, 0.0)], 0.0)
    exact = (np.log(g.coord[0]) - np.log(5.0)) / (np.log(1.0) - np.log(5.0))
    e = np.abs(u - exact).max() / H ** 2
    _print("cylinder shell", e)
    assert e < 0.01


def test_unstructured_cylinder():
    _print("case cyl_unstruct.sg")
    g = circular_symmetric(cases.sg_grid("cyl_unstruct.sg"))
    u = _solve(g, [(2, 0.0)], 1.0)
    err = np.abs(u - 0.25 * (25.0 - g.coord[0] ** 2)).max()
    _print("unstructured cylinder", err)
    assert err < 0.0012


def test_cylindrical_device_setup_equals_host_and_volume():
    """The cylindrical factors built on the GPU equal the host build bit for bit, and the
    node volumes add up to the volume of the solid of revolution (Pappus)."""
    _print("case cylindrical setup")
    g = circular_symmetric(cases.sg_grid("cyl_unstruct.sg"))
    ctx = vfvm.System(g, nspec=1).ctx
    vols = []
    try:
        for on_device in (False, True):
            ctx.set_setup(on_device)
            s = vfvm.System(g, flux=vfvm.LinearFlux(1.0), species=[1])
            s.boundary_dirichlet(1, 2, 0.0)
            s.handle()
            vols.append(s.nodevolumes())
    finally:
        ctx.set_setup("auto")
    assert np.array_equal(vols[0], vols[1])
    # total volume: the solid of revolution of the meshed region, 2 pi int r dA
    P = g.coord[:, g.cellnodes]
    area = 0.5 * np.abs((P[0, 1] - P[0, 0]) * (P[1, 2] - P[1, 0]) - (P[0, 2] - P[0, 0]) * (P[1, 1] - P[1, 0]))
    rc = P[0].mean(axis=0)
    assert vols[0].sum() == pytest.approx((2 * np.pi * area * rc).sum(), rel=1e-12)


def test_cylindrical_geometry_update_equals_fresh_system():
    """update_grid on a cylindrical system recomputes the cylindrical factors: bit for bit
    those of a cylindrical system created with the moved coordinates."""
    _print("case cylindrical update_grid")
    from test_gpu_geometry import _asm, _moved, _same
    g0 = circular_symmetric(cases.sg_grid("cyl_unstruct.sg"))
    s0 = vfvm.System(g0, flux=vfvm.LinearFlux(1.0), species=[1])
    s0.boundary_dirichlet(1, 2, 0.0)
    s0.set_compact_graph(False)
    g1 = _moved(g0, 0.15, 4)
    s0.update_grid(g1)
    s1 = vfvm.System(g1, flux=vfvm.LinearFlux(1.0), species=[1])
    s1.boundary_dirichlet(1, 2, 0.0)
    s1.set_compact_graph(False)
    U = np.random.default_rng(8).uniform(0.0, 1.0, (1, g0.num_nodes))
    assert _same(_asm(s0, U), _asm(s1, U))
    assert np.array_equal(s0.nodevolumes(), s1.nodevolumes())


def test_batch_of_cylindrical_systems_matches_single_solves():
    """Cylindrical systems (Example203's unstructured cylinder with different sources)
    solved as one batch give the single-system solutions."""
    _print("case cylindrical batch")
    g = circular_symmetric(cases.sg_grid("cyl_unstruct.sg"))
    systems = []
    for f in (1.0, 2.5, -0.7):
        s = vfvm.System(g, flux=vfvm.LinearFlux(1.0), reaction=vfvm.SinhReaction(), species=[1],
                        source=np.full((1, g.num_nodes), f, order="F"))
        s.boundary_dirichlet(1, 2, 0.0)
        systems.append(s)
    c = vfvm.SolverControl(reltol_linear=1e-12, abstol_linear=0.0, maxiters_linear=5000)
    Ub = vfvm.Batch(systems).solve(0.0, c)
    for s, ub in zip(systems, Ub):
        us = vfvm.solve(s, inival=0.0, control=c)
        assert np.abs(vfvm.dofs(ub) - vfvm.dofs(us)).max() <= 1e-9
