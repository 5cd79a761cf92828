"""Small-signal impedance (src/vfvm_impedance.jl, Example150_Impedance1D.jl) on the
GPU against the oracle (its assembly, a direct complex solve per frequency, the
test function and measurement rows formed from the oracle's matrices) and against
the closed-form impedance of the continuous Example150 problem."""
import math

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import cases
import oracle as O
import vfvm
from vfvm.grids import linspace, simplexgrid

pytestmark = pytest.mark.gpu

OMEGAS = 1e-3 * 1.1 ** np.arange(113)  # Example150: w from 1e-3 by factors 1.1 below 50
ZTOL = 1e-7  # impedance vs the oracle's direct solve (Krylov tolerance 1e-12 x condition)


def _print(*a):
    print(*a, flush=True)


def _problem(g, C=1.0, R=1.0, D=1.0, excited=1, meas=2):
    """C u_t - (D u_x)_x + R u = 0, u = 1 on region `excited`, u = 0 on region `meas`
    (Example150_Impedance1D.jl:38-88; storage coefficient C = 1)."""
    assert C == 1.0
    phys = dict(flux=vfvm.LinearFlux(D), reaction=vfvm.PowerReaction(m=1, k=R), storage=vfvm.IdentityStorage())
    sys, osys = cases.pair(g, 1, phys, legacy=[(1, excited, 1.0), (1, meas, 0.0)])
    return sys, osys, vfvm.Physics(**phys)


def _oracle(g, osys, ph, omegas, excited=1, meas=2):
    n = g.num_nodes
    U0, _ = osys.solve_step(np.zeros(n))
    osys.assemble(U0, U0, tstep=1e30)
    J0 = osys.matrix_scipy().tocsc()
    vol = osys.nodevolumes()
    b = np.zeros(n)
    for f, r in zip(g.bfacenodes.T, g.bfaceregions):
        if r == excited:
            for k in f:
                b[k] += 1e30
    tsys = O.OracleSystem(g, 1, O.physics("linear", D0=1.0),
                          {(excited, 0): ("dirichlet_legacy", 0.0), (meas, 0): ("dirichlet_legacy", 1.0)}, O.EDGEWISE)
    T, _ = tsys.solve_step(np.zeros(n))
    nsys = O.OracleSystem(g, 1, cases.oracle_physics(ph, 1), {}, O.EDGEWISE)
    nsys.assemble(U0, U0, tstep=math.inf)
    Jn = nsys.matrix_scipy().tocsc()
    dstdy = -(Jn.T @ T)
    dtran = -T * vol
    Z, UZ = [], []
    for w in omegas:
        x = spla.spsolve((J0 + 1j * w * sp.diags(vol)).tocsc(), b.astype(complex))
        UZ.append(x)
        Z.append(1.0 / (dstdy @ x + 1j * w * (dtran @ x)))
    return U0, T, dstdy, dtran, np.array(Z), UZ


def _gpu(sys, omegas, excited=1, meas=2):
    c = vfvm.SolverControl(reltol_linear=1e-13, abstol_linear=0.0, maxiters_linear=5000, precon="ilu0")
    U0 = vfvm.solve(sys, inival=0.0, control=c)
    T = vfvm.testfunction(sys, [excited], [meas])
    ds, dt = vfvm.measurement_derivatives(sys, U0, T, 1)
    isys = vfvm.ImpedanceSystem(sys, U0, 1, excited)
    Z = isys.impedances(omegas, ds, dt)
    return U0, T, ds, dt, Z, isys


def _rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("n", [41, 201])
def test_impedance_1d_matches_oracle_and_closed_form(n):
    _print("case impedance 1d", n)
    g = simplexgrid(linspace(0.0, 1.0, n))
    sys, osys, ph = _problem(g)
    U0o, To, dso, dto, Zo, UZo = _oracle(g, osys, ph, OMEGAS)
    U0, T, ds, dt, Z, isys = _gpu(sys, OMEGAS)
    _print("iterations", isys.iterations.min(), isys.iterations.max())
    assert _rel(vfvm.dofs(U0), U0o) <= 1e-9
    assert _rel(T, To) <= 1e-9
    assert _rel(ds, dso) <= 1e-9 and _rel(dt, dto) <= 1e-9
    err = np.abs(Z - Zo) / np.abs(Zo)
    _print("max rel err vs oracle", err.max())
    # BiCGStab stops at a relative residual of 1e-12 of the row-scaled system; the error
    # bound is that times its condition number, about 2 / (h^2 (R + w C)) <= 1e5 here
    assert err.max() <= ZTOL
    # one frequency's solution vector
    uz = isys.solve(OMEGAS[40])
    assert _rel(uz, UZo[40]) <= ZTOL
    # the continuous problem: 1 / (2 D z / (e^zL - e^-zL)), z = sqrt(i w C / D + R / D)
    zz = np.sqrt(1j * OMEGAS + 1.0)
    exact = 1.0 / (2.0 * zz / (np.exp(zz) - np.exp(-zz)))
    derr = np.abs(Z - exact) / np.abs(exact)
    _print("max rel err vs closed form", derr.max())
    assert derr.max() <= (2e-2 if n == 41 else 1e-3)


def test_impedance_unstructured_matches_oracle():
    _print("case impedance rz2d")
    g = cases.sg_grid("rz2d.sg")
    sys, osys, ph = _problem(g, D=0.5, R=2.0)
    om = OMEGAS[::8]
    _, _, ds, dt, Z, isys = _gpu(sys, om)
    _, _, dso, dto, Zo, _ = _oracle(g, osys, ph, om)
    assert _rel(ds, dso) <= 1e-9 and _rel(dt, dto) <= 1e-9
    err = np.abs(Z - Zo) / np.abs(Zo)
    _print("max rel err vs oracle", err.max(), "iterations", isys.iterations.max())
    assert err.max() <= ZTOL


def test_dudp_matches_oracle_difference():
    """dF/dlambda (bias value_lambda of drift-diffusion contacts; an embedded source)
    against the oracle's residual difference (F is affine in lambda)."""
    _print("case dudp")
    sys1, osys1 = cases.drift_diffusion_2d(9, lambda2=1e-2)
    X = linspace(0.0, 1.0, 17)
    g = simplexgrid(X, X)
    bcs = [vfvm.Dirichlet(species=1, region=1, value=0.0, value_lambda=0.5),
           vfvm.Dirichlet(species=1, region=3, value=1.0, value_lambda=-2.0)]
    src = np.full((1, g.num_nodes), 3.0, order="F")
    sys2, osys2 = cases.pair(g, 1, dict(flux=vfvm.LinearFlux(0.1), reaction=vfvm.SinhReaction()), bcs, source=src,
                             embed_source=True)
    for sys, osys in ((sys1, osys1), (sys2, osys2)):
        n = sys.grid.num_nodes * sys.num_species
        u = np.random.default_rng(2).uniform(0.3, 1.0, n)
        F1 = osys.assemble(u, u, tstep=math.inf, embed=0.2)
        F2 = osys.assemble(u, u, tstep=math.inf, embed=0.7)
        want = (F2 - F1) / 0.5
        got = vfvm.assemble_dudp(sys)
        pen = np.abs(want) > 1e20
        for mask in (pen, ~pen):
            if mask.any():
                sc = np.abs(want[mask]).max()
                assert np.abs(got[mask] - want[mask]).max() <= 1e-9 * sc


def test_parameter_excited_impedance_equals_boundary_excited():
    """Example151 (boundary value = the parameter, ImpedanceSystem(sys, U0; lambda0))
    and Example150 (legacy Dirichlet excitation) give the same impedance."""
    _print("case impedance parameter excitation")
    g = simplexgrid(linspace(0.0, 1.0, 101))
    sysA, _, _ = _problem(g)
    _, _, dsA, dtA, ZA, _ = _gpu(sysA, OMEGAS)
    phys = dict(flux=vfvm.LinearFlux(1.0), reaction=vfvm.PowerReaction(m=1, k=1.0), storage=vfvm.IdentityStorage())
    bcs = [vfvm.Dirichlet(species=1, region=1, value=1.0, value_lambda=1.0),
           vfvm.Dirichlet(species=1, region=2, value=0.0)]
    sysB, _ = cases.pair(g, 1, phys, bcs, oracle=False)
    c = vfvm.SolverControl(reltol_linear=1e-13, abstol_linear=0.0, maxiters_linear=5000, precon="ilu0")
    U0 = vfvm.solve(sysB, inival=0.0, control=c)
    T = vfvm.testfunction(sysB, [1], [2])
    ds, dt = vfvm.measurement_derivatives(sysB, U0, T, 1)
    isys = vfvm.ImpedanceSystem(sysB, U0, lambda0=0.0)
    ZB = isys.impedances(OMEGAS, ds, dt)
    err = np.abs(ZB - ZA) / np.abs(ZA)
    _print("max rel diff", err.max())
    assert err.max() <= 1e-10


def test_parameter_derivative_matches_central_difference():
    """du/dlambda of a stationary nonlinear Poisson solution (boundary value on region 2
    = 1 + 0.5 lambda) against the central difference of solutions at lambda = +-h."""
    _print("case parameter derivative")
    X = linspace(0.0, 1.0, 33)
    g = simplexgrid(X, X)
    phys = dict(flux=vfvm.LinearFlux(0.5), reaction=vfvm.SinhReaction())
    src = np.full((1, g.num_nodes), 1.0, order="F")
    c = vfvm.SolverControl(reltol_linear=1e-13, abstol_linear=0.0, maxiters_linear=5000, precon="ilu0")

    def build(v):
        bcs = [vfvm.Dirichlet(species=1, region=4, value=0.0),
               vfvm.Dirichlet(species=1, region=2, value=v, value_lambda=0.5)]
        return cases.pair(g, 1, phys, bcs, source=src, oracle=False)[0]

    h = 1e-3
    s0, sp_, sm = build(1.0), build(1.0 + 0.5 * h), build(1.0 - 0.5 * h)
    U0 = vfvm.solve(s0, inival=0.0, control=c)
    Up = vfvm.solve(sp_, inival=0.0, control=c)
    Um = vfvm.solve(sm, inival=0.0, control=c)
    fd = (vfvm.dofs(Up) - vfvm.dofs(Um)) / (2 * h)
    d = vfvm.parameter_derivative(s0, U0, 0.0, c)
    err = np.abs(d - fd).max() / np.abs(fd).max()
    _print("sensitivity vs central difference", err)
    assert err <= 1e-5


def test_testfunction_integral_matches_oracle_and_closed_form():
    """integrate(system, T, U) on Example150's steady state: against the oracle's residual
    without boundary conditions weighted by T, and against the exact boundary flux
    u'(1) = -1/sinh(1) of -u'' + u = 0, u(0) = 1, u(1) = 0."""
    _print("case testfunction integral")
    g = simplexgrid(linspace(0.0, 1.0, 201))
    sys, osys, ph = _problem(g)
    c = vfvm.SolverControl(reltol_linear=1e-13, abstol_linear=0.0, maxiters_linear=5000, precon="ilu0")
    U0 = vfvm.solve(sys, inival=0.0, control=c)
    T = vfvm.testfunction(sys, [1], [2])
    I = vfvm.integrate_testfunction(sys, T, U0)[0]
    nsys = O.OracleSystem(g, 1, cases.oracle_physics(ph, 1), {}, O.EDGEWISE)
    F = nsys.assemble(vfvm.dofs(U0), vfvm.dofs(U0), tstep=math.inf, embed=1.0)
    Io = float(T @ F)
    _print("integral", I, "oracle", Io, "exact", -1.0 / math.sinh(1.0))
    assert abs(I - Io) <= 1e-12 * abs(Io)
    assert abs(I + 1.0 / math.sinh(1.0)) <= 1e-3
    # an implicit-Euler step's integral adds the storage change: oracle with tstep
    Uold = 0.5 * vfvm.dofs(U0)
    It = vfvm.integrate_testfunction(sys, T, U0, Uold, 0.1)[0]
    Ft = nsys.assemble(vfvm.dofs(U0), Uold, tstep=0.1, embed=1.0)
    Ito = float(T @ Ft)
    _print("transient integral", It, "oracle", Ito)
    assert abs(It - Ito) <= 1e-12 * abs(Ito) and abs(It - I) > 1e-3
