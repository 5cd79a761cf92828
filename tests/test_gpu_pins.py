"""GPU checks of what the oracle comparison alone leaves open: the drift-diffusion
residual against the physics (thermodynamic equilibrium, the two-node boundary-value
problem), the Dirichlet initialisation (corner nodes: the last boundary face visited
wins, src/vfvm_system.jl:947-1003) and the Newton history (niter, updatenorm;
src/vfvm_history.jl:10-37) against the oracle's."""
import numpy as np
import pytest

import cases
import dd_pins
import oracle as O
import vfvm
from vfvm.grids import linspace, simplexgrid

pytestmark = pytest.mark.gpu


def _print(*a):
    print(*a, flush=True)


def _dd_system(g, zn=-1.0, zp=1.0):
    sys, _ = cases.pair(g, 3, dict(flux=vfvm.DriftDiffusionFlux(1e-2, zn, zp),
                                   reaction=vfvm.DriftDiffusionReaction(1.0)), oracle=False)
    return sys


@pytest.mark.parametrize("dim", [1, 2])
def test_dd_zero_bias_equilibrium_gpu(dim):
    """n = e^psi, p = e^-psi: the carrier rows of the HIP residual vanish to rounding for
    arbitrary psi (every Bernoulli branch is hit with psi in [-30, 30] on a coarse grid);
    with the charge numbers flipped they do not."""
    _print("case dd equilibrium", dim)
    g = simplexgrid(*([linspace(0, 1, 9)] * dim))
    rng = np.random.default_rng(21)
    _, ef = O.OracleSystem(g, 1, O.physics("linear", D0=1.0), {}, O.EDGEWISE).edges()
    for psi_range in (3.0, 30.0):
        U = dd_pins.equilibrium_state(g.num_nodes, rng, psi_range)
        F, _ = vfvm.evaluate_residual_and_jacobian(_dd_system(g), U)
        Ff, _ = vfvm.evaluate_residual_and_jacobian(_dd_system(g, zn=1.0, zp=-1.0), U)
        for ispec in (1, 2):
            scale = np.abs(U[ispec]).max() * np.abs(ef).max()
            err = np.abs(np.asarray(F)[ispec]).max()
            _print("psi range", psi_range, "species", ispec, "max residual", err, "scale", scale)
            assert err <= 1e-13 * scale
            assert np.abs(np.asarray(Ff)[ispec]).max() >= 1e-2 * scale


@pytest.mark.parametrize("psiK,psiL,nK,nL", [(0.3, -0.1, 1.2, 0.4), (-2.0, 1.5, 0.5, 3.0),
                                             (0.01, 0.0, 2.0, 1.0), (40.0, -30.0, 1e-3, 2.0)])
def test_dd_two_node_current_gpu(psiK, psiL, nK, nL):
    """Two-node 1D grid: the HIP carrier residuals equal +j and -j, the current of the
    exact 1D boundary-value problem (ODE integrated numerically)."""
    _print("case dd two-node", psiK, psiL, nK, nL)
    h = 0.37
    g = simplexgrid(np.array([0.0, h]))
    U = np.array([[psiK, psiL], [nK, nL], [1.0 / nK, 1.0 / nL]], order="F")
    F, _ = vfvm.evaluate_residual_and_jacobian(_dd_system(g), U)
    F = np.asarray(F)
    for ispec, z in ((1, -1.0), (2, 1.0)):
        j = dd_pins.sg_current_ode(U[ispec, 0], U[ispec, 1], psiK, psiL, z, h)
        sc = max(abs(j), abs(U[ispec]).max() / h)
        assert abs(F[ispec, 0] - j) <= 1e-9 * sc, (ispec, F[ispec, 0], j)
        assert abs(F[ispec, 1] + j) <= 1e-9 * sc, (ispec, F[ispec, 1], -j)


def test_initialize_dirichlet_matches_oracle():
    """_initialize_dirichlet! on the HIP path equals the oracle's bit for bit: four sides
    with distinct values (each corner node lies on two boundary regions, and the last
    boundary face visited sets it), lambda-dependent values, a Robin side left alone,
    and 3 species with a legacy Dirichlet condition."""
    _print("case initialize_dirichlet")
    X = linspace(0.0, 1.0, 9)
    g = simplexgrid(X, X)
    rng = np.random.default_rng(4)
    bcs = [vfvm.Dirichlet(species=1, region=1, value=0.25, value_lambda=0.5),
           vfvm.Dirichlet(species=1, region=2, value=0.5),
           vfvm.Robin(species=1, region=3, factor=2.0, value=7.0),
           vfvm.Dirichlet(species=1, region=4, value=1.0, value_lambda=-1.0)]
    sys, osys = cases.pair(g, 1, dict(flux=vfvm.LinearFlux(1.0)), bcs)
    for lam in (0.0, 0.7):
        U = rng.uniform(-3, 3, (1, g.num_nodes))
        Ug = vfvm.initialize_dirichlet(sys, U, embed=lam)
        Uo = osys.initialize_dirichlet(vfvm.dofs(U), embed=lam)
        assert np.array_equal(vfvm.dofs(Ug), Uo)
        corners = [k for k in range(g.num_nodes) if np.sum(np.isin(g.bfacenodes, k).any(axis=0)) > 0
                   and len(set(g.bfaceregions[np.isin(g.bfacenodes, k).any(axis=0)])) == 2]
        assert len(corners) == 4
        # a corner between two Dirichlet sides carries one of the two values, not the old U
        for k in corners:
            assert vfvm.dofs(Ug)[k] != U[0, k]
    sys, osys = cases.drift_diffusion_2d(9, lambda2=1e-2)
    sys.boundary_dirichlet(2, 1, 0.125)
    osys.bc[(1 - 1) * 3 + 1].kind = O.BC["dirichlet_legacy"]
    osys.bc[(1 - 1) * 3 + 1].value = 0.125
    O.lib().orc_set_bc(osys._h, osys.bc)
    U = np.asfortranarray(rng.uniform(0.1, 2.0, (3, g.num_nodes)))
    Ug = vfvm.initialize_dirichlet(sys, U, embed=0.3)
    Uo = osys.initialize_dirichlet(vfvm.dofs(U), embed=0.3)
    assert np.array_equal(vfvm.dofs(Ug), Uo)


def test_newton_history_matches_oracle():
    """cfg0 (101 x 101 nonlinear Poisson): the device Newton loop takes the oracle's
    number of iterations, and each iteration's update norm agrees with the oracle's
    (direct solves there, ILU(0) BiCGStab to 1e-13 here)."""
    _print("case newton history")
    sys, osys = cases.nonlinear_poisson_2d(101)
    Uo, ho = osys.solve_step(np.zeros(sys.grid.num_nodes))
    c = vfvm.SolverControl(reltol_linear=1e-13, abstol_linear=0.0, maxiters_linear=4000, precon="ilu0")
    vfvm.solve(sys, inival=0.0, control=c)
    h = sys.history
    uo = np.array(ho.updatenorm[:ho.niter])
    ug = np.array(h.updatenorm[:h.niter])
    _print("niter gpu", h.niter, "oracle", ho.niter)
    _print("updatenorm gpu", ug)
    _print("updatenorm oracle", uo)
    assert h.niter == ho.niter
    big = uo > 1e-6
    assert np.all(np.abs(ug[big] - uo[big]) <= 1e-7 * uo[big])
    assert np.all(np.abs(ug[~big] - uo[~big]) <= 1e-10)
