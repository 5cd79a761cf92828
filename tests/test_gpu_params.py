"""Registry coefficients that depend on the system parameter lambda, with dF/dlambda by
dual-number parameters (the reference's params lanes, src/vfvm_assembly.jl:116-120,
194-198):

* examples/Example430_ParameterDerivativesStationary.jl:258 known answer 489.3432830184927
  (the sum of dI/dp over p in 0.1:0.05:2), computed as the reference's runh does: dI/dp =
  (dI/du) du/dp + dI/dp|_u with du/dp = -J^-1 dF/dp;
* dF/dlambda against the oracle's differences of F in lambda: exact for the coefficients
  that enter linearly, to 1e-7 for the drift-diffusion charge numbers inside the
  Bernoulli functions."""
import numpy as np
import pytest

import cases
import oracle as O
import vfvm
from vfvm.grids import simplexgrid

pytestmark = pytest.mark.gpu

TESTVAL = 489.3432830184927
P = 0.1 + 0.05 * np.arange(39)


def _print(*a):
    print(*a, flush=True)


def example430_system(n=10):
    X = np.arange(n + 1) / n
    g = simplexgrid(X, X)
    ph = vfvm.Physics(flux=vfvm.PowerFlux(2, 1.0), dflux=vfvm.PowerFlux(2, 1.0),  # D(p) = 1 + p
                      reaction=vfvm.PowerReaction(5, 0.0), dreaction=vfvm.PowerReaction(5, 1.0),  # k(p) = p
                      bcondition=[vfvm.Dirichlet(species=1, region=1, value=0.0),
                                  vfvm.Dirichlet(species=1, region=3, value=0.0, value_lambda=1.0)])
    return g, vfvm.System(g, ph, species=[1])


def test_example430_known_answer_on_gpu():
    _print("case Example430")
    g, sys = example430_system()
    c = vfvm.SolverControl(precon="ilu0", reltol_linear=1e-13, abstol_linear=0.0, maxiters_linear=2000)
    T = vfvm.testfunction(sys, [1], [3])
    total = 0.0
    for p in P:
        U = vfvm.solve(sys, inival=0.5, control=c, params=[p])
        assert sys.history.residual_norm <= 1e-10
        dudp = vfvm.parameter_derivative(sys, U, p, c)
        dstdy, _ = vfvm.measurement_derivatives(sys, U, T, 1, params=[p])  # d(-I)/du
        dIdp_p = float(T @ vfvm.assemble_dudp_at(sys, U, p, with_bc=False))
        total += -float(dstdy @ vfvm.dofs(dudp)) + dIdp_p
    _print("sum dI/dp", repr(total), "testval", TESTVAL)
    assert total == pytest.approx(TESTVAL, rel=1.4901161193847656e-08)


def test_dudp_at_matches_oracle_differences():
    _print("case dudp_at vs oracle")
    g, sys = example430_system()
    import test_oracle_params as TP
    _, osys, onobc, _ = TP.example430_oracle()
    U = np.random.default_rng(7).uniform(0.2, 1.0, (1, g.num_nodes))
    lam = 0.7
    # Dirichlet rows: region 1 (value independent of lambda), region 3 (value_lambda = 1)
    dir1 = np.unique(g.bfacenodes[:, g.bfaceregions == 1])
    dir3 = np.unique(g.bfacenodes[:, g.bfaceregions == 3])
    free = np.ones(g.num_nodes, bool)
    free[dir1] = free[dir3] = False
    d_nobc = vfvm.assemble_dudp_at(sys, U, lam, with_bc=False)
    for with_bc, o in ((True, osys), (False, onobc)):
        d = vfvm.assemble_dudp_at(sys, U, lam, with_bc=with_bc)
        # F is affine in lambda here: a difference over a unit step is exact up to rounding
        ref = o.assemble(vfvm.dofs(U), embed=lam + 0.5) - o.assemble(vfvm.dofs(U), embed=lam - 0.5)
        rows = free if with_bc else np.ones(g.num_nodes, bool)
        sc = np.abs(ref[rows]).max()
        _print("with_bc", with_bc, "max err", np.abs(d[rows] - ref[rows]).max() / sc)
        assert np.abs(d[rows] - ref[rows]).max() <= 1e-10 * sc
    # on the penalty rows the difference of two 1e30-sized residuals keeps only the
    # penalty part; the dual lanes give the exact derivative: the penalty term's
    # (1e30 * fac * value_lambda on region 3, 0 on region 1) plus the bulk terms'
    d = vfvm.assemble_dudp_at(sys, U, lam, with_bc=True)
    assert np.array_equal(d[dir1], d_nobc[dir1])
    ref = osys.assemble(vfvm.dofs(U), embed=lam + 0.5) - osys.assemble(vfvm.dofs(U), embed=lam - 0.5)
    assert np.all(np.abs(ref[dir3]) > 1e20)
    assert np.allclose(d[dir3], ref[dir3], rtol=1e-10)


def test_dudp_at_drift_diffusion_charges_match_oracle():
    """lambda in lambda^2, in the charge numbers (inside the Bernoulli functions) and in
    n_i: dF/dlambda by dual parameters against central differences of the oracle."""
    _print("case dudp_at drift-diffusion")
    sys, osys = cases.drift_diffusion_2d(17, lambda2=1e-2)
    dp = np.zeros(16)
    dp[7], dp[8], dp[9], dp[10], dp[11] = 0.01, -0.3, 0.2, 0.5, 0.25
    sys.physics.dflux = vfvm.DriftDiffusionFlux(dp[7], dp[8], dp[9])
    sys.physics.dreaction = vfvm.DriftDiffusionReaction(0.0)
    sys._drop()
    # DriftDiffusionReaction(ni) fills n_i^2 and n_i; set the derivative slots directly
    h = sys.handle()
    from vfvm import _capi
    _capi.check(_capi.lib().vfvm_system_set_param_dependence(h, _capi.dptr(np.ascontiguousarray(dp))))
    osys.set_param_dependence(dp)
    rng = np.random.default_rng(3)
    U = cases.dd_initial(sys.grid)
    U[0] += rng.uniform(-1.5, 1.5, sys.grid.num_nodes)
    lam, eps = 0.4, 1e-6
    d = vfvm.assemble_dudp_at(sys, U, lam, with_bc=False)
    # the oracle's F without boundary conditions
    onobc = O.OracleSystem(sys.grid, 3, cases.oracle_physics(sys.physics, 3), {}, O.EDGEWISE,
                           np.asarray(sys._source_array()).reshape(-1, order="F"))
    onobc.set_param_dependence(dp)
    ref = (onobc.assemble(vfvm.dofs(U), embed=lam + eps) - onobc.assemble(vfvm.dofs(U), embed=lam - eps)) / (2 * eps)
    err = np.abs(d - ref).max() / np.abs(ref).max()
    _print("relative error vs central differences", err)
    assert err <= 1e-7
