"""Pins the oracle's lambda-dependent registry coefficients with the reference's own known
answer: examples/Example430_ParameterDerivativesStationary.jl:258 testval
489.3432830184927, the sum over p in 0.1:0.05:2 of dI/dp, I(p) the test-function flux
(integrate(sys, tfc, sol), T = 0 on region 1 and 1 on region 3) of the solution of
flux (1 + p)(u_k^2 - u_l^2), reaction p u^5, u = 0 on region 1 and u = p on region 3.
Here dI/dp comes from central differences of the oracle's I(p) (direct solves)."""
import numpy as np
import pytest

import oracle as O
from vfvm.grids import simplexgrid

TESTVAL = 489.3432830184927
P = 0.1 + 0.05 * np.arange(39)  # 0.1:0.05:2


def example430_oracle(n=10):
    X = np.arange(n + 1) / n
    g = simplexgrid(X, X)
    ph = O.physics("power", "pow", D0=1.0, m=2, k=0.0, rm=5)
    dp = np.zeros(16)
    dp[0] = 1.0  # D(p) = 1 + p
    dp[5] = 1.0  # k(p) = p
    s = O.OracleSystem(g, 1, ph, {(1, 0): ("dirichlet", 0.0), (3, 0): ("dirichlet", 0.0, 0.0, 1.0)}, O.EDGEWISE)
    s.set_param_dependence(dp)
    nobc = O.OracleSystem(g, 1, ph, {}, O.EDGEWISE)  # integrate sees no boundary conditions
    nobc.set_param_dependence(dp)
    lap = O.OracleSystem(g, 1, O.physics("linear", D0=1.0), {(1, 0): ("dirichlet", 0.0), (3, 0): ("dirichlet", 1.0)},
                         O.EDGEWISE)
    T, _ = lap.solve_step(np.zeros(g.num_nodes))
    return g, s, nobc, T


def test_example430_known_answer_by_central_differences():
    g, s, nobc, T = example430_oracle()
    c = O.control(reltol=1e-14, abstol=1e-14)
    u0 = np.full(g.num_nodes, 0.5)

    def I(p):
        U, _ = s.solve_step(u0, c, embed=p)
        return float(T @ nobc.assemble(U, embed=p))

    h = 1e-5
    total = sum((I(p + h) - I(p - h)) / (2 * h) for p in P)
    assert total == pytest.approx(TESTVAL, rel=1e-7)


def test_param_dependence_is_affine_in_lambda():
    """p_j(lambda) = p_j + dp_j lambda enters the power flux and the power reaction
    linearly, so F(U; lambda) is affine in lambda: F(1) - F(0) = F(0) - F(-1)."""
    g, s, nobc, T = example430_oracle()
    U = np.random.default_rng(4).uniform(0.2, 1.0, g.num_nodes)
    F = [nobc.assemble(U, embed=lam) for lam in (-1.0, 0.0, 1.0)]
    assert np.abs((F[2] - F[1]) - (F[1] - F[0])).max() <= 1e-12 * np.abs(F[1]).max()
    assert np.abs(F[2] - F[1]).max() > 1e-3 * np.abs(F[1]).max()
