"""The benchmark configurations' problems as device systems (the reference's examples
they follow are named per builder).  The test suite builds the oracle's twin of the same
systems; nothing here touches the oracle."""
from __future__ import annotations

import math

import numpy as np

from .grids import linspace, simplexgrid
from .physics import (Dirichlet, DriftDiffusionFlux, DriftDiffusionReaction, IdentityStorage, LinearFlux,
                      NonlinearDiffusionFlux, Physics, SinhReaction)
from .system import System


def nonlinear_poisson_2d(n=101, eps=1e-2, f=1.0, values=(0.0, 1.0, 1.0, 0.0)):
    """Nonlinear Poisson on an n x n grid of the unit square: flux eps (u_k - u_l), reaction
    e^u - e^-u - f (examples/Example105_NonlinearPoisson1D.jl:41-60 in 2D); Dirichlet values
    on south, east, north, west (boundary regions 1..4)."""
    X = linspace(0.0, 1.0, n)
    g = simplexgrid(X, X)
    bcs = [Dirichlet(species=1, region=r + 1, value=v) for r, v in enumerate(values)]
    src = np.full((1, g.num_nodes), f, order="F")
    ph = Physics(flux=LinearFlux(eps), reaction=SinhReaction(), source=src, bcondition=bcs)
    return System(g, ph, species=[1])


def nonlinear_diffusion_3d(n=17):
    """3D nonlinear diffusion D(u) = 1 + u^2 (diffusion_flux, src/vfvm_physics.jl:478-480),
    storage u, u = 0 on all six faces, on an n^3 grid of the unit cube."""
    X = linspace(0.0, 1.0, n)
    g = simplexgrid(X, X, X)
    bcs = [Dirichlet(species=1, region=r, value=0.0) for r in range(1, 7)]
    ph = Physics(flux=NonlinearDiffusionFlux(1.0, 1.0), storage=IdentityStorage(), bcondition=bcs)
    return System(g, ph, species=[1])


def initial_sine_3d(grid):
    """u0 = 0.5 sin(pi x) sin(pi y) sin(pi z)"""
    x, y, z = grid.coord
    return 0.5 * np.sin(math.pi * x) * np.sin(math.pi * y) * np.sin(math.pi * z)


def drift_diffusion_2d(n=33, lambda2=1e-4):
    """Bipolar drift-diffusion (psi, n, p) on an n x n grid of the unit square
    (examples/Example161_BipolarDriftDiffusionCurrent.jl:109-171 with the classic
    Scharfetter-Gummel flux): doping +1 for x < 0.5 else -1, ohmic contacts on the west
    (x = 0) and east (x = 1) sides, the east contact biased by the embedding parameter."""
    X = linspace(0.0, 1.0, n)
    g = simplexgrid(X, X)
    x = g.coord[0]
    C = np.where(x < 0.5, 1.0, -1.0)
    src = np.zeros((3, g.num_nodes), order="F")
    src[0] = C
    bcs = []
    for reg, c, bias in ((4, 1.0, 0.0), (2, -1.0, 1.0)):
        nn = 0.5 * (c + math.sqrt(c * c + 4.0))
        pp = 1.0 / nn
        psi = math.asinh(c / 2.0)
        bcs += [Dirichlet(species=1, region=reg, value=psi, value_lambda=bias),
                Dirichlet(species=2, region=reg, value=nn),
                Dirichlet(species=3, region=reg, value=pp)]
    ph = Physics(flux=DriftDiffusionFlux(lambda2, -1.0, 1.0), reaction=DriftDiffusionReaction(1.0), source=src,
                 bcondition=bcs)
    return System(g, ph, species=[1, 2, 3])


def dd_initial(grid):
    """The zero-bias equilibrium of the doping: psi = asinh(C / 2), n = e^psi, p = 1 / n."""
    x = grid.coord[0]
    C = np.where(x < 0.5, 1.0, -1.0)
    U = np.zeros((3, grid.num_nodes), order="F")
    U[0] = np.arcsinh(C / 2.0)
    U[1] = 0.5 * (C + np.sqrt(C * C + 4))
    U[2] = 1.0 / U[1]
    return U


def batch_member(seed, sizes=(33, 65, 129, 257)):
    """One instance of the batched nonlinear-Poisson family: N from `sizes`, eps
    log-uniform in [1e-3, 1], Dirichlet values uniform in [-1, 1] per side, constant
    source uniform in [-5, 5]; all drawn from numpy.random.default_rng(seed)."""
    rng = np.random.default_rng(seed)
    n = int(rng.choice(sizes))
    eps = float(10.0 ** rng.uniform(-3.0, 0.0))
    vals = rng.uniform(-1.0, 1.0, 4)
    f = float(rng.uniform(-5.0, 5.0))
    return n, eps, vals, f
