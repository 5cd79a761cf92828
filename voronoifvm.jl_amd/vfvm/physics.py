"""The fixed registry of device callbacks standing in for VoronoiFVM's Physics
closures (src/vfvm_physics.jl:67-182).

GPU code cannot call Julia (or Python) closures, so a Physics is assembled from
registry entries.  Each entry names the reference example whose callback it
restates; the parameter slots match vfvm_physics_t.p in include/vfvm_hip.h.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Callable, Optional, Sequence, Union

import numpy as np

FLUX_NONE, FLUX_LINEAR, FLUX_POWER, FLUX_NLDIFF, FLUX_DD = range(5)
REACT_NONE, REACT_SINH, REACT_POW, REACT_EXCHANGE, REACT_DD = range(5)
STOR_NONE, STOR_IDENTITY = range(2)
BC_NONE, BC_DIRICHLET, BC_DIRICHLET_LEGACY, BC_ROBIN = range(4)


@dataclass
class LinearFlux:
    """f[i] = D_i (u[i,1] - u[i,2]); Example105_NonlinearPoisson1D.jl:41-44."""
    D: Union[float, Sequence[float]] = 1.0

    def fill(self, p, nspec):
        D = np.broadcast_to(np.asarray(self.D, dtype=float), (nspec,))
        p[0:nspec] = D
        return FLUX_LINEAR


@dataclass
class PowerFlux:
    """f[1] = D (u[1,1]^m - u[1,2]^m); Example106_NonlinearDiffusion1D.jl:57-60."""
    m: int = 2
    D: float = 1.0

    def fill(self, p, nspec):
        p[0] = self.D
        p[4] = float(int(self.m))
        return FLUX_POWER


@dataclass
class NonlinearDiffusionFlux:
    """f[1] = D(ubar)(u[1,1] - u[1,2]), D(u) = d0 + d2 u^2, ubar the edge mean
    (diffusion_flux, src/vfvm_physics.jl:478-480)."""
    d0: float = 1.0
    d2: float = 1.0

    def fill(self, p, nspec):
        p[0] = self.d0
        p[3] = self.d2
        return FLUX_NLDIFF


@dataclass
class DriftDiffusionFlux:
    """psi: lambda^2 (psi_k - psi_l); carriers c with charge z: Scharfetter-Gummel
    B(-z dpsi) c_k - B(z dpsi) c_l (Example160_UnipolarDriftDiffusion1D.jl:43-50)."""
    lambda2: float = 1e-4
    zn: float = -1.0
    zp: float = 1.0

    def fill(self, p, nspec):
        p[7], p[8], p[9] = self.lambda2, self.zn, self.zp
        return FLUX_DD


@dataclass
class SinhReaction:
    """f[1] = k (e^u - e^-u); Example105_NonlinearPoisson1D.jl:56-59."""
    k: float = 1.0

    def fill(self, p, nspec):
        p[5] = self.k
        return REACT_SINH


@dataclass
class PowerReaction:
    """f[1] = k u^m; Example207_NonlinearPoisson2D.jl:34-36."""
    m: int = 2
    k: float = 1.0

    def fill(self, p, nspec):
        p[5] = self.k
        p[6] = float(int(self.m))
        return REACT_POW


@dataclass
class ExchangeReaction:
    """f[1] = k(u1 - u2), f[2] = k(u2 - u1); Example210:25-30."""
    k: float = 1.0

    def fill(self, p, nspec):
        p[5] = self.k
        return REACT_EXCHANGE


@dataclass
class DriftDiffusionReaction:
    """psi: n - p (doping enters as source); n, p: (np - ni^2)/(n + p + 2 ni)."""
    ni: float = 1.0

    def fill(self, p, nspec):
        p[10], p[11] = self.ni * self.ni, self.ni
        return REACT_DD


@dataclass
class IdentityStorage:
    """f[i] = u[i]; Example106_NonlinearDiffusion1D.jl:63-66."""

    def fill(self, p, nspec):
        return STOR_IDENTITY


@dataclass
class Dirichlet:
    """boundary_dirichlet!(y, u, bnode; species, region, value) in a bcondition
    callback: penalty 1e30 times the boundary node factor (src/vfvm_physics.jl:487-506).
    value(lambda) = value + value_lambda * lambda (embedding)."""
    species: int = 1
    region: int = 1
    value: float = 0.0
    value_lambda: float = 0.0


@dataclass
class Robin:
    """boundary_robin!(y, u, bnode; species, region, factor, value): factor u - value."""
    species: int = 1
    region: int = 1
    factor: float = 0.0
    value: float = 0.0


def Neumann(species=1, region=1, value=0.0):
    """boundary_neumann!(y, u, bnode; species, region, value): -value (src/vfvm_physics.jl:533)."""
    return Robin(species=species, region=region, factor=0.0, value=value)


# ---- boundary physics (assemble_bnodes / assemble_bedges, src/vfvm_assembly.jl:318-510) ----

@dataclass
class BoundaryReaction:
    """breaction on boundary region `region`: f_i = sum_j matrix[i][j] u_j (species in
    1-based order, rows = equations), e.g. the exchange of
    Example220_NonlinearPoisson2D_BoundarySpecies.jl:25-33 or
    Example311_HeatEquation_BoundaryDiffusion.jl:66-75."""
    region: int = 1
    matrix: Sequence[Sequence[float]] = ()


@dataclass
class BoundaryStorage:
    """bstorage on `region`: f_i = u_i for `species` (Example220:33-37)."""
    region: int = 1
    species: int = 1


@dataclass
class BoundaryFlux:
    """bflux on the boundary edges of `region`: f_i = D (u_i,K - u_i,L) for `species`
    (Example230_BoundaryFlux.jl:125-130, Example311:58-64)."""
    region: int = 1
    species: int = 1
    D: float = 1.0


@dataclass
class BoundarySource:
    """bsource on `region` for `species`: per-node values (array over all nodes) or a
    callable of the node coordinates (Example230:145-156)."""
    region: int = 1
    species: int = 1
    values: Union[np.ndarray, Callable, float] = 0.0


@dataclass
class BoundaryReactionField:
    """breaction term c_K u_i on `region` for `species` with per-node c (array or callable
    of the coordinates), e.g. Example230:132-143."""
    region: int = 1
    species: int = 1
    values: Union[np.ndarray, Callable, float] = 0.0


@dataclass
class ReactionField:
    """bulk reaction term c_K u_i for `species` with per-node c (array or callable of the
    coordinates), a reaction callback reading node data (Example230:63-66)."""
    species: int = 1
    values: Union[np.ndarray, Callable, float] = 0.0


@dataclass
class GenericMatrix:
    """Physics(generic = f) in registry form: a constant matrix G over the dofs (the
    order of dofs(U)); every assembly adds G U to F and G to the Jacobian
    (_eval_and_assemble_generic_operator, src/vfvm_assembly.jl:648-677), as
    examples/Example405_GenericOperator.jl's linear upwind operator does.  G: a scipy
    sparse matrix or a dense array whose entries lie in the system's sparsity pattern
    (the pattern holds the grid's edge couplings when the physics has a flux; with a
    generic operator and no flux the system takes a flux with zero coefficients, so G
    may couple neighbouring nodes, as Example405's operator does)."""
    matrix: object

    def coo(self):
        """1-based dof rows, columns and values of the stored entries"""
        import scipy.sparse as sp
        A = sp.coo_matrix(self.matrix)
        return (np.ascontiguousarray(A.row.astype(np.int64) + 1), np.ascontiguousarray(A.col.astype(np.int64) + 1),
                np.ascontiguousarray(A.data.astype(np.float64)))


@dataclass
class Physics:
    """VoronoiFVM.Physics(; flux, reaction, storage, source, bcondition, breaction,
    bstorage, bflux, bsource).

    embed_source: the source callback returns embedparam(node) * f, so that
    solve(...; embed) ramps the source with the embedding parameter.
    breaction: BoundaryReaction / BoundaryReactionField entries; bstorage:
    BoundaryStorage; bflux: BoundaryFlux; bsource: BoundarySource; reaction_fields:
    ReactionField entries added to the bulk reaction."""
    flux: Optional[object] = None
    reaction: Optional[object] = None
    storage: Optional[object] = None
    source: Optional[Union[np.ndarray, Callable]] = None
    bcondition: list = field(default_factory=list)
    embed_source: bool = False
    breaction: list = field(default_factory=list)
    bstorage: list = field(default_factory=list)
    bflux: list = field(default_factory=list)
    bsource: list = field(default_factory=list)
    reaction_fields: list = field(default_factory=list)
    generic: Optional[GenericMatrix] = None  # the generic operator (a constant matrix)
    # registry objects whose coefficients are the derivatives of flux's and reaction's
    # coefficients with respect to the system parameter lambda (the coefficient is
    # c + dc lambda; Example430_ParameterDerivativesStationary.jl:142-160 passes its p
    # this way).  Exponents do not depend on lambda.
    dflux: Optional[object] = None
    dreaction: Optional[object] = None

    def dparam_block(self, nspec):
        """d(parameter slot)/dlambda, 16 values (zero where nothing depends on lambda)"""
        dp = np.zeros(16)
        if self.dflux is not None:
            self.dflux.fill(dp, nspec)
        if self.dreaction is not None:
            self.dreaction.fill(dp, nspec)
        dp[4] = dp[6] = 0.0  # integer exponents
        dp[12] = 0.0  # the embedded-source flag
        return dp

    def has_boundary_physics(self):
        return bool(self.breaction or self.bstorage or self.bflux or self.bsource)

    def param_block(self, nspec):
        p = np.zeros(16)
        flux = self.flux.fill(p, nspec) if self.flux is not None else FLUX_NONE
        if self.flux is None and self.generic is not None:
            # the edge couplings a generic operator acts through stay in the pattern: a
            # flux with zero coefficients (it adds nothing)
            flux = LinearFlux(0.0).fill(p, nspec)
        react = self.reaction.fill(p, nspec) if self.reaction is not None else REACT_NONE
        stor = self.storage.fill(p, nspec) if self.storage is not None else STOR_NONE
        if self.embed_source:
            p[12] = 1.0
        return flux, react, stor, p
