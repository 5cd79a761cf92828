"""Builds the same problem twice: as a vfvm.System (HIP path) and as an
OracleSystem (CPU restatement), from one description.  The oracle is imported
only when an oracle system is asked for (oracle=True), so that the GPU legs of
the benchmark can build their problems here without loading it."""
import math

import numpy as np

import vfvm
from vfvm import problems
from vfvm.grids import linspace, simplexgrid

ORACLE_BC = {vfvm.Dirichlet: "dirichlet", vfvm.Robin: "robin"}


def oracle_physics(ph, nspec):
    """The oracle's parameter block from the same registry objects."""
    import oracle as O  # imported only where an oracle system is built
    flux, react, stor, p = ph.param_block(nspec)
    oph = O.Physics()
    oph.flux, oph.reaction, oph.storage = flux, react, stor
    for k in range(16):
        oph.p[k] = p[k]
    return oph


def pair(grid, nspec, phys_kw, bcs=(), legacy=(), source=None, oracle=True, embed_source=False):
    """phys_kw: dict(flux=..., reaction=..., storage=...) of vfvm registry objects.
    bcs: vfvm.Dirichlet / vfvm.Robin entries; legacy: (species, region, value) Dirichlet.
    With oracle=False only the device system is built (second item None)."""
    phys_kw = dict(phys_kw)
    embed_source = embed_source or phys_kw.pop("embed_source", False)
    ph = vfvm.Physics(bcondition=list(bcs), source=source, embed_source=embed_source, **phys_kw)
    sys = vfvm.System(grid, ph, species=list(range(1, nspec + 1)))
    for (sp, reg, val) in legacy:
        sys.boundary_dirichlet(sp, reg, val)
    if not oracle:
        return sys, None
    oph = oracle_physics(ph, nspec)
    obcs = {}
    for b in bcs:
        if isinstance(b, vfvm.Dirichlet):
            obcs[(b.region, b.species - 1)] = ("dirichlet", b.value, 0.0, b.value_lambda)
        else:
            obcs[(b.region, b.species - 1)] = ("robin", b.value, b.factor)
    for (sp, reg, val) in legacy:
        obcs[(reg, sp - 1)] = ("dirichlet_legacy", val)
    osrc = None if source is None else np.asarray(source, dtype=float).reshape(-1, order="F")
    import oracle as O
    osys = O.OracleSystem(grid, nspec, oph, obcs, O.EDGEWISE, osrc)
    return sys, osys


def oracle_twin(sys):
    """The oracle's system for a vfvm.System built from registry entries (physics block,
    bcondition entries, legacy boundary conditions, source array)."""
    import oracle as O
    ph = sys.physics
    ns = sys.num_species
    obcs = {}
    for b in ph.bcondition:
        if isinstance(b, vfvm.Dirichlet):
            obcs[(b.region, b.species - 1)] = ("dirichlet", b.value, 0.0, b.value_lambda)
        else:
            obcs[(b.region, b.species - 1)] = ("robin", b.value, b.factor)
    for (reg, sp), (kind, val, fac) in sys._legacy.items():
        obcs[(reg, sp - 1)] = ("dirichlet_legacy", val) if kind == vfvm.physics.BC_DIRICHLET_LEGACY else \
            ("robin", val, fac)
    src = sys._source_array()
    osrc = None if src is None else np.asarray(src, dtype=float).reshape(-1, order="F")
    return O.OracleSystem(sys.grid, ns, oracle_physics(ph, ns), obcs, O.EDGEWISE, osrc)


def nonlinear_poisson_2d(n=101, eps=1e-2, f=1.0, values=(0.0, 1.0, 1.0, 0.0), oracle=True):
    """Nonlinear Poisson: flux eps (u_k - u_l), reaction e^u - e^-u - f; Dirichlet
    values on south, east, north, west (regions 1..4)."""
    sys = problems.nonlinear_poisson_2d(n, eps, f, values)
    return sys, (oracle_twin(sys) if oracle else None)


def sg_grid(name):
    """One of the reference's unstructured grid assets, kept as fixtures in tests/golden."""
    import os
    from vfvm.grids import read_sg
    return read_sg(os.path.join(os.path.dirname(__file__), "golden", name))


def nonlinear_poisson_sg(name, eps=0.1, f=1.0):
    """Nonlinear Poisson (as nonlinear_poisson_2d) on an unstructured .sg grid: u = 0 on
    boundary face region 1, u = 1 on region 2, homogeneous Neumann elsewhere."""
    return nonlinear_poisson_on(sg_grid(name), eps, f)


def nonlinear_poisson_on(g, eps=0.1, f=1.0):
    """Nonlinear Poisson on any grid: u = 0 on boundary face region 1, u = 1 on region 2."""
    bcs = [vfvm.Dirichlet(species=1, region=1, value=0.0), vfvm.Dirichlet(species=1, region=2, value=1.0)]
    src = np.full((1, g.num_nodes), f, order="F")
    return pair(g, 1, dict(flux=vfvm.LinearFlux(eps), reaction=vfvm.SinhReaction()), bcs, source=src)


def nonlinear_diffusion_3d(n=17, oracle=True):
    """3D nonlinear diffusion D(u) = 1 + u^2, storage u, u = 0 on all faces."""
    sys = problems.nonlinear_diffusion_3d(n)
    return sys, (oracle_twin(sys) if oracle else None)


initial_sine_3d = problems.initial_sine_3d


def drift_diffusion_2d(n=33, lambda2=1e-4, oracle=True):
    """Bipolar drift-diffusion (psi, n, p), doping +1 for x < 0.5 else -1, ohmic
    contacts on west (x = 0) and east (x = 1); east contact biased by lambda."""
    sys = problems.drift_diffusion_2d(n, lambda2)
    return sys, (oracle_twin(sys) if oracle else None)


dd_initial = problems.dd_initial


def penalty_rows(osys):
    """dof mask of rows carrying a Dirichlet penalty (diagonal >= 1e20)."""
    A = osys.matrix_scipy()
    return np.abs(A.diagonal()) > 1e20


def compare_assembly(sys, osys, U, Uold=None, tstep=math.inf, embed=0.0, rtol=1e-12):
    """Assemble on both paths; pattern must be identical, values within rtol (normwise,
    penalty and non-penalty rows separately).  Returns the max relative errors."""
    Uold = U if Uold is None else Uold
    Fo = osys.assemble(vfvm.dofs(U), vfvm.dofs(Uold), tstep=tstep, embed=embed)
    Ao = osys.matrix_scipy().tocsc()
    Ao.sort_indices()
    import ctypes as C
    from vfvm import _capi
    L = _capi.lib()
    h = sys.handle()
    ndof, nnz, _ = sys.sizes()
    Ud = np.ascontiguousarray(vfvm.dofs(U), dtype=float)
    Uod = np.ascontiguousarray(vfvm.dofs(Uold), dtype=float)
    F = np.zeros(ndof)
    nz = np.zeros(nnz)
    _capi.check(L.vfvm_assemble(h, _capi.dptr(Ud), _capi.dptr(Uod), 0.0, tstep, embed, _capi.dptr(F), _capi.dptr(nz)))
    cp, rv = sys.pattern()
    assert np.array_equal(cp - 1, Ao.indptr), "colptr differs from the oracle pattern"
    assert np.array_equal(rv - 1, Ao.indices), "rowval differs from the oracle pattern"
    pen = np.abs(Ao.diagonal()) > 1e20
    rows = np.repeat(np.arange(ndof), 0)  # placeholder
    # entrywise row index of each CSC value
    rowidx = Ao.indices
    errs = {}
    for name, mask in (("free", ~pen), ("penalty", pen)):
        if mask.sum() == 0:
            continue
        fo, fg = Fo[mask], F[mask]
        sc = max(np.abs(fo).max(), 1e-300)
        errs["F_" + name] = np.abs(fo - fg).max() / sc
        em = mask[rowidx]
        jo, jg = Ao.data[em], nz[em]
        sc = max(np.abs(jo).max(), 1e-300)
        errs["J_" + name] = np.abs(jo - jg).max() / sc
    for k, v in errs.items():
        assert v <= rtol, (k, v)
    return errs


batch_member = problems.batch_member
