"""Host-side mirror of VoronoiFVM's System / solve surface, driving the HIP library.

Mirrors (paths relative to the reference root):
  System(grid; physics..., species)          src/vfvm_system.jl:215-261
  enable_species!                              src/vfvm_system.jl:433-470
  boundary_dirichlet!/neumann!/robin!(sys,..)  src/vfvm_system.jl:854-933 (legacy BC API)
  unknowns(sys; inival)                        src/vfvm_system.jl:1086-1110
  solve(sys; inival, control, tstep, times, embed)  src/vfvm_solver.jl:551-668
  evaluate_residual_and_jacobian               src/vfvm_solver.jl:224-260
Arrays of unknowns are nspec x nnodes in Fortran order, so that
``U.ravel(order="F")`` is the reference's ``dofs(U)`` order.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _capi
from .control import SolverControl
from .grids import Grid
from .physics import (BC_DIRICHLET, BC_DIRICHLET_LEGACY, BC_ROBIN, BoundaryFlux, BoundaryReaction,
                      BoundaryReactionField, BoundarySource, BoundaryStorage, Dirichlet, Physics, ReactionField,
                      Robin)


@dataclass
class NewtonSolverHistory:
    """src/vfvm_history.jl:10-37 subset."""
    nlu: int = 0
    nlin: int = 0
    niter: int = 0
    time: float = 0.0
    updatenorm: tuple = ()
    residual_norm: float = float("nan")  # ||F||_inf over non-penalty rows at the solution
    fp64_fallbacks: int = 0  # linear solves done in FP64 although FP32 refinement was requested


class System:
    """VoronoiFVM.System(grid, physics; species, is_linear, assembly).

    assembly: "edgewise" (factors summed per edge; no Jacobian entry for an edge whose
    factor is zero) or "cellwise" (the reference's default: an entry for every cell edge;
    values equal to rounding), src/vfvm_system.jl:215-227."""

    def __init__(self, grid: Grid, physics: Optional[Physics] = None, species=None, is_linear=False,
                 context=None, nspec=None, assembly="edgewise", **physics_kw):
        self.grid = grid
        self.physics = physics if physics is not None else Physics(**physics_kw)
        self.is_linear = is_linear
        self.species = sorted(set(species or []))
        if nspec is not None:
            self.species = list(range(1, nspec + 1))
        if assembly not in _capi.ASSEMBLY:
            raise ValueError("assembly must be 'edgewise' or 'cellwise'")
        self.assembly = assembly
        self._bspecies = {}  # boundary species -> boundary regions
        self._legacy = {}  # (region, species) -> (kind, value, factor)
        self._ctx = context
        self._h = None
        self._bc_dirty = True
        self.history = None

    # --- species and legacy boundary conditions -------------------------------------------------
    def enable_species(self, ispec, regions=None):
        """enable_species!(sys, ispec, regions) (src/vfvm_system.jl:433-455).  The registry
        path assembles a bulk species in every cell region: `regions` must name all of the
        grid's cell regions (or be None)."""
        if ispec in self._bspecies:
            raise ValueError(f"species {ispec} is already a boundary species")
        if regions is not None:
            allreg = set(int(r) for r in np.unique(self.grid.cellregions))
            if set(int(r) for r in regions) != allreg:
                raise ValueError("a bulk species enabled in only some cell regions is not supported: "
                                 f"regions {sorted(regions)} vs the grid's {sorted(allreg)}")
        if ispec not in self.species:
            self.species = sorted(self.species + [ispec])
            self._drop()

    def enable_boundary_species(self, ispec, bregions):
        """enable_boundary_species!(sys, ispec, bregions) (src/vfvm_system.jl:492-515)."""
        if ispec in self.species:
            raise ValueError(f"species {ispec} is already a bulk species")
        self._bspecies[ispec] = sorted(set(int(r) for r in bregions))
        self._drop()

    @property
    def num_species(self):
        return max(list(self.species) + list(self._bspecies)) if (self.species or self._bspecies) else 0

    def _species_table(self):
        ns = self.num_species
        tab = (_capi.SpeciesT * ns)()
        for i in range(1, ns + 1):
            e = tab[i - 1]
            if i in self._bspecies:
                e.kind = _capi.SPECIES_BOUNDARY
                for r in self._bspecies[i]:
                    if not 1 <= r <= 32:
                        raise ValueError("boundary region numbers must lie in 1..32")
                    e.bregions |= 1 << (r - 1)
            elif i in self.species:
                e.kind = _capi.SPECIES_BULK
            else:
                e.kind = _capi.SPECIES_NONE
        return tab

    def _bphys_table(self):
        ph = self.physics
        if not ph.has_boundary_physics():
            return None
        ns, nb = self.num_species, self._nbreg()
        tab = (_capi.BPhysT * nb)()

        def entry(r):
            if not 1 <= r <= nb:
                raise ValueError(f"boundary region {r} outside 1..{nb}")
            return tab[r - 1]

        for b in ph.breaction:
            if isinstance(b, BoundaryReaction):
                e = entry(b.region)
                M = np.asarray(b.matrix, dtype=float)
                if M.shape != (ns, ns):
                    raise ValueError(f"BoundaryReaction matrix must be {ns} x {ns}")
                for i in range(ns):
                    for j in range(ns):
                        e.react[i][j] = M[i, j]
            elif isinstance(b, BoundaryReactionField):
                entry(b.region).fields = 1
            else:
                raise TypeError(f"not a boundary reaction of the registry: {b!r}")
        for b in ph.bstorage:
            entry(b.region).storage |= 1 << (b.species - 1)
        for b in ph.bflux:
            entry(b.region).bflux[b.species - 1] = float(b.D)
        for b in ph.bsource:
            entry(b.region).fields = 1
        return tab

    def _node_values(self, values):
        n = self.grid.num_nodes
        if callable(values):
            return np.array([values(*self.grid.coord[:, k]) for k in range(n)], dtype=float)
        return np.broadcast_to(np.asarray(values, dtype=float), (n,)).copy()

    def _field_arrays(self):
        """(rcoef, bsrc, bcoef), nspec x nnodes Fortran arrays or None.  Boundary fields are
        set on the nodes of their region's boundary faces."""
        ph = self.physics
        ns, n = self.num_species, self.grid.num_nodes
        g = self.grid

        def region_nodes(r):
            return np.unique(g.bfacenodes[:, g.bfaceregions == r])

        rcoef = bsrc = bcoef = None
        for f in ph.reaction_fields:
            rcoef = np.zeros((ns, n), order="F") if rcoef is None else rcoef
            rcoef[f.species - 1] += self._node_values(f.values)
        for lst, kind in ((ph.bsource, BoundarySource), (ph.breaction, BoundaryReactionField)):
            for f in lst:
                if not isinstance(f, kind):
                    continue
                nodes = region_nodes(f.region)
                vals = self._node_values(f.values)
                if kind is BoundarySource:
                    bsrc = np.zeros((ns, n), order="F") if bsrc is None else bsrc
                    bsrc[f.species - 1, nodes] = vals[nodes]
                else:
                    bcoef = np.zeros((ns, n), order="F") if bcoef is None else bcoef
                    bcoef[f.species - 1, nodes] = vals[nodes]
        return rcoef, bsrc, bcoef

    def node_species(self):
        """node_dof activity: bit i - 1 of entry K is set if species i is active at node K."""
        m = np.zeros(self.grid.num_nodes, np.uint8)
        _capi.check(_capi.lib().vfvm_node_species(self.handle(), m.ctypes.data_as(C.POINTER(C.c_uint8))),
                    self.ctx.handle)
        return m

    def boundary_dirichlet(self, ispec, ibc, v):
        self.enable_species(ispec)
        self._legacy[(ibc, ispec)] = (BC_DIRICHLET_LEGACY, float(v), 0.0)
        self._bc_dirty = True

    def boundary_robin(self, ispec, ibc, alpha, v):
        self.enable_species(ispec)
        self._legacy[(ibc, ispec)] = (BC_ROBIN, float(v), float(alpha))
        self._bc_dirty = True

    def boundary_neumann(self, ispec, ibc, v):
        self.boundary_robin(ispec, ibc, 0.0, v)

    # --- device handle --------------------------------------------------------------------------
    def _nbreg(self):
        return max(self.grid.num_bfaceregions, 1)

    def _bc_table(self):
        ns, nb = self.num_species, self._nbreg()
        tab = (_capi.BCT * (nb * ns))()
        for bc in self.physics.bcondition:
            if bc.region > nb or bc.species > ns:
                continue
            e = tab[(bc.region - 1) * ns + bc.species - 1]
            if isinstance(bc, Dirichlet):
                e.kind, e.value, e.value_lambda = BC_DIRICHLET, bc.value, bc.value_lambda
            elif isinstance(bc, Robin):
                e.kind, e.value, e.factor = BC_ROBIN, bc.value, bc.factor
        for (reg, sp), (kind, val, fac) in self._legacy.items():
            if reg <= nb and sp <= ns:
                e = tab[(reg - 1) * ns + sp - 1]
                e.kind, e.value, e.factor = kind, val, fac
        return tab

    def _source_array(self):
        src = self.physics.source
        if src is None:
            return None
        ns, n = self.num_species, self.grid.num_nodes
        if callable(src):
            vals = np.array([src(*self.grid.coord[:, k]) for k in range(n)], dtype=float)
            src = vals
        src = np.asarray(src, dtype=float)
        if src.ndim == 1 and src.size == n:
            full = np.zeros((ns, n), order="F")
            full[0] = src
            src = full
        return np.asfortranarray(src.reshape(ns, n, order="F") if src.ndim == 1 else src)

    def _drop(self):
        if self._h is not None and _capi._LIB is not None:
            _capi._LIB.vfvm_system_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self._drop()
        except Exception:
            pass

    @property
    def ctx(self):
        if self._ctx is None:
            self._ctx = _capi.default_context()
        return self._ctx

    def handle(self):
        """_complete!: build the device system on first use (src/vfvm_system.jl:530-596)."""
        if self._h is None:
            L = _capi.lib()
            g = self.grid
            ns = self.num_species
            if ns < 1:
                raise ValueError("No species enabled. Call enable_species at least once.")
            flux, react, stor, p = self.physics.param_block(ns)
            ph = _capi.PhysicsT()
            ph.flux, ph.reaction, ph.storage = flux, react, stor
            for k in range(16):
                ph.p[k] = p[k]
            coord = np.asfortranarray(g.coord, dtype=np.float64)
            cn = np.asfortranarray(g.cellnodes.astype(np.int64) + 1)
            cr = np.ascontiguousarray(g.cellregions.astype(np.int64))
            bn = np.asfortranarray(g.bfacenodes.astype(np.int64) + 1)
            br = np.ascontiguousarray(g.bfaceregions.astype(np.int64))
            tab = self._bc_table()
            spec = self._species_table()
            bph = self._bphys_table()
            h = C.c_void_p()
            rc = L.vfvm_system_create_ex(self.ctx.handle, g.dim, g.num_nodes, _capi.dptr(coord), g.num_cells,
                                         _capi.lptr(cn), _capi.lptr(cr), g.num_bfaces, _capi.lptr(bn),
                                         _capi.lptr(br), ns, C.byref(ph), self._nbreg(), tab, int(g.coordsys),
                                         spec, bph, _capi.ASSEMBLY[self.assembly], C.byref(h))
            _capi.check(rc, self.ctx.handle)
            self._h = h
            src = self._source_array()
            if src is not None:
                _capi.check(L.vfvm_system_set_source(h, _capi.dptr(src)), self.ctx.handle)
            dp = self.physics.dparam_block(ns)
            if np.any(dp != 0.0):
                _capi.check(L.vfvm_system_set_param_dependence(h, _capi.dptr(np.ascontiguousarray(dp))),
                            self.ctx.handle)
            fields = self._field_arrays()
            if any(f is not None for f in fields):
                _capi.check(L.vfvm_system_set_fields(h, *[_capi.dptr(f) if f is not None else None for f in fields]),
                            self.ctx.handle)
            if self.physics.generic is not None:
                r, c, v = self.physics.generic.coo()
                _capi.check(L.vfvm_system_set_generic_matrix(h, v.size, _capi.lptr(r), _capi.lptr(c), _capi.dptr(v)),
                            self.ctx.handle)
            self._bc_dirty = False
        elif self._bc_dirty:
            _capi.check(_capi.lib().vfvm_system_set_bc(self._h, self._bc_table()), self.ctx.handle)
            self._bc_dirty = False
        return self._h

    def set_source(self, src):
        self.physics.source = src
        if self._h is not None:
            arr = self._source_array()
            _capi.check(_capi.lib().vfvm_system_set_source(self._h, _capi.dptr(arr) if arr is not None else None),
                        self.ctx.handle)

    def update_bcondition(self, bcondition):
        self.physics.bcondition = list(bcondition)
        self._bc_dirty = True

    def sizes(self):
        nd, nz, ne = C.c_int64(), C.c_int64(), C.c_int64()
        _capi.check(_capi.lib().vfvm_system_sizes(self.handle(), C.byref(nd), C.byref(nz), C.byref(ne)))
        return nd.value, nz.value, ne.value

    def autotune(self):
        a, s = C.c_int32(), C.c_int32()
        _capi.check(_capi.lib().vfvm_autotune(self.handle(), C.byref(a), C.byref(s)), self.ctx.handle)
        return a.value, s.value

    def set_compact_graph(self, enable: bool):
        """Assembly reads the compact node graph (16-bit neighbour offsets, factor
        table; default where it fits) or int32 neighbours and FP64 factors."""
        _capi.check(_capi.lib().vfvm_set_compact_graph(self.handle(), int(bool(enable))), self.ctx.handle)

    def set_sell_index(self, bits: int):
        """SpMV column indices: 16-bit offsets from the row (default where they fit) or
        32-bit absolute columns."""
        _capi.check(_capi.lib().vfvm_set_sell_index(self.handle(), int(bits)), self.ctx.handle)

    def set_groups(self, asm_group: int, spmv_group: int = 8):
        """Choose the assembly kernel variant (lanes per node; 1 and 0: one node per lane
        with LDS staging) and the lanes per row of the CSR residual kernels."""
        _capi.check(_capi.lib().vfvm_set_groups(self.handle(), asm_group, spmv_group), self.ctx.handle)

    def nodevolumes(self):
        v = np.zeros(self.grid.num_nodes)
        _capi.check(_capi.lib().vfvm_nodevolumes(self.handle(), _capi.dptr(v)), self.ctx.handle)
        return v

    def eval_rhs_device(self, u, du, t=0.0):
        """eval_rhs! on device memory: u and du are DeviceArrays of ndof entries."""
        _capi.check(_capi.lib().vfvm_eval_rhs_device(self.handle(), u.ptr, float(t), du.ptr), self.ctx.handle)

    def eval_jacobian_device(self, u, nzval_csr, t=0.0):
        """eval_jacobian! on device memory: the values of -dF/du in the row-major order of
        pattern_csr() into the DeviceArray nzval_csr (nnz entries)."""
        _capi.check(_capi.lib().vfvm_eval_jacobian_device(self.handle(), u.ptr, float(t), nzval_csr.ptr),
                    self.ctx.handle)

    def update_grid_device(self, coord):
        """update_grid! with the new coordinates (dim x nnodes, column-major) in a
        DeviceArray of the system's context; the cells and boundary faces stay those of
        the system's grid.  The grid object then holds the new coordinates (read back from
        the device), so a callable source and a later rebuild of the device system see
        the updated geometry.  Raises ValueError for an array of another size or context."""
        g = self.grid
        n = g.dim * g.num_nodes
        if getattr(coord, "n", None) != n or getattr(coord, "ptr", None) is None:
            raise ValueError(f"update_grid_device needs a DeviceArray of dim x nnodes = {n} coordinates")
        if getattr(coord, "ctx", None) is not self.ctx:
            raise ValueError("update_grid_device needs coordinates in the system's context")
        _capi.check(_capi.lib().vfvm_update_grid_device(self.handle(), coord.ptr), self.ctx.handle)
        import dataclasses
        xy = np.ascontiguousarray(coord.numpy().reshape((g.dim, g.num_nodes), order="F"))
        self.grid = dataclasses.replace(g, coord=xy)

    def update_grid(self, grid: Grid):
        """update_grid!(system; grid) (src/vfvm_system.jl:607-612): `grid` has the cells and
        boundary faces of the system's grid and new coordinates; the form factors are
        recomputed on the device.  Raises VFVMError (geometry unchanged) when an edge with
        no Jacobian entry (zero factor at creation) would get a nonzero factor."""
        g = self.grid

        def same(a, b):
            return a is b or np.array_equal(a, b)

        if (grid.dim != g.dim or grid.num_nodes != g.num_nodes or not same(grid.cellnodes, g.cellnodes)
                or not same(grid.bfacenodes, g.bfacenodes) or not same(grid.bfaceregions, g.bfaceregions)
                or grid.coordsys != g.coordsys):
            raise ValueError("update_grid needs the same cells and boundary faces (only coordinates may change)")
        coord = np.asfortranarray(grid.coord, dtype=np.float64)
        if self._h is not None:
            _capi.check(_capi.lib().vfvm_update_grid(self._h, _capi.dptr(coord)), self.ctx.handle)
        self.grid = grid

    def pattern(self):
        """CSC pattern, 1-based (colptr, rowval) as SparseMatrixCSC(state.matrix) holds it."""
        ndof, nnz, _ = self.sizes()
        cp = np.zeros(ndof + 1, np.int64)
        rv = np.zeros(max(nnz, 1), np.int64)
        _capi.check(_capi.lib().vfvm_pattern(self.handle(), _capi.lptr(cp), _capi.lptr(rv)))
        return cp, rv[:nnz]

    def pattern_csr(self):
        """The device CSR pattern, 0-based (rowptr, colind): the order of the *_device
        Jacobian values."""
        import ctypes as C
        ndof, nnz, _ = self.sizes()
        rp = np.zeros(ndof + 1, np.int32)
        ci = np.zeros(max(nnz, 1), np.int32)
        ip = C.POINTER(C.c_int32)
        _capi.check(_capi.lib().vfvm_pattern_csr(self.handle(), rp.ctypes.data_as(ip), ci.ctypes.data_as(ip)))
        return rp, ci[:nnz]


def unknowns(sys: System, inival=0.0):
    """unknowns(sys; inival): nspec x nnodes, Fortran order."""
    U = np.zeros((sys.num_species, sys.grid.num_nodes), order="F")
    U[...] = inival
    return U


def dofs(U):
    return np.asarray(U).reshape(-1, order="F")


def _as_dofs(sys, U):
    a = np.asarray(U, dtype=np.float64)
    if a.ndim == 2:
        a = a.reshape(-1, order="F")
    if a.size != sys.num_species * sys.grid.num_nodes:
        raise ValueError("wrong size of unknowns")
    return np.ascontiguousarray(a)


def _from_dofs(sys, v):
    return np.asfortranarray(v.reshape(sys.num_species, sys.grid.num_nodes, order="F"))


def evaluate_residual_and_jacobian(sys: System, u, t=0.0, tstep=math.inf, embed=0.0, init_dirichlet=False):
    """Residual and Jacobian at u (src/vfvm_solver.jl:224-260).  Returns
    (F as nspec x nnodes, scipy CSC matrix)."""
    import scipy.sparse as sp
    L = _capi.lib()
    h = sys.handle()
    U = _as_dofs(sys, u).copy()
    if init_dirichlet:
        _capi.check(L.vfvm_initialize_dirichlet(h, _capi.dptr(U), t, embed), sys.ctx.handle)
    ndof, nnz, _ = sys.sizes()
    F = np.zeros(ndof)
    nz = np.zeros(max(nnz, 1))
    _capi.check(L.vfvm_assemble(h, _capi.dptr(U), _capi.dptr(U), t, tstep, embed, _capi.dptr(F), _capi.dptr(nz)),
                 sys.ctx.handle)
    cp, rv = sys.pattern()
    A = sp.csc_matrix((nz[:nnz], rv - 1, cp - 1), shape=(ndof, ndof))
    return _from_dofs(sys, F), A


def linsolve(sys: System, b, control: Optional[SolverControl] = None, **kwargs):
    """_solve_linear! (src/vfvm_linsolve.jl:6-61) on the Jacobian of the system's last
    assembly (evaluate_residual_and_jacobian, or the last Newton step): x with J x = b and
    the solver statistics (iterations, converged, residual norms).  The preconditioner is
    built from that Jacobian.  Raises LinearSolverError if the tolerance is not reached."""
    c = _control(control, kwargs)
    rhs = _as_dofs(sys, b).copy()
    x = np.zeros_like(rhs)
    st = _capi.LinStatsT()
    rc = _capi.lib().vfvm_linsolve(sys.handle(), _capi.dptr(rhs), _capi.dptr(x), c.to_c(), st)
    stats = dict(iters=st.iters, refinements=st.refinements, converged=bool(st.converged),
                 fp64_fallback=bool(st.fp64_fallback), resnorm0=st.resnorm0, resnorm=st.resnorm)
    _capi.check(rc, sys.ctx.handle)
    return x, stats


def ilu0_factor(sys: System):
    """The ILU(0) factor of the last linear solve with precon "ilu0" or "ilu0_natural"
    (ILUZero.ILU0Precon's l and u), one value per stored entry of the system's CSR dof
    pattern (System.pattern_csr order): L strictly below the diagonal of the factor's
    ordering (unit diagonal not stored), U on and above it, of the Jacobian as assembled."""
    nnz = sys.sizes()[1]
    lu = np.zeros(max(nnz, 1))
    _capi.check(_capi.lib().vfvm_ilu0_factor(sys.handle(), _capi.dptr(lu)), sys.ctx.handle)
    return lu[:nnz]


def initialize_dirichlet(sys: System, u, t=0.0, embed=0.0):
    """_initialize_dirichlet!(U, system, data) (src/vfvm_system.jl:947-1003): a copy of u
    with the Dirichlet values of time t and embedding parameter `embed` set on the
    boundary nodes (the last boundary face visited sets a node on several regions)."""
    U = _as_dofs(sys, u).copy()
    _capi.check(_capi.lib().vfvm_initialize_dirichlet(sys.handle(), _capi.dptr(U), float(t), float(embed)),
                sys.ctx.handle)
    return _from_dofs(sys, U)


def eval_rhs(sys: System, u, t=0.0):
    """eval_rhs!(du, u, state, t) (src/vfvm_diffeq_interface.jl:31-36): du = -F(u) with
    F assembled at U = Uold = u, tstep = Inf.  u and the result are dof vectors."""
    U = np.ascontiguousarray(np.asarray(u, dtype=np.float64).reshape(-1))
    ndof, _, _ = sys.sizes()
    if U.size != ndof:
        raise ValueError(f"u must have {ndof} entries")
    du = np.zeros(ndof)
    _capi.check(_capi.lib().vfvm_eval_rhs(sys.handle(), _capi.dptr(U), float(t), _capi.dptr(du)), sys.ctx.handle)
    return du


def eval_jacobian(sys: System, u, t=0.0):
    """eval_jacobian!(J, u, state, t) (src/vfvm_diffeq_interface.jl:44-50): -dF/du at u as a
    scipy CSC matrix with the system's pattern."""
    import scipy.sparse as sp
    U = np.ascontiguousarray(np.asarray(u, dtype=np.float64).reshape(-1))
    ndof, nnz, _ = sys.sizes()
    if U.size != ndof:
        raise ValueError(f"u must have {ndof} entries")
    nz = np.zeros(max(nnz, 1))
    _capi.check(_capi.lib().vfvm_eval_jacobian(sys.handle(), _capi.dptr(U), float(t), _capi.dptr(nz)),
                sys.ctx.handle)
    cp, rv = sys.pattern()
    return sp.csc_matrix((nz[:nnz], rv - 1, cp - 1), shape=(ndof, ndof))


def mass_matrix(sys: System):
    """mass_matrix(state) (src/vfvm_diffeq_interface.jl:58-96): diagonal for every physics
    of the registry, returned as a scipy sparse diagonal matrix (the reference returns a
    Diagonal when the matrix is diagonal)."""
    import scipy.sparse as sp
    ndof, _, _ = sys.sizes()
    d = np.zeros(ndof)
    _capi.check(_capi.lib().vfvm_mass_matrix(sys.handle(), _capi.dptr(d)), sys.ctx.handle)
    return sp.diags(d, format="dia")


def _control(control, kwargs):
    c = SolverControl() if control is None else control
    c.update(**{k: v for k, v in kwargs.items() if k not in ("times", "embed")})
    return c


class TransientSolution:
    """Accepted times (or embedding parameters) t and solutions u (nspec x nnodes
    each), as the reference's TransientSolution (src/vfvm_transientsolution.jl)."""

    def __init__(self, t, u, history):
        self.t = np.asarray(t)
        self.u = list(u)
        self.history = history

    def __len__(self):
        return len(self.t)

    def __getitem__(self, i):
        return self.u[i]


def solve(sys: System, inival=0.0, control: Optional[SolverControl] = None, time=0.0, tstep=math.inf,
          times=None, embed=None, store_all=True, params=None, **kwargs):
    """solve(sys; inival, control, time, tstep, times, embed, params) (src/vfvm_solver.jl:551-668).
    params: [lambda], the value of the system parameter for a single solve.

    Stationary and implicit-Euler-step solves run the device Newton loop and return
    the solution (nspec x nnodes).  `times` (implicit Euler) and `embed` (parameter
    embedding) run solve_transient! with its step-size control (the Δt resp. Δp
    fields of the control; `fixed_timesteps` gives fixed steps) and return a
    TransientSolution with the initial value and every accepted step (only the last
    one with store_all=False)."""
    c = _control(control, kwargs)
    L = _capi.lib()
    h = sys.handle()
    if np.isscalar(inival):
        inival = unknowns(sys, inival)
    U0 = _as_dofs(sys, inival)
    cc = c.to_c(sys.is_linear)
    if times is not None or embed is not None:
        transient = times is not None
        lam = np.ascontiguousarray(np.asarray(times if transient else embed, dtype=np.float64))
        sc = c.to_step_c(transient)
        # output entries: the start and every accepted step (bounded by the smallest step)
        span = float(lam[-1] - lam[0]) if lam.size > 1 else 0.0
        max_out = int(min(1_000_000, span / max(sc.d_min, 1e-300) + 2 * lam.size + 8)) if store_all else 1
        out = np.zeros_like(U0)
        out_t = np.zeros(max_out)
        out_u = np.zeros((max_out, U0.size)) if store_all else None
        nst = C.c_int64()
        hist = _capi.HistoryT()
        rc = L.vfvm_solve_transient(h, _capi.dptr(U0), lam.size, _capi.dptr(lam), int(transient), C.byref(cc),
                                    C.byref(sc), _capi.dptr(out), max_out, _capi.dptr(out_t),
                                    _capi.dptr(out_u) if store_all else None, C.byref(nst), C.byref(hist))
        sys.history = NewtonSolverHistory(niter=hist.niter, nlin=hist.nlin, time=hist.time,
                                          residual_norm=hist.residual_norm, fp64_fallbacks=hist.fp64_fallbacks)
        _capi.check(rc, sys.ctx.handle)
        if not hist.converged:  # handle_exceptions: the steps done so far (src/vfvm_solver.jl:413-440)
            import warnings
            warnings.warn("solve: " + _capi.lib().vfvm_last_error(sys.ctx.handle).decode(errors="replace"))
        n = min(int(nst.value) + 1, max_out)  # entries: start + accepted steps
        if store_all:
            ts = list(out_t[:n])
            us = [_from_dofs(sys, out_u[k]) for k in range(n)]
        else:
            ts, us = [float(lam[-1])], [_from_dofs(sys, out)]
        return TransientSolution(ts, us, sys.history)
    U = np.zeros_like(U0)
    hist = _capi.HistoryT()
    lam = 0.0 if params is None else float(np.atleast_1d(params)[0])
    rc = L.vfvm_newton(h, _capi.dptr(U), _capi.dptr(U0), time, tstep, lam, C.byref(cc), C.byref(hist))
    sys.history = NewtonSolverHistory(nlin=hist.nlin, niter=hist.niter, time=hist.time,
                                      updatenorm=tuple(hist.updatenorm[:min(hist.niter, 64)]),
                                      residual_norm=hist.residual_norm, fp64_fallbacks=hist.fp64_fallbacks)
    _capi.check(rc, sys.ctx.handle)
    return _from_dofs(sys, U)


def integrate(sys: System, U):
    """integrate(sys, U)[ispec, region] for a single-region grid: sum_K |omega_K| U[i, K]
    (src/vfvm_postprocess.jl:18-97)."""
    vol = sys.nodevolumes()
    return (np.asarray(U) * vol[None, :]).sum(axis=1, keepdims=True)


def enable_boundary_species(sys, ispec, bregions):
    return sys.enable_boundary_species(ispec, bregions)


def enable_species(sys, ispec, regions=None):
    sys.enable_species(ispec, regions)


def boundary_dirichlet(sys, ispec, ibc, v):
    sys.boundary_dirichlet(ispec, ibc, v)


def boundary_neumann(sys, ispec, ibc, v):
    sys.boundary_neumann(ispec, ibc, v)


def boundary_robin(sys, ispec, ibc, alpha, v):
    sys.boundary_robin(ispec, ibc, alpha, v)


class Batch:
    """Many independent systems solved together (vfvm_batch_*): the reference's
    one-SystemState-per-thread pattern (examples/Example440_ParallelState.jl:36-45)
    packed into single device launches.  Members must share species count and
    physics kinds; parameters, sources and boundary values may differ."""

    def __init__(self, systems, autotune=True):
        self.systems = list(systems)
        self.ctx = self.systems[0].ctx
        L = _capi.lib()
        hs = (C.c_void_p * len(self.systems))(*[s.handle() for s in self.systems])
        h = C.c_void_p()
        _capi.check(L.vfvm_batch_create(self.ctx.handle, len(self.systems), hs, C.byref(h)), self.ctx.handle)
        self._h = h
        self.groups = None
        if autotune:
            a, g = C.c_int32(), C.c_int32()
            _capi.check(L.vfvm_batch_autotune(h, C.byref(a), C.byref(g)), self.ctx.handle)
            self.groups = (a.value, g.value)
        self.histories = []

    def __del__(self):
        if getattr(self, "_h", None) is not None and _capi._LIB is not None:
            _capi._LIB.vfvm_batch_destroy(self._h)
            self._h = None

    def set_graphs(self, enable: bool):
        """Run the batched Krylov iteration blocks as captured HIP graphs (default) or
        launch their kernels one by one."""
        _capi.check(_capi.lib().vfvm_batch_set_graphs(self._h, int(enable)), self.ctx.handle)

    def set_member_solver(self, block: int):
        """Batched Krylov solves as one block of `block` threads per member (256, 512 or
        1024: the member's whole BiCGStab with its AMG cycles in one block, one launch per
        linear solve), or 0: one launch per BiCGStab step for all members."""
        _capi.check(_capi.lib().vfvm_batch_set_member_solver(self._h, int(block)), self.ctx.handle)

    def set_compact_graph(self, enable: bool):
        """The batched assembly's node graph: compact (default where it fits) or plain."""
        _capi.check(_capi.lib().vfvm_batch_set_compact_graph(self._h, int(bool(enable))), self.ctx.handle)

    def set_sell_index(self, bits: int):
        """The batched SpMV's column indices: 16-bit offsets or 32-bit columns."""
        _capi.check(_capi.lib().vfvm_batch_set_sell_index(self._h, int(bits)), self.ctx.handle)

    def sizes(self):
        """Total dofs and the member offsets of the packed (member-order) dof vector."""
        nd = C.c_int64()
        off = np.zeros(len(self.systems) + 1, dtype=np.int64)
        _capi.check(_capi.lib().vfvm_batch_sizes(self._h, C.byref(nd), off.ctypes.data_as(_capi._lp)),
                    self.ctx.handle)
        return nd.value, off

    def solve_device(self, U_dev, Uold_dev, control: Optional[SolverControl] = None, time=0.0, tstep=math.inf,
                     embed=0.0, raise_on_failure=True, **kwargs):
        """solve for every member with start values Uold_dev and solutions U_dev in device
        memory (vfvm.DeviceArray of the packed member-order dofs, see sizes())."""
        c = _control(control, kwargs)
        n = len(self.systems)
        hist = (_capi.HistoryT * n)()
        cc = c.to_c(self.systems[0].is_linear)
        rc = _capi.lib().vfvm_batch_newton_device(self._h, U_dev.ptr, Uold_dev.ptr, time, tstep, embed, C.byref(cc),
                                                  hist)
        self.histories = [NewtonSolverHistory(nlin=h.nlin, niter=h.niter, time=h.time,
                                              updatenorm=tuple(h.updatenorm[:min(h.niter, 64)]),
                                              residual_norm=h.residual_norm) for h in hist]
        if raise_on_failure:
            _capi.check(rc, self.ctx.handle)
        return U_dev

    def update_member(self, m, source=None, bcondition=None):
        """Replace member m's source and/or boundary conditions inside the batch."""
        s = self.systems[m]
        src_p = None
        if source is not None:
            s.physics.source = source
            arr = s._source_array()
            src_p = _capi.dptr(arr)
        tab = None
        if bcondition is not None:
            s.physics.bcondition = list(bcondition)
            tab = s._bc_table()
        _capi.check(_capi.lib().vfvm_batch_update_member(self._h, m, src_p, tab), self.ctx.handle)

    def solve(self, inivals=0.0, control: Optional[SolverControl] = None, time=0.0, tstep=math.inf, embed=0.0,
              raise_on_failure=True):
        """solve_step! for every member; returns the list of solutions (nspec x nnodes)."""
        c = SolverControl() if control is None else control
        n = len(self.systems)
        if np.isscalar(inivals):
            inivals = [unknowns(s, inivals) for s in self.systems]
        U0 = [np.ascontiguousarray(_as_dofs(s, u)) for s, u in zip(self.systems, inivals)]
        U = [np.zeros_like(u) for u in U0]
        pu = (_capi._dp * n)(*[_capi.dptr(u) for u in U])
        p0 = (_capi._dp * n)(*[_capi.dptr(u) for u in U0])
        hist = (_capi.HistoryT * n)()
        cc = c.to_c(self.systems[0].is_linear)
        rc = _capi.lib().vfvm_batch_newton(self._h, pu, p0, time, tstep, embed, C.byref(cc), hist)
        self.histories = [NewtonSolverHistory(nlin=h.nlin, niter=h.niter, time=h.time,
                                              updatenorm=tuple(h.updatenorm[:min(h.niter, 64)]),
                                              residual_norm=h.residual_norm) for h in hist]
        if raise_on_failure:
            _capi.check(rc, self.ctx.handle)
        return [_from_dofs(s, u) for s, u in zip(self.systems, U)]
