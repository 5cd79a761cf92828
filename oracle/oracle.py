"""ctypes loader for the CPU Float64 restatement (liboracle): TEST INFRASTRUCTURE.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module.  It is the checker, never the thing measured or shipped.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

FLUX = {"none": 0, "linear": 1, "power": 2, "nldiff": 3, "dd": 4}
REACT = {"none": 0, "sinh": 1, "pow": 2, "exchange": 3, "dd": 4}
STOR = {"none": 0, "identity": 1}
BC = {"none": 0, "dirichlet": 1, "dirichlet_legacy": 2, "robin": 3}
LIN = {"direct": 0, "bicgstab_jacobi": 1, "bicgstab_ilu0": 2, "gmres_blockjacobi": 3}
CELLWISE, EDGEWISE = 0, 1


class Physics(C.Structure):
    _fields_ = [("flux", C.c_int32), ("reaction", C.c_int32), ("storage", C.c_int32),
                ("pad", C.c_int32), ("p", C.c_double * 16)]


class BCEntry(C.Structure):
    _fields_ = [("kind", C.c_int32), ("pad", C.c_int32), ("value", C.c_double),
                ("value_lambda", C.c_double), ("factor", C.c_double)]


class Control(C.Structure):
    _fields_ = [("abstol", C.c_double), ("reltol", C.c_double), ("tol_round", C.c_double),
                ("tol_mono", C.c_double), ("damp_initial", C.c_double), ("damp_growth", C.c_double),
                ("maxiters", C.c_int32), ("max_round", C.c_int32), ("updatecontrol", C.c_int32),
                ("method_linear", C.c_int32), ("reltol_linear", C.c_double),
                ("abstol_linear", C.c_double), ("maxiters_linear", C.c_int32),
                ("gmres_restart", C.c_int32), ("is_linear", C.c_int32),
                ("factorize_every_newtonstep", C.c_int32), ("factorize_every_timestep", C.c_int32),
                ("pad", C.c_int32), ("dt", C.c_double), ("dt_min", C.c_double),
                ("dt_max", C.c_double), ("dt_grow", C.c_double), ("dt_decrease", C.c_double),
                ("dp", C.c_double), ("dp_min", C.c_double), ("dp_max", C.c_double),
                ("dp_grow", C.c_double), ("dp_decrease", C.c_double), ("du_opt", C.c_double),
                ("du_max_factor", C.c_double), ("num_final_steps", C.c_int32),
                ("force_first_step", C.c_int32), ("handle_exceptions", C.c_int32), ("pad2", C.c_int32)]


class BPhysics(C.Structure):
    """orc_bphysics: breaction matrix, bflux coefficients, bstorage bits, fields flag."""
    _fields_ = [("react", (C.c_double * 3) * 3), ("bflux", C.c_double * 3), ("storage", C.c_int32),
                ("fields", C.c_int32)]


SPEC = {"bulk": 0, "boundary": 1, "none": 2}


class History(C.Structure):
    _fields_ = [("niter", C.c_int32), ("nlin_total", C.c_int32), ("nlu", C.c_int32),
                ("converged", C.c_int32), ("last_update_norm", C.c_double),
                ("updatenorm", C.c_double * 128)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libvfvm_oracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle library not built: {path} (run `make -C oracle`)")
        L = C.CDLL(path)
        dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
        ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
        lp = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.c_int, C.c_int, dp, C.c_int, ip, ip, C.c_int, ip, ip, C.c_int,
                                 C.POINTER(Physics), C.c_int, C.POINTER(BCEntry), C.c_int]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_set_source.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_set_bc.argtypes = [C.c_void_p, C.POINTER(BCEntry)]
        L.orc_set_species.argtypes = [C.c_void_p, ip, np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")]
        L.orc_set_bphysics.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_set_bfields.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_set_rcoef.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_set_param_dependence.argtypes = [C.c_void_p, C.c_void_p]
        L.orc_set_generic.argtypes = [C.c_void_p, C.c_int, np.ctypeslib.ndpointer(np.int32, flags="C"),
                                      np.ctypeslib.ndpointer(np.int32, flags="C"), dp]
        L.orc_node_species.argtypes = [C.c_void_p, np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")]
        L.orc_last_error.restype = C.c_char_p
        L.orc_last_error.argtypes = [C.c_void_p]
        L.orc_num_edges.argtypes = [C.c_void_p]
        L.orc_edges.argtypes = [C.c_void_p, ip, dp]
        L.orc_nodevolumes.argtypes = [C.c_void_p, dp]
        L.orc_assemble.argtypes = [C.c_void_p, dp, dp, C.c_double, C.c_double, C.c_double, dp]
        L.orc_matrix_nnz.restype = C.c_int64
        L.orc_matrix_nnz.argtypes = [C.c_void_p]
        L.orc_matrix_csc.argtypes = [C.c_void_p, lp, ip, dp]
        L.orc_matrix_reset.argtypes = [C.c_void_p]
        L.orc_initialize_dirichlet.argtypes = [C.c_void_p, dp, C.c_double, C.c_double]
        L.orc_solve_step.argtypes = [C.c_void_p, dp, dp, C.POINTER(Control), C.c_double, C.c_double,
                                     C.c_double, C.POINTER(History)]
        L.orc_solve_transient.argtypes = [C.c_void_p, dp, C.c_int, dp, C.c_int, C.POINTER(Control),
                                          C.c_double, dp, C.POINTER(C.c_int), C.POINTER(C.c_int), dp, C.c_int]
        L.orc_linsolve.argtypes = [C.c_void_p, C.POINTER(Control), dp, dp, C.POINTER(C.c_int)]
        L.orc_ilu0.argtypes = [C.c_int, np.ctypeslib.ndpointer(np.int64, flags="C"),
                               np.ctypeslib.ndpointer(np.int32, flags="C"), dp, dp]
        L.orc_default_control.argtypes = [C.POINTER(Control)]
        L.orc_bernoulli.restype = C.c_double
        L.orc_bernoulli.argtypes = [C.c_double]
        L.orc_bernoulli_pm.argtypes = [C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_bernoulli_dual.argtypes = [C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_cellfactors.argtypes = [C.c_int, dp, ip, dp, dp]
        L.orc_bfacefactors.argtypes = [C.c_int, dp, ip, dp]
        _LIB = L
    return _LIB


def control(**kw) -> Control:
    c = Control()
    lib().orc_default_control(C.byref(c))
    for k, v in kw.items():
        if k == "method_linear" and isinstance(v, str):
            v = LIN[v]
        setattr(c, k, v)
    return c


def physics(flux="none", reaction="none", storage="none", **params) -> Physics:
    """Parameter slots follow ORC_P_* in vfvm_oracle.h."""
    slots = {"D0": 0, "D1": 1, "D2sp": 2, "d2": 3, "m": 4, "k": 5, "rm": 6, "lambda2": 7,
             "zn": 8, "zp": 9, "ni2": 10, "ni": 11, "embed_source": 12}
    ph = Physics()
    ph.flux, ph.reaction, ph.storage = FLUX[flux], REACT[reaction], STOR[storage]
    for k, v in params.items():
        ph.p[slots[k]] = float(v)
    return ph


class OracleSystem:
    """One restated VoronoiFVM system (grid + physics + BC table)."""

    def __init__(self, grid, nspec, phys: Physics, bcs=None, assembly=EDGEWISE, source=None, species=None,
                 bphysics=None, bfields=None):
        """species: per species ("bulk",) / ("boundary", [bregions]) / ("none",) (default: all
        bulk); bphysics: {bregion: dict(react=3x3, bflux=3, storage=[species 0-based], fields=bool)};
        bfields: (bsrc, bcoef), each nspec x nnodes species fastest or None."""
        self._init(grid, nspec, phys, bcs, assembly, source)
        if species is not None:
            kind = np.zeros(nspec, np.int32)
            breg = np.zeros(nspec, np.uint32)
            for i, sp in enumerate(species):
                kind[i] = SPEC[sp[0]]
                for r in (sp[1] if len(sp) > 1 else ()):
                    breg[i] |= np.uint32(1 << (r - 1))
            lib().orc_set_species(self._h, kind, breg)
        if bphysics is not None:
            arr = (BPhysics * self.nbreg)()
            for r, d in bphysics.items():
                e = arr[r - 1]
                for i, row in enumerate(d.get("react", ())):
                    for j, v in enumerate(row):
                        e.react[i][j] = float(v)
                for i, v in enumerate(d.get("bflux", ())):
                    e.bflux[i] = float(v)
                for i in d.get("storage", ()):
                    e.storage |= 1 << i
                e.fields = int(bool(d.get("fields", False)))
            self._bphys = arr
            lib().orc_set_bphysics(self._h, C.cast(arr, C.c_void_p))
        if bfields is not None:
            self._bf = [None if a is None else np.ascontiguousarray(a, dtype=np.float64).ravel() for a in bfields]
            lib().orc_set_bfields(self._h, *[None if a is None else a.ctypes.data for a in self._bf])

    def set_rcoef(self, rcoef):
        """per-node bulk reaction coefficient (nspec x nnodes, species fastest) or None"""
        self._rc = None if rcoef is None else np.ascontiguousarray(rcoef, dtype=np.float64).ravel()
        lib().orc_set_rcoef(self._h, None if self._rc is None else self._rc.ctypes.data)

    def set_param_dependence(self, dp):
        """registry parameter slots p_j(lambda) = p_j + dp[j] lambda (16 values) or None"""
        self._dp = None if dp is None else np.ascontiguousarray(dp, dtype=np.float64)
        lib().orc_set_param_dependence(self._h, None if self._dp is None else self._dp.ctypes.data)

    def set_generic(self, matrix):
        """generic operator as a constant matrix over the dofs (scipy sparse or dense):
        F += G U, J += G"""
        import scipy.sparse as sp
        A = sp.coo_matrix(matrix)
        self._gen = (np.ascontiguousarray(A.row, dtype=np.int32), np.ascontiguousarray(A.col, dtype=np.int32),
                     np.ascontiguousarray(A.data, dtype=np.float64))
        lib().orc_set_generic(self._h, A.nnz, *self._gen)

    def node_species(self):
        m = np.zeros(self.grid.num_nodes, np.uint8)
        lib().orc_node_species(self._h, m)
        return m

    def _init(self, grid, nspec, phys, bcs, assembly, source):
        L = lib()
        self.grid = grid
        self.nspec = nspec
        self.nbreg = max(grid.num_bfaceregions, 1)
        self.bc = (BCEntry * (self.nbreg * nspec))()
        for (breg, ispec), spec in (bcs or {}).items():
            e = self.bc[(breg - 1) * nspec + ispec]
            e.kind = BC[spec[0]]
            e.value = float(spec[1]) if len(spec) > 1 else 0.0
            e.factor = float(spec[2]) if len(spec) > 2 else 0.0
            e.value_lambda = float(spec[3]) if len(spec) > 3 else 0.0
        self.phys = phys
        self._h = L.orc_create(grid.dim, grid.num_nodes, np.ascontiguousarray(grid.coord.T).ravel(),
                               grid.num_cells, np.ascontiguousarray(grid.cellnodes.T).ravel(),
                               np.ascontiguousarray(grid.cellregions), grid.num_bfaces,
                               np.ascontiguousarray(grid.bfacenodes.T).ravel(),
                               np.ascontiguousarray(grid.bfaceregions), nspec, C.byref(phys),
                               self.nbreg, self.bc, assembly)
        if not self._h:
            raise RuntimeError("orc_create failed")
        self._src = None
        if source is not None:
            self.set_source(source)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_destroy(self._h)
            self._h = None

    @property
    def ndof(self):
        return self.nspec * self.grid.num_nodes

    def set_source(self, src):
        src = np.ascontiguousarray(np.asarray(src, dtype=np.float64).reshape(self.ndof))
        self._src = src
        lib().orc_set_source(self._h, src.ctypes.data)

    def set_bc_value(self, breg, ispec, value):
        self.bc[(breg - 1) * self.nspec + ispec].value = float(value)
        lib().orc_set_bc(self._h, self.bc)

    def error(self):
        return lib().orc_last_error(self._h).decode()

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(f"oracle error {rc}: {self.error()}")

    def edges(self):
        E = lib().orc_num_edges(self._h)
        en = np.zeros(2 * E, np.int32)
        ef = np.zeros(E, np.float64)
        lib().orc_edges(self._h, en, ef)
        return en.reshape(E, 2), ef

    def nodevolumes(self):
        v = np.zeros(self.grid.num_nodes)
        lib().orc_nodevolumes(self._h, v)
        return v

    def assemble(self, U, Uold=None, time=0.0, tstep=np.inf, embed=0.0):
        U = np.ascontiguousarray(U, dtype=np.float64).ravel()
        Uold = U if Uold is None else np.ascontiguousarray(Uold, dtype=np.float64).ravel()
        F = np.zeros(self.ndof)
        self._check(lib().orc_assemble(self._h, U, Uold, time, tstep, embed, F))
        return F

    def matrix_csc(self):
        """(colptr, rowval, nzval), 0-based, rows sorted within columns."""
        nnz = lib().orc_matrix_nnz(self._h)
        cp = np.zeros(self.ndof + 1, np.int64)
        rv = np.zeros(max(nnz, 1), np.int32)
        nz = np.zeros(max(nnz, 1), np.float64)
        lib().orc_matrix_csc(self._h, cp, rv, nz)
        return cp, rv[:nnz], nz[:nnz]

    def matrix_scipy(self):
        import scipy.sparse as sp
        cp, rv, nz = self.matrix_csc()
        return sp.csc_matrix((nz, rv, cp), shape=(self.ndof, self.ndof))

    def evaluate_residual_and_jacobian(self, U, time=0.0, tstep=np.inf, embed=0.0, init_dirichlet=False):
        U = np.array(U, dtype=np.float64).ravel()
        if init_dirichlet:
            lib().orc_initialize_dirichlet(self._h, U, time, embed)
        F = self.assemble(U, U, time, tstep, embed)
        return F, self.matrix_scipy()

    def initialize_dirichlet(self, U, time=0.0, embed=0.0):
        U = np.array(U, dtype=np.float64).ravel()
        lib().orc_initialize_dirichlet(self._h, U, time, embed)
        return U

    def solve_step(self, inival, control=None, time=0.0, tstep=np.inf, embed=0.0):
        control = control or globals()["control"]()
        Uold = np.ascontiguousarray(np.asarray(inival, dtype=np.float64).ravel())
        U = np.zeros(self.ndof)
        h = History()
        self._check(lib().orc_solve_step(self._h, U, Uold, C.byref(control), time, tstep, embed, C.byref(h)))
        return U, h

    def solve_transient(self, inival, lambdas, control=None, transient=True, time=0.0, return_steps=False,
                        max_steps=100000):
        """solve_transient! (step control of src/vfvm_solver.jl:267-537).  Returns the
        final solution, the number of accepted steps, the Newton iterations in total
        [and the accepted times / parameters]."""
        control = control or globals()["control"]()
        inival = np.ascontiguousarray(np.asarray(inival, dtype=np.float64).ravel())
        lam = np.ascontiguousarray(np.asarray(lambdas, dtype=np.float64))
        out = np.zeros(self.ndof)
        steps = np.zeros(max_steps)
        ns, nn = C.c_int(0), C.c_int(0)
        self._check(lib().orc_solve_transient(self._h, inival, lam.size, lam, int(transient), C.byref(control),
                                              time, out, C.byref(ns), C.byref(nn), steps, max_steps))
        if return_steps:
            return out, ns.value, nn.value, steps[:min(ns.value, max_steps)].copy()
        return out, ns.value, nn.value

    def linsolve(self, b, control=None):
        control = control or globals()["control"]()
        x = np.zeros(self.ndof)
        it = C.c_int(0)
        self._check(lib().orc_linsolve(self._h, C.byref(control),
                                       np.ascontiguousarray(b, dtype=np.float64), x, C.byref(it)))
        return x, it.value


def ilu0(rowptr, colind, val):
    """ILUZero's natural-order ilu0 of a CSR matrix with sorted rows: the L (strictly
    lower, unit diagonal implied) and U values in the stored positions."""
    rp = np.ascontiguousarray(rowptr, dtype=np.int64)
    ci = np.ascontiguousarray(colind, dtype=np.int32)
    v = np.ascontiguousarray(val, dtype=np.float64)
    lu = np.zeros(v.size)
    rc = lib().orc_ilu0(rp.size - 1, rp, ci, v, lu)
    if rc:
        raise RuntimeError(f"oracle ilu0: error {rc}")
    return lu


def bernoulli(x: float) -> float:
    return lib().orc_bernoulli(float(x))


def bernoulli_pm(x: float):
    a, b = C.c_double(), C.c_double()
    lib().orc_bernoulli_pm(float(x), C.byref(a), C.byref(b))
    return a.value, b.value


def bernoulli_dual(x: float):
    a, b = C.c_double(), C.c_double()
    lib().orc_bernoulli_dual(float(x), C.byref(a), C.byref(b))
    return a.value, b.value
