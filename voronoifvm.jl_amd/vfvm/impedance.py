"""Small-signal impedance: host mirror of src/vfvm_impedance.jl and the test
functions of src/vfvm_testfunctions.jl.  The work is done by libvfvm_hip.so
(vfvm_impedance_*, vfvm_measurement_derivatives): the Jacobian and the
measurement rows are assembled on the GPU and the frequency-domain systems of a
whole sweep are solved there together."""
import ctypes as C
import math

import numpy as np

from . import _capi
from .control import SolverControl
from .physics import LinearFlux
from .system import System, dofs, solve


def testfunction(sys: System, bc0, bc1, reltol_linear=1e-12):
    """testfunction(TestFunctionFactory(sys), bc0, bc1) (src/vfvm_testfunctions.jl:46-106):
    T with -Laplace T = 0, T = 0 on the boundary regions bc0, T = 1 on bc1, homogeneous
    Neumann elsewhere; nnodes values."""
    tf = System(sys.grid, flux=LinearFlux(1.0), species=[1], is_linear=True, context=sys.ctx)
    for r in bc1:
        tf.boundary_dirichlet(1, r, 1.0)
    for r in bc0:
        tf.boundary_dirichlet(1, r, 0.0)
    c = SolverControl(reltol_linear=reltol_linear, abstol_linear=0.0, maxiters_linear=20000, precon="ilu0")
    T = solve(tf, inival=0.0, control=c)
    return np.ascontiguousarray(dofs(T), dtype=np.float64)


def integrate_testfunction(sys: System, T, U, Uold=None, tstep=math.inf, params=None):
    """integrate(system, T, U) for a stationary solution (src/vfvm_testfunctions.jl:110-130),
    or integrate(system, T, U, Uold, tstep) for an implicit-Euler step (lines 194-210): per
    species I_flux + I_react - I_src (+ (I_stor - I_oldstor) / tstep) weighted by the test
    function T; with T from testfunction(sys, bc0, bc1) the flux through the bc1 boundary.
    params: [lambda], the system parameter (default: the reference's embedding value 1)."""
    u = _dofvec(U)
    uo = None if Uold is None else _dofvec(Uold)
    T = np.ascontiguousarray(T, dtype=np.float64)
    out = np.zeros(sys.num_species)
    lam = 1.0 if params is None else float(np.atleast_1d(params)[0])
    _capi.check(_capi.lib().vfvm_integrate_testfunction_at(sys.handle(), _capi.dptr(u),
                                                           None if uo is None else _capi.dptr(uo), float(tstep),
                                                           _capi.dptr(T), lam, _capi.dptr(out)), sys.ctx.handle)
    return out


def measurement_derivatives(sys: System, U0, T, spec, params=None):
    """measurement_derivative (src/vfvm_impedance.jl:235-252) of
    U -> -integrate_stdy(sys, T, U)[spec] and U -> -integrate_tran(sys, T, U)[spec] at U0
    (and the system parameter params[0], default 0): two dof vectors (dmeas_stdy, dmeas_tran)."""
    ndof, _, _ = sys.sizes()
    u = np.ascontiguousarray(dofs(U0) if np.ndim(U0) == 2 else np.asarray(U0, dtype=np.float64).reshape(-1))
    T = np.ascontiguousarray(T, dtype=np.float64)
    ds, dt = np.zeros(ndof), np.zeros(ndof)
    lam = 0.0 if params is None else float(np.atleast_1d(params)[0])
    _capi.check(_capi.lib().vfvm_measurement_derivatives_at(sys.handle(), _capi.dptr(u), _capi.dptr(T), int(spec),
                                                            lam, _capi.dptr(ds), _capi.dptr(dt)), sys.ctx.handle)
    return ds, dt


def _dofvec(U):
    return np.ascontiguousarray(dofs(U) if np.ndim(U) == 2 else np.asarray(U, dtype=np.float64).reshape(-1))


def assemble_dudp(sys: System):
    """dF/dlambda for the system's parameter lambda (the reference's state.dudp[1] with
    params = [lambda]; src/vfvm_assembly.jl:116-120, 194-198, 398-405): a dof vector."""
    ndof, _, _ = sys.sizes()
    d = np.zeros(ndof)
    _capi.check(_capi.lib().vfvm_assemble_dudp(sys.handle(), _capi.dptr(d)), sys.ctx.handle)
    return d


def assemble_dudp_at(sys: System, U, lam=0.0, with_bc=True):
    """state.dudp[1] = dF/dlambda at U and lambda, including the flux and reaction terms
    whose coefficients depend on lambda (Physics.dflux / dreaction); with_bc=False leaves
    out the boundary conditions (the derivative of a test-function integral)."""
    u = _dofvec(U)
    d = np.zeros(u.size)
    _capi.check(_capi.lib().vfvm_assemble_dudp_at(sys.handle(), _capi.dptr(u), float(lam), int(bool(with_bc)),
                                                  _capi.dptr(d)), sys.ctx.handle)
    return d


def parameter_derivative(sys: System, U, lam=0.0, control=None):
    """du/dlambda = -J(U)^-1 dF/dlambda at a stationary solution U for parameter lam."""
    c = SolverControl() if control is None else control
    cc = c.to_c(sys.is_linear)
    u = _dofvec(U)
    out = np.zeros(u.size)
    _capi.check(_capi.lib().vfvm_parameter_derivative(sys.handle(), _capi.dptr(u), float(lam), C.byref(cc),
                                                      _capi.dptr(out)), sys.ctx.handle)
    return out


class ImpedanceSystem:
    """ImpedanceSystem(system, U0, excited_spec, excited_bc) (src/vfvm_impedance.jl:159-186),
    or, without excited_spec / excited_bc, ImpedanceSystem(system, U0; lambda0) (lines
    49-152) excited through the system parameter lambda (F = -dF/dlambda)."""

    def __init__(self, sys: System, U0, excited_spec=None, excited_bc=None, lambda0=0.0):
        self.sys = sys
        self.U0 = U0
        u = _dofvec(U0)
        h = C.c_void_p()
        if excited_spec is None and excited_bc is None:
            rc = _capi.lib().vfvm_impedance_create_param(sys.handle(), _capi.dptr(u), float(lambda0), C.byref(h))
        else:
            rc = _capi.lib().vfvm_impedance_create(sys.handle(), _capi.dptr(u), int(excited_spec), int(excited_bc),
                                                   C.byref(h))
        _capi.check(rc, sys.ctx.handle)
        self._h = h
        self.ndof = u.size
        self.iterations = None

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value and _capi._LIB is not None:
            _capi._LIB.vfvm_impedance_destroy(self._h)
        self._h = None

    def __del__(self):
        self.close()

    def _sweep(self, omegas, dstdy=None, dtran=None, reltol=1e-12, maxiters=20000, want_uz=False):
        om = np.ascontiguousarray(np.atleast_1d(np.asarray(omegas, dtype=np.float64)))
        nw = om.size
        uz = np.zeros((nw, self.ndof), dtype=np.complex128) if want_uz else None
        z = np.zeros(nw, dtype=np.complex128) if dstdy is not None else None
        its = np.zeros(nw, dtype=np.int32)
        ds = None if dstdy is None else np.ascontiguousarray(dstdy, dtype=np.float64)
        dt = None if dtran is None else np.ascontiguousarray(dtran, dtype=np.float64)
        vp = lambda a: None if a is None else a.ctypes.data_as(C.c_void_p)
        _capi.check(_capi.lib().vfvm_impedance_sweep(self._h, nw, _capi.dptr(om), vp(ds), vp(dt), float(reltol),
                                                     int(maxiters), vp(uz), vp(z),
                                                     its.ctypes.data_as(C.POINTER(C.c_int32))),
                    self.sys.ctx.handle)
        self.iterations = its
        return uz, z

    def solve(self, omega, reltol=1e-12, maxiters=20000):
        """solve!(UZ, isys, w) (src/vfvm_impedance.jl:205-232): the complex dof vector UZ."""
        uz, _ = self._sweep([omega], reltol=reltol, maxiters=maxiters, want_uz=True)
        return uz[0]

    def impedances(self, omegas, dmeas_stdy, dmeas_tran, reltol=1e-12, maxiters=20000):
        """impedance(isys, w, U0, dstdy, dtran) (src/vfvm_impedance.jl:268-288) for every w
        of the sweep, solved together on the GPU."""
        _, z = self._sweep(omegas, dmeas_stdy, dmeas_tran, reltol=reltol, maxiters=maxiters)
        return z


def impedance(isys: ImpedanceSystem, omega, U0, dmeas_stdy, dmeas_tran):
    """impedance(impedance_system, w, U0, dmeas_stdy, dmeas_tran) for one frequency."""
    return isys.impedances([omega], dmeas_stdy, dmeas_tran)[0]
