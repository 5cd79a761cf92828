"""ctypes binding of libvfvm_hip.so (the C ABI declared in include/vfvm_hip.h).

The library is the product; if it is missing or no GPU is present, every
compute call raises -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBPATH = os.path.join(_PKG, "libvfvm_hip.so")
_LIB = None

OK, ERR_CONVERGENCE, ERR_ASSEMBLY, ERR_LINEAR, ERR_ARG, ERR_HIP = range(6)


class PhysicsT(C.Structure):
    _fields_ = [("flux", C.c_int32), ("reaction", C.c_int32), ("storage", C.c_int32), ("pad", C.c_int32),
                ("p", C.c_double * 16)]


class BCT(C.Structure):
    _fields_ = [("kind", C.c_int32), ("pad", C.c_int32), ("value", C.c_double), ("value_lambda", C.c_double),
                ("factor", C.c_double)]


class SpeciesT(C.Structure):
    _fields_ = [("kind", C.c_int32), ("bregions", C.c_uint32)]


class BPhysT(C.Structure):
    _fields_ = [("react", (C.c_double * 3) * 3), ("bflux", C.c_double * 3), ("storage", C.c_int32),
                ("fields", C.c_int32)]


SPECIES_BULK, SPECIES_BOUNDARY, SPECIES_NONE = range(3)
ASSEMBLY = {"edgewise": 0, "cellwise": 1}


class ControlT(C.Structure):
    _fields_ = [("abstol", C.c_double), ("reltol", C.c_double), ("tol_round", C.c_double), ("tol_mono", C.c_double),
                ("damp_initial", C.c_double), ("damp_growth", C.c_double), ("maxiters", C.c_int32),
                ("max_round", C.c_int32), ("updatecontrol", C.c_int32), ("is_linear", C.c_int32),
                ("reltol_linear", C.c_double), ("abstol_linear", C.c_double), ("maxiters_linear", C.c_int32),
                ("krylov", C.c_int32), ("precon", C.c_int32), ("precision", C.c_int32),
                ("gmres_restart", C.c_int32), ("max_refinements", C.c_int32), ("reuse_precon", C.c_int32),
                ("check_every", C.c_int32), ("inner_reltol", C.c_double)]


class HistoryT(C.Structure):
    _fields_ = [("niter", C.c_int32), ("nlin", C.c_int32), ("nlu", C.c_int32), ("converged", C.c_int32),
                ("time", C.c_double), ("tasm", C.c_double), ("tlinsolve", C.c_double),
                ("last_update_norm", C.c_double), ("residual_norm", C.c_double), ("updatenorm", C.c_double * 64),
                ("fp64_fallbacks", C.c_int32), ("pad", C.c_int32)]


class StepControlT(C.Structure):
    _fields_ = [("d", C.c_double), ("d_min", C.c_double), ("d_max", C.c_double), ("d_grow", C.c_double),
                ("d_decrease", C.c_double), ("du_opt", C.c_double), ("du_max_factor", C.c_double),
                ("num_final_steps", C.c_int32), ("force_first_step", C.c_int32),
                ("handle_exceptions", C.c_int32), ("pad", C.c_int32)]


class LinStatsT(C.Structure):
    _fields_ = [("iters", C.c_int32), ("refinements", C.c_int32), ("converged", C.c_int32),
                ("fp64_fallback", C.c_int32),
                ("resnorm0", C.c_double), ("resnorm", C.c_double)]


# exported entry points and their ctypes signatures (kept in the order of include/vfvm_hip.h)
_vp, _dp, _lp, _ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int64), C.POINTER(C.c_int32)
SIGNATURES = {
    "vfvm_default_control": (None, [C.POINTER(ControlT)]),
    "vfvm_ctx_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
    "vfvm_ctx_destroy": (C.c_int, [_vp]),
    "vfvm_last_error": (C.c_char_p, [_vp]),
    "vfvm_device_alloc": (C.c_int, [_vp, C.c_int64, C.POINTER(_vp)]),
    "vfvm_device_free": (C.c_int, [_vp, _vp]),
    "vfvm_device_copy": (C.c_int, [_vp, _vp, _vp, C.c_int64, C.c_int]),
    "vfvm_system_create": (C.c_int, [_vp, C.c_int, C.c_int64, _dp, C.c_int64, _lp, _lp, C.c_int64, _lp, _lp, C.c_int,
                                     C.POINTER(PhysicsT), C.c_int, C.POINTER(BCT), C.POINTER(_vp)]),
    "vfvm_system_create_cs": (C.c_int, [_vp, C.c_int, C.c_int64, _dp, C.c_int64, _lp, _lp, C.c_int64, _lp, _lp,
                                        C.c_int, C.POINTER(PhysicsT), C.c_int, C.POINTER(BCT), C.c_int,
                                        C.POINTER(_vp)]),
    "vfvm_system_create_ex": (C.c_int, [_vp, C.c_int, C.c_int64, _dp, C.c_int64, _lp, _lp, C.c_int64, _lp, _lp,
                                        C.c_int, C.POINTER(PhysicsT), C.c_int, C.POINTER(BCT), C.c_int,
                                        C.POINTER(SpeciesT), C.POINTER(BPhysT), C.c_int, C.POINTER(_vp)]),
    "vfvm_system_set_fields": (C.c_int, [_vp, _dp, _dp, _dp]),
    "vfvm_node_species": (C.c_int, [_vp, C.POINTER(C.c_uint8)]),
    "vfvm_system_destroy": (C.c_int, [_vp]),
    "vfvm_system_set_source": (C.c_int, [_vp, _dp]),
    "vfvm_system_set_bc": (C.c_int, [_vp, C.POINTER(BCT)]),
    "vfvm_system_sizes": (C.c_int, [_vp, _lp, _lp, _lp]),
    "vfvm_nodevolumes": (C.c_int, [_vp, _dp]),
    "vfvm_update_grid": (C.c_int, [_vp, _dp]),
    "vfvm_update_grid_device": (C.c_int, [_vp, _vp]),
    "vfvm_ctx_set_setup": (C.c_int, [_vp, C.c_int]),
    "vfvm_impedance_create": (C.c_int, [_vp, _dp, C.c_int, C.c_int, C.POINTER(_vp)]),
    "vfvm_impedance_destroy": (C.c_int, [_vp]),
    "vfvm_impedance_create_param": (C.c_int, [_vp, _dp, C.c_double, C.POINTER(_vp)]),
    "vfvm_assemble_dudp": (C.c_int, [_vp, _dp]),
    "vfvm_system_set_param_dependence": (C.c_int, [_vp, _dp]),
    "vfvm_assemble_dudp_at": (C.c_int, [_vp, _dp, C.c_double, C.c_int, _dp]),
    "vfvm_parameter_derivative": (C.c_int, [_vp, _dp, C.c_double, C.POINTER(ControlT), _dp]),
    "vfvm_measurement_derivatives": (C.c_int, [_vp, _dp, _dp, C.c_int, _dp, _dp]),
    "vfvm_integrate_testfunction": (C.c_int, [_vp, _dp, _dp, C.c_double, _dp, _dp]),
    "vfvm_integrate_testfunction_at": (C.c_int, [_vp, _dp, _dp, C.c_double, _dp, C.c_double, _dp]),
    "vfvm_measurement_derivatives_at": (C.c_int, [_vp, _dp, _dp, C.c_int, C.c_double, _dp, _dp]),
    "vfvm_impedance_sweep": (C.c_int, [_vp, C.c_int64, _dp, _vp, _vp, C.c_double, C.c_int32, _vp, _vp,
                                       C.POINTER(C.c_int32)]),
    "vfvm_pattern": (C.c_int, [_vp, _lp, _lp]),
    "vfvm_initialize_dirichlet": (C.c_int, [_vp, _dp, C.c_double, C.c_double]),
    "vfvm_assemble": (C.c_int, [_vp, _dp, _dp, C.c_double, C.c_double, C.c_double, _dp, _dp]),
    "vfvm_linsolve": (C.c_int, [_vp, _dp, _dp, C.POINTER(ControlT), C.POINTER(LinStatsT)]),
    "vfvm_amg_aggregate": (C.c_int, [C.c_int64, _ip, _ip, _ip, _lp]),
    "vfvm_amg_levels": (C.c_int, [_vp, C.POINTER(C.c_int), _lp]),
    "vfvm_ilu0_factor": (C.c_int, [_vp, _dp]),
    "vfvm_amg_level": (C.c_int, [_vp, C.c_int, _lp, _ip, _ip, _dp, _ip, _ip, _dp, _ip, _dp]),
    "vfvm_newton": (C.c_int, [_vp, _dp, _dp, C.c_double, C.c_double, C.c_double, C.POINTER(ControlT),
                              C.POINTER(HistoryT)]),
    "vfvm_solve_fixed_steps": (C.c_int, [_vp, _dp, C.c_int64, _dp, C.c_int, C.POINTER(ControlT), _dp,
                                         C.POINTER(HistoryT)]),
    "vfvm_default_stepcontrol": (None, [C.c_int, C.POINTER(StepControlT)]),
    "vfvm_solve_transient": (C.c_int, [_vp, _dp, C.c_int64, _dp, C.c_int, C.POINTER(ControlT),
                                       C.POINTER(StepControlT), _dp, C.c_int64, _dp, _dp, C.POINTER(C.c_int64),
                                       C.POINTER(HistoryT)]),
    "vfvm_residual_norm": (C.c_int, [_vp, _dp, _dp, C.c_double, C.c_double, C.c_double, _dp]),
    "vfvm_residual_norm_device_noalloc": (C.c_int, [_vp, _vp, _vp, C.c_double, C.c_double, C.c_double, _vp]),
    "vfvm_autotune": (C.c_int, [_vp, _ip, _ip]),
    "vfvm_set_groups": (C.c_int, [_vp, C.c_int32, C.c_int32]),
    "vfvm_set_sell_index": (C.c_int, [_vp, C.c_int]),
    "vfvm_set_compact_graph": (C.c_int, [_vp, C.c_int]),
    "vfvm_time_kernel": (C.c_int, [_vp, C.c_int, C.c_int, _dp]),
    "vfvm_probe_arm": (C.c_int, [_vp, C.c_int, C.c_int64]),
    "vfvm_probe_read": (C.c_int, [_vp, C.c_int, _dp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
    "vfvm_pattern_csr": (C.c_int, [_vp, _ip, _ip]),
    "vfvm_mpnn_create": (C.c_int, [_vp, C.c_int, C.c_int, C.c_uint64, C.POINTER(_vp)]),
    "vfvm_mpnn_destroy": (C.c_int, [_vp]),
    "vfvm_mpnn_params": (C.c_int, [_vp, _lp, C.POINTER(C.c_float)]),
    "vfvm_mpnn_set_params": (C.c_int, [_vp, C.POINTER(C.c_float)]),
    "vfvm_mpnn_train": (C.c_int, [_vp, C.c_int64, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int64, C.c_int64,
                                  C.c_float, C.POINTER(C.c_float)]),
    "vfvm_mpnn_infer": (C.c_int, [_vp, C.c_int64, C.POINTER(C.c_float), _dp]),
    "vfvm_mpnn_infer_sides": (C.c_int, [_vp, C.c_int64, _dp, _dp, C.c_double, _dp]),
    "vfvm_mpnn_infer_sides_device": (C.c_int, [_vp, C.c_int64, _vp, _vp, C.c_double, _vp]),
    "vfvm_mpnn_loss_grad": (C.c_int, [_vp, C.c_int64, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                      C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "vfvm_batch_create": (C.c_int, [_vp, C.c_int64, C.POINTER(_vp), C.POINTER(_vp)]),
    "vfvm_batch_destroy": (C.c_int, [_vp]),
    "vfvm_batch_set_graphs": (C.c_int, [_vp, C.c_int]),
    "vfvm_batch_set_member_solver": (C.c_int, [_vp, C.c_int]),
    "vfvm_system_set_generic_matrix": (C.c_int, [_vp, C.c_int64, _lp, _lp, _dp]),
    "vfvm_batch_set_sell_index": (C.c_int, [_vp, C.c_int]),
    "vfvm_batch_set_compact_graph": (C.c_int, [_vp, C.c_int]),
    "vfvm_batch_time_kernel": (C.c_int, [_vp, C.c_int, C.c_int, _dp]),
    "vfvm_mpnn_time_layer": (C.c_int, [_vp, C.c_int64, C.c_int, _dp]),
    "vfvm_batch_update_member": (C.c_int, [_vp, C.c_int64, _dp, C.POINTER(BCT)]),
    "vfvm_batch_autotune": (C.c_int, [_vp, _ip, _ip]),
    "vfvm_batch_newton": (C.c_int, [_vp, C.POINTER(_dp), C.POINTER(_dp), C.c_double, C.c_double, C.c_double,
                                    C.POINTER(ControlT), C.POINTER(HistoryT)]),
    "vfvm_batch_newton_device": (C.c_int, [_vp, _vp, _vp, C.c_double, C.c_double, C.c_double, C.POINTER(ControlT),
                                           C.POINTER(HistoryT)]),
    "vfvm_batch_sizes": (C.c_int, [_vp, _lp, _lp]),
    "vfvm_assemble_device": (C.c_int, [_vp, _vp, _vp, C.c_double, C.c_double, C.c_double, _vp, _vp]),
    "vfvm_eval_rhs": (C.c_int, [_vp, _dp, C.c_double, _dp]),
    "vfvm_eval_jacobian": (C.c_int, [_vp, _dp, C.c_double, _dp]),
    "vfvm_mass_matrix": (C.c_int, [_vp, _dp]),
    "vfvm_eval_rhs_device": (C.c_int, [_vp, _vp, C.c_double, _vp]),
    "vfvm_eval_jacobian_device": (C.c_int, [_vp, _vp, C.c_double, _vp]),
}


class VFVMError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class ConvergenceError(VFVMError):
    """Mirrors VoronoiFVM.ConvergenceError (src/vfvm_logging_exceptions.jl:6)."""


class AssemblyError(VFVMError):
    """Mirrors VoronoiFVM.AssemblyError (src/vfvm_logging_exceptions.jl:14)."""


class LinearSolverError(VFVMError):
    """Mirrors VoronoiFVM.LinearSolverError (src/vfvm_logging_exceptions.jl:22)."""


_ERRTYPES = {ERR_CONVERGENCE: ConvergenceError, ERR_ASSEMBLY: AssemblyError, ERR_LINEAR: LinearSolverError}


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError(f"HIP library not built: {LIBPATH} (run __graft_entry__.build())")
        L = C.CDLL(LIBPATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def check(rc, ctx=None):
    if rc != OK:
        msg = lib().vfvm_last_error(ctx).decode(errors="replace")
        raise _ERRTYPES.get(rc, VFVMError)(rc, msg)


def dptr(a: np.ndarray):
    if a.dtype != np.float64 or not (a.flags["C_CONTIGUOUS"] or a.flags["F_CONTIGUOUS"]):
        raise TypeError("expected a contiguous float64 array")
    return a.ctypes.data_as(_dp)


def lptr(a: np.ndarray):
    assert a.dtype == np.int64
    return a.ctypes.data_as(_lp)


def iptr(a: np.ndarray):
    assert a.dtype == np.int32
    return a.ctypes.data_as(_ip)


class Context:
    """One device and one HIP stream (vfvm_ctx)."""

    def __init__(self, device: int = 0):
        h = _vp()
        rc = lib().vfvm_ctx_create(device, C.byref(h))
        if rc != OK:
            raise VFVMError(rc, lib().vfvm_last_error(None).decode())
        self.handle = h

    def __del__(self):
        if getattr(self, "handle", None) and _LIB is not None:
            _LIB.vfvm_ctx_destroy(self.handle)
            self.handle = None

    def set_setup(self, where):
        """Where new systems build their node graph and form factors: True (the GPU),
        False (the host) or "auto" (default: the GPU for grids of at least 2^18 cell-node
        incidences); all give the same system bit for bit."""
        mode = 2 if where == "auto" else int(bool(where))
        check(lib().vfvm_ctx_set_setup(self.handle, mode), self.handle)

    def probe(self, *kernels: int, max_launches: int = 50000):
        """Context manager timing every launch of the given kernel classes (KERNEL_*)
        inside the calls it wraps, with HIP events on the launching stream.  Yields a
        dict {kernel: {total_ms, launches, avg_ms, work}} filled on exit; work is
        (rows, nonzeros, 0) for SpMV and (nodes, adjacency entries, Jacobian slots)
        for assembly, summed over the timed launches."""
        return _Probe(self, kernels, max_launches)


(KERNEL_ASSEMBLE, KERNEL_RESIDUAL, KERNEL_SPMV64, KERNEL_SPMV32, KERNEL_HALFSTEP64, KERNEL_HALFSTEP32, KERNEL_AMG64,
 KERNEL_AMG32, KERNEL_MSOLVE64, KERNEL_MSOLVE32) = range(10)


class _Probe:
    def __init__(self, ctx, kernels, max_launches):
        self.ctx, self.kernels, self.max_launches = ctx, kernels, max_launches
        self.result = {}

    def __enter__(self):
        for k in self.kernels:
            check(lib().vfvm_probe_arm(self.ctx.handle, k, self.max_launches), self.ctx.handle)
        return self.result

    def __exit__(self, *exc):
        for k in self.kernels:
            t = C.c_double()
            n = C.c_int64()
            w = (C.c_int64 * 3)()
            check(lib().vfvm_probe_read(self.ctx.handle, k, C.byref(t), C.byref(n), w), self.ctx.handle)
            self.result[k] = dict(total_ms=t.value, launches=n.value, avg_ms=t.value / max(n.value, 1),
                                  work=tuple(w))
        return False


_DEFAULT_CTX = None


def default_context() -> Context:
    global _DEFAULT_CTX
    if _DEFAULT_CTX is None:
        _DEFAULT_CTX = Context(0)
    return _DEFAULT_CTX


class DeviceArray:
    """A float64 array in the device memory of a context (vfvm_device_alloc), for the
    *_device entry points: inputs stay resident between calls and results are read
    back only when wanted."""

    def __init__(self, ctx, n):
        self.ctx, self.n = ctx, int(n)
        p = C.c_void_p()
        check(lib().vfvm_device_alloc(ctx.handle, 8 * self.n, C.byref(p)), ctx.handle)
        self.ptr = p

    @classmethod
    def from_numpy(cls, ctx, a):
        a = np.ascontiguousarray(a, dtype=np.float64).ravel()
        d = cls(ctx, a.size)
        d.upload(a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64).ravel()
        if a.size != self.n:
            raise ValueError("size mismatch")
        check(lib().vfvm_device_copy(self.ctx.handle, self.ptr, a.ctypes.data_as(_vp), 8 * self.n, 0),
              self.ctx.handle)

    def numpy(self):
        out = np.empty(self.n)
        check(lib().vfvm_device_copy(self.ctx.handle, out.ctypes.data_as(_vp), self.ptr, 8 * self.n, 1),
              self.ctx.handle)
        return out

    def free(self):
        if getattr(self, "ptr", None) is not None and self.ptr.value and _LIB is not None:
            _LIB.vfvm_device_free(self.ctx.handle, self.ptr)
        self.ptr = None

    def __del__(self):
        self.free()
