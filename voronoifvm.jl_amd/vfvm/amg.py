"""The smoothed aggregation AMG preconditioner (SolverControl(precon="amg")): the
reference's KrylovJL_BICGSTAB(precs = SmoothedAggregationPreconBuilder())
(examples/DevEx003_Solvers.jl:149-156).  These functions read its hierarchy back for
inspection: the aggregation of a pattern (host code of the library) and, after a solve
or linsolve with precon="amg", each level's operator A, prolongator P, aggregates and
Jacobi weight omega (AlgebraicMultigrid.jl's MultiLevel levels)."""
import ctypes as C

import numpy as np

from . import _capi


def aggregate(rowptr, colind):
    """Standard aggregation of a 0-based CSR pattern (its diagonal ignored): (agg, count)."""
    rp = np.ascontiguousarray(rowptr, dtype=np.int32)
    ci = np.ascontiguousarray(colind, dtype=np.int32)
    n = rp.size - 1
    agg = np.zeros(max(n, 1), np.int32)
    na = C.c_int64(0)
    ip = C.POINTER(C.c_int32)
    _capi.check(_capi.lib().vfvm_amg_aggregate(n, rp.ctypes.data_as(ip), ci.ctypes.data_as(ip), agg.ctypes.data_as(ip),
                                               C.byref(na)))
    return agg[:n], int(na.value)


def hierarchy(sys):
    """Levels of the system's current AMG hierarchy: a list of dicts with A (scipy CSR; level
    0 is the row-scaled Jacobian diag(1/diag J) J), P (CSR rows x coarse rows, None on the
    coarsest level), agg (None on the coarsest level) and omega."""
    import scipy.sparse as sp
    L = _capi.lib()
    h = sys.handle()
    nl = C.c_int(0)
    rows = np.zeros(16, np.int64)
    _capi.check(L.vfvm_amg_levels(h, C.byref(nl), _capi.lptr(rows)), sys.ctx.handle)
    ip = C.POINTER(C.c_int32)
    out = []
    for lvl in range(nl.value):
        sz = np.zeros(4, np.int64)
        null = C.POINTER(C.c_int32)()
        _capi.check(L.vfvm_amg_level(h, lvl, _capi.lptr(sz), null, null, C.POINTER(C.c_double)(), null, null,
                                     C.POINTER(C.c_double)(), null, C.POINTER(C.c_double)()), sys.ctx.handle)
        n, nnz, nc, nnzp = (int(v) for v in sz)
        arp = np.zeros(n + 1, np.int32)
        aci = np.zeros(max(nnz, 1), np.int32)
        av = np.zeros(max(nnz, 1))
        prp = np.zeros(n + 1, np.int32)
        pci = np.zeros(max(nnzp, 1), np.int32)
        pv = np.zeros(max(nnzp, 1))
        agg = np.zeros(max(n, 1), np.int32)
        om = C.c_double(0.0)
        _capi.check(L.vfvm_amg_level(h, lvl, _capi.lptr(sz), arp.ctypes.data_as(ip), aci.ctypes.data_as(ip),
                                     _capi.dptr(av), prp.ctypes.data_as(ip), pci.ctypes.data_as(ip), _capi.dptr(pv),
                                     agg.ctypes.data_as(ip), C.byref(om)), sys.ctx.handle)
        A = sp.csr_matrix((av[:nnz], aci[:nnz], arp), shape=(n, n))
        P = sp.csr_matrix((pv[:nnzp], pci[:nnzp], prp), shape=(n, nc)) if nc else None
        out.append(dict(A=A, P=P, agg=agg[:n] if nc else None, omega=om.value))
    return out
