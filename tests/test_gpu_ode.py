"""ODE/DAE form of the discrete system (src/vfvm_diffeq_interface.jl): eval_rhs!,
eval_jacobian! and mass_matrix through the C ABI, against the oracle's assembly
at tstep = Inf and against the implicit-Euler identity
    F_dt(u; uold) = M (u - uold) / dt - rhs(u),   dF_dt/du = M / dt - J(u)
that ties the three to the (parity-tested) transient assembly."""
import ctypes as C
import math

import numpy as np
import pytest

import cases
import vfvm
from vfvm import _capi

pytestmark = pytest.mark.gpu

RTOL = 1e-12


def _print(*a):
    print(*a, flush=True)


def _cases():
    return [("nldiff3d", lambda: cases.nonlinear_diffusion_3d(7), (0.2, 1.0)),
            ("dd2d", lambda: cases.drift_diffusion_2d(9, lambda2=1e-2), (0.2, 1.0)),
            ("rz2d", lambda: cases.nonlinear_poisson_sg("rz2d.sg"), (-1.0, 1.0))]


def _close(a, b, pen, rtol=RTOL):
    for mask in (pen, ~pen):
        if mask.any():
            sc = max(np.abs(b[mask]).max(), 1e-300)
            assert np.abs(a[mask] - b[mask]).max() <= rtol * sc


@pytest.mark.parametrize("case", [0, 1, 2])
def test_rhs_and_jacobian_match_oracle(case):
    name, build, (lo, hi) = _cases()[case]
    _print("case ode rhs/jacobian", name)
    sys, osys = build()
    ns, n = sys.num_species, sys.grid.num_nodes
    u = np.random.default_rng(case).uniform(lo, hi, ns * n)
    Fo = osys.assemble(u, u, tstep=math.inf, embed=0.0)
    Ao = osys.matrix_scipy().tocsc()
    Ao.sort_indices()
    pen = np.abs(Ao.diagonal()) > 1e20
    du = vfvm.eval_rhs(sys, u)
    _close(du, -Fo, pen)
    J = vfvm.eval_jacobian(sys, u)
    assert np.array_equal(J.indptr, Ao.indptr) and np.array_equal(J.indices, Ao.indices)
    rowpen = pen[Ao.indices]
    _close(J.data, -Ao.data, rowpen)
    # device forms: the same numbers, bit for bit
    ctx = sys.ctx
    ud = vfvm.DeviceArray.from_numpy(ctx, u)
    dud = vfvm.DeviceArray(ctx, u.size)
    sys.eval_rhs_device(ud, dud)
    assert np.array_equal(dud.numpy(), du)
    ndof, nnz, _ = sys.sizes()
    jd = vfvm.DeviceArray(ctx, nnz)
    sys.eval_jacobian_device(ud, jd)
    rp, ci = sys.pattern_csr()
    import scipy.sparse as sp
    Jcsr = sp.csr_matrix((jd.numpy(), ci, rp), shape=(ndof, ndof)).tocsc()
    Jcsr.sort_indices()
    assert np.array_equal(Jcsr.data, J.data)


@pytest.mark.parametrize("case", [0, 1, 2])
def test_mass_matrix_and_implicit_euler_identity(case):
    name, build, (lo, hi) = _cases()[case]
    _print("case ode mass matrix", name)
    sys, _ = build()
    ns, n = sys.num_species, sys.grid.num_nodes
    rng = np.random.default_rng(10 + case)
    u = rng.uniform(lo, hi, ns * n)
    uold = rng.uniform(lo, hi, ns * n)
    M = vfvm.mass_matrix(sys).diagonal()
    has_storage = name == "nldiff3d"
    want = np.repeat(sys.nodevolumes(), ns) if has_storage else np.zeros(ns * n)
    assert np.array_equal(M, want)
    dt = 0.05
    ndof, nnz, _ = sys.sizes()
    F = np.zeros(ndof)
    nz = np.zeros(nnz)
    L = _capi.lib()
    _capi.check(L.vfvm_assemble(sys.handle(), _capi.dptr(u), _capi.dptr(uold), 0.0, dt, 0.0, _capi.dptr(F),
                                _capi.dptr(nz)))
    rhs = vfvm.eval_rhs(sys, u)
    J = vfvm.eval_jacobian(sys, u)
    A = J.copy()
    A.data[:] = nz
    pen = np.abs(A.diagonal()) > 1e20
    _close(F, M * (u - uold) / dt - rhs, pen)
    import scipy.sparse as sp
    W = (sp.diags(M / dt) - J).tocsc()
    W.sort_indices()
    A.sort_indices()
    # compare on A's pattern (W has the same pattern: M is diagonal, the diagonal is stored)
    Wv = np.asarray(W.todense())[A.indices, np.repeat(np.arange(ndof), np.diff(A.indptr))]
    _close(A.data, Wv, pen[A.indices])
