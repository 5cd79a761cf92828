"""vfvm_residual_norm_device_noalloc (the entry point that allocates nothing) against
vfvm_residual_norm, through the C ABI: two calls in a row on one GPU, the second on a
smaller system, each bit for bit equal to the allocating entry point's norm; and the
refusal of a system beyond the static array."""
import ctypes as C

import numpy as np
import pytest

import cases
import vfvm
from vfvm import _capi

pytestmark = pytest.mark.gpu

MAX_DOF = 1 << 22


def _norms(sys, U, Uold, tstep, embed):
    """(allocating entry point's norm, no-alloc entry point's norm) at U, Uold"""
    L = _capi.lib()
    h = sys.handle()
    Ud = np.ascontiguousarray(vfvm.dofs(U), dtype=float)
    Uod = np.ascontiguousarray(vfvm.dofs(Uold), dtype=float)
    ref = C.c_double()
    _capi.check(L.vfvm_residual_norm(h, _capi.dptr(Ud), _capi.dptr(Uod), 0.0, tstep, embed, C.byref(ref)),
                sys.ctx.handle)
    dU = vfvm.DeviceArray.from_numpy(sys.ctx, Ud)
    dUo = vfvm.DeviceArray.from_numpy(sys.ctx, Uod)
    dn = vfvm.DeviceArray.from_numpy(sys.ctx, np.full(1, -7.0))
    _capi.check(L.vfvm_residual_norm_device_noalloc(h, dU.ptr, dUo.ptr, 0.0, tstep, embed, dn.ptr), sys.ctx.handle)
    got = dn.numpy()[0]  # the copy synchronises the handle's stream
    return ref.value, got


def test_noalloc_residual_norm_equals_allocating_entry_point():
    print("case noalloc: 65^2 drift-diffusion, then 33^2 Poisson, then 9^3 transient", flush=True)
    rng = np.random.default_rng(3)
    sys, _ = cases.drift_diffusion_2d(65, lambda2=1e-2, oracle=False)
    U = cases.dd_initial(sys.grid)
    U[0] += rng.uniform(-1, 1, sys.grid.num_nodes)
    ref, got = _norms(sys, U, U, np.inf, 0.4)
    print("dd 65^2", ref, got, flush=True)
    assert ref > 0 and got == ref
    sys2, _ = cases.nonlinear_poisson_2d(33, oracle=False)  # second call: a smaller input
    U2 = rng.uniform(-1, 1, (1, sys2.grid.num_nodes))
    ref2, got2 = _norms(sys2, U2, U2, np.inf, 0.0)
    print("poisson 33^2", ref2, got2, flush=True)
    assert ref2 > 0 and got2 == ref2
    sys3, _ = cases.nonlinear_diffusion_3d(9)
    u0 = cases.initial_sine_3d(sys3.grid).reshape(1, -1)
    ref3, got3 = _norms(sys3, 1.2 * u0, u0, 1e-2, 0.0)
    print("nldiff 9^3 transient", ref3, got3, flush=True)
    assert ref3 > 0 and got3 == ref3


def test_noalloc_refuses_systems_beyond_its_array():
    print("case noalloc refusal: 3 species on 1183^2 nodes", flush=True)
    n = 1183  # 3 * 1183^2 = 4,198,467 dofs > 2^22
    assert 3 * n * n > MAX_DOF
    sys, _ = cases.drift_diffusion_2d(n, lambda2=1e-2, oracle=False)
    L = _capi.lib()
    dn = vfvm.DeviceArray(sys.ctx, 1)
    dU = vfvm.DeviceArray(sys.ctx, 3 * n * n)  # full size, though the size check comes before any launch
    rc = L.vfvm_residual_norm_device_noalloc(sys.handle(), dU.ptr, None, 0.0, np.inf, 0.0, dn.ptr)
    assert rc == 4
    assert "exceed" in _capi.lib().vfvm_last_error(sys.ctx.handle).decode()
