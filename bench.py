"""Benchmark of the MI355X assemble-and-solve path.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME] [--no-cpu]

Prints ONE JSON line (metric, value, unit, ... roofline, cpu_baseline).  The
value counts converged solves: Newton to ||F||_inf <= 1e-10 over non-penalty
rows.  Inputs are resident on the device before the timed region; only the
solve calls are timed.  Workloads:

  cfg4 (default)  1,024 held-out 129x129 nonlinear-Poisson inputs: learned initial
                  guess + batches of 512 + FP32 BiCGStab with a smoothed-aggregation
                  AMG preconditioner, against the plain FP64 GPU Newton from zero on
                  the same inputs; a step is one batch of 512
  cfg0            101^2 nonlinear Poisson (the reference's CPU-runnable case); a step
                  is one solve
  cfg1            129^3 nonlinear diffusion, implicit Euler; a step is one time step
  cfg2            512^2 bipolar drift-diffusion bias sweep; a step is one sweep
  cfg3            256 seeded nonlinear-Poisson instances of uneven size; a step is
                  one batched solve of all 256
  copy            1 GiB device-to-device copy: the measured HBM bandwidth reference

The roofline entries come from the library's kernel probe: right after the timed
solves, the same solves run once more with every launch of the kernel bracketed
by HIP events on the library's stream (batched Krylov blocks then launch kernel
by kernel instead of as HIP graphs), and the work those launches did is counted
alongside.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
# oracle/ and tests/ serve the cpu_baseline legs only (the checker and its problem twins)
for _p in (os.path.join(ROOT, "voronoifvm.jl_amd"), os.path.join(ROOT, "oracle"), os.path.join(ROOT, "tests")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import ctypes as C  # noqa: E402

import vfvm  # noqa: E402
from vfvm import _capi  # noqa: E402

METRIC = "converged solves/s at ||F||<=1e-10: surrogate+batched+FP32 vs plain GPU Newton"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def spmv_bytes(rows, nnz, sv):
    """Algorithmic bytes of y = A x: values + column indices, row pointers, x and y once."""
    return nnz * (sv + 4) + (rows + 1) * 4 + 2 * rows * sv


def assembly_bytes(nspec, nodes, adj, slots, transient=True):
    """Algorithmic bytes of one Jacobian assembly: U, F (and Uold), the node factor,
    per edge its factor and two int32 endpoints, and the Jacobian values."""
    return 8 * nspec * nodes * (3 if transient else 2) + 8 * nodes + (adj // 2) * (8 + 8) + 8 * slots


def probe_bytes(kind, probe, sv=4, nspec=1, transient=True):
    """Algorithmic bytes per launch from the work the timed launches did."""
    w, n = probe["work"], max(probe["launches"], 1)
    if kind == "spmv":
        return spmv_bytes(w[0] / n, w[1] / n, sv)
    if kind == "amg_sweep":
        # out = x + omega (b - A x) on the AMG's own SELL copy of the operator: FP16 values,
        # column indices shared by the members of one pattern (read from the L2, not
        # counted), b and x read and out written in the Krylov precision
        return 2 * w[1] / n + 3 * sv * w[0] / n
    if kind == "msolve":
        # one launch = every member's BiCGStab solve; per iteration of a member: two
        # operator applications (values sv + 2-byte column offset per nonzero; x gathered
        # once, y written, one dot vector read: 3 sv per row), two AMG V-cycles (the
        # cycle's matrix entry reads at 2 bytes (FP16) or 8, eight finest-level vector
        # passes), and the vector updates (p: 4 vectors, s: 3, x and r: 8)
        rows, nnz, cyc = w[0] / n, w[1] / n, w[2] / n
        vb = 2 if sv == 4 else 8
        return 2 * nnz * (sv + 2) + 2 * 3 * sv * rows + 2 * vb * cyc + 2 * 8 * sv * rows + 15 * sv * rows
    return assembly_bytes(nspec, w[0] / n, w[1] / n, w[2] / n, transient)


def roofline_hbm(algorithmic_bytes, probe, traffic, name):
    """Roofline entry of an HBM-bound kernel from its live per-launch durations."""
    t = probe["avg_ms"]
    gbs = algorithmic_bytes / (t * 1e-3) / 1e9
    log(f"{name}: {probe['launches']} launches, {t * 1e3:.1f} us/launch, {gbs:.0f} GB/s "
        f"({100 * gbs / HBM_PEAK_GBS:.1f}% of 8 TB/s)")
    return dict(bound="hbm", achieved=round(gbs, 1), peak=HBM_PEAK_GBS, unit="GB/s", frac=round(gbs / HBM_PEAK_GBS, 4),
                traffic=traffic, algorithmic_bytes=int(algorithmic_bytes), avg_ms=round(t, 5),
                launches=probe["launches"])


def read_traffic(name):
    """HBM bytes per launch measured with rocprofv3 --pmc (profiles/pmc_traffic.json), or None."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    try:
        with open(path) as f:
            return json.load(f).get(name, {}).get("bytes_per_launch")
    except (OSError, ValueError):
        return None


# linear tolerance of every form: SolverControl's default reltol_linear
# (src/vfvm_solvercontrol.jl:111-116).  FP32 refinement rounds reduce the residual by
# INNER_RELTOL each with Jacobi or ILU(0) (an FP32 BiCGStab asked for 1e-4 at once takes
# 1.6-1.8x the iterations of the FP64 one on cfg1 and cfg3); the AMG-preconditioned cfg4
# solve aims at reltol_linear in one round
RELTOL_LINEAR = 1e-4
INNER_RELTOL = 1e-2

# --------------------------------------------------------------------------- cfg0
def run_cfg0(args):
    """The reference's CPU-runnable case: 101 x 101 nonlinear Poisson (eps = 1e-2, f = 1,
    Dirichlet 0 on west/south and 1 on east/north, zero initial value), plain FP64 device
    Newton with BiCGStab + ILU(0); a step is one solve."""
    from vfvm import problems
    s = problems.nonlinear_poisson_2d(101)
    s.handle()
    s.autotune()
    ctrl = vfvm.SolverControl(precon="ilu0", reltol_linear=1e-8, abstol_linear=0.0, maxiters_linear=1000)
    K, W = args.steps, args.warmup
    for _ in range(max(W, 1)):
        vfvm.solve(s, inival=0.0, control=ctrl)
    t1 = time.perf_counter()
    for _ in range(K):
        vfvm.solve(s, inival=0.0, control=ctrl)
        if not s.history.residual_norm <= 1e-10:
            raise RuntimeError(f"cfg0: ||F||_inf = {s.history.residual_norm:.3e}")
    el = time.perf_counter() - t1
    log(f"cfg0: {K} solves in {el:.3f}s -> {K / el:.1f} solves/s, newton {s.history.niter}, krylov {s.history.nlin}")
    return dict(value=K / el, ms_per_step=1e3 * el / K, roofline=None,
                config=dict(workload="cfg0: 2D nonlinear Poisson 101x101 (10,201 unknowns), eps=1e-2, f=1, FP64 "
                                     "device Newton with BiCGStab + ILU(0) (the reference solves it with UMFPACK)",
                            unknowns=s.grid.num_nodes),
                dtype="f64")


def cpu_cfg0(args):
    """The oracle on the same problem: Newton with banded direct solves (UMFPACK's role)."""
    import cases
    s, o = cases.nonlinear_poisson_2d(101)
    n = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < 10.0:
        o.solve_step(np.zeros(s.grid.num_nodes))
        n += 1
    el = time.perf_counter() - t0
    return dict(value=n / el, unit="solves/s", cores=1, kind="port",
                sample=f"{n} solves of the cfg0 problem with the oracle's banded direct solver, {el:.2f} s")


# --------------------------------------------------------------------------- cfg1
def build_cfg1(n):
    """3D nonlinear diffusion D(u) = 1 + u^2 on the unit cube, storage u, u = 0 on the boundary."""
    X = vfvm.linspace(0.0, 1.0, n)
    g = vfvm.simplexgrid(X, X, X)
    bcs = [vfvm.Dirichlet(species=1, region=r, value=0.0) for r in range(1, 7)]
    ph = vfvm.Physics(flux=vfvm.NonlinearDiffusionFlux(1.0, 1.0), storage=vfvm.IdentityStorage(), bcondition=bcs)
    s = vfvm.System(g, ph, species=[1])
    x, y, z = g.coord
    u0 = 0.5 * np.sin(math.pi * x) * np.sin(math.pi * y) * np.sin(math.pi * z)
    return s, u0.reshape(1, -1)


def cfg1_control(precision):
    """BiCGStab + ILU(0), fixed implicit-Euler steps of 1e-2 (fixed_timesteps!)."""
    c = vfvm.SolverControl(krylov="bicgstab", precon="ilu0", precision=precision, reltol_linear=RELTOL_LINEAR,
                           abstol_linear=0.0, maxiters_linear=1000, inner_reltol=INNER_RELTOL, max_refinements=10)
    return vfvm.fixed_timesteps(c, 1e-2)


def run_cfg1(args):
    n = 129
    t0 = time.perf_counter()
    s, u0 = build_cfg1(n)
    s.handle()
    ndof, nnz, nedges = s.sizes()
    create_s = time.perf_counter() - t0
    log(f"cfg1: {n}^3 grid, ndof={ndof} nnz={nnz} edges={nedges}, setup {create_s:.1f}s")
    # the same system with the node graph and form factors built on the host
    s.ctx.set_setup(False)
    try:
        t1 = time.perf_counter()
        s_host, _ = build_cfg1(n)
        s_host.handle()
        create_host_s = time.perf_counter() - t1
    finally:
        s.ctx.set_setup("auto")
    del s_host
    log(f"system creation: {create_s:.2f} s with the device setup, {create_host_s:.2f} s with the host setup")
    # update_grid on the device: every form factor, node volume and boundary factor
    # recomputed for the (same) coordinates; then the compact factor table rebuilt
    tg = []
    for _ in range(3):
        t1 = time.perf_counter()
        s.update_grid(s.grid)
        tg.append(time.perf_counter() - t1)
    s.set_compact_graph(True)
    log(f"device geometry update: {1e3 * min(tg):.1f} ms (system creation {create_s:.2f} s)")
    ag, sg = s.autotune()
    log(f"autotuned groups: assembly {ag}, spmv {sg}")
    dt = 1e-2
    K, W = args.steps, args.warmup
    res = {}
    probes = {}
    for label, prec, kernels in (("plain_fp64", "fp64", (_capi.KERNEL_SPMV64,)),
                                 ("fp32_refine", "fp32", (_capi.KERNEL_SPMV32, _capi.KERNEL_ASSEMBLE))):
        ctrl = cfg1_control(prec)
        if W > 0:
            warm = np.cumsum([0.0] + [dt] * W)
            vfvm.solve(s, inival=np.asfortranarray(u0), control=ctrl, times=warm, store_all=False)
        times = np.cumsum([0.0] + [dt] * K)
        t1 = time.perf_counter()
        U = vfvm.solve(s, inival=np.asfortranarray(u0), control=ctrl, times=times, store_all=False).u[-1]
        el = time.perf_counter() - t1
        # probe pass: the same solve again with the kernels' launches bracketed by HIP events
        with s.ctx.probe(*kernels) as pr:
            vfvm.solve(s, inival=np.asfortranarray(u0), control=ctrl, times=times, store_all=False)
        probes.update(pr)
        # converged = every step met ||F|| <= 1e-10: check the final state's residual
        Ud = np.ascontiguousarray(vfvm.dofs(U))
        if not s.history.residual_norm <= 1e-10:
            raise RuntimeError(f"{label}: a step ended with ||F||_inf = {s.history.residual_norm:.3e} > 1e-10")
        res[label] = dict(seconds=el, solves_per_s=K / el, newton=s.history.niter, krylov=s.history.nlin,
                          max_residual=s.history.residual_norm, U=Ud)
        log(f"{label}: {K} steps in {el:.3f}s -> {K / el:.3f} solves/s, newton iterations {s.history.niter}, "
            f"krylov iterations {s.history.nlin}")
    # residual of the last step at the returned solution (tstep = dt, Uold = previous step)
    err32 = float(np.abs(res["fp32_refine"]["U"] - res["plain_fp64"]["U"]).max())
    # roofline of the hot kernels at the final state
    n_, E = s.grid.num_nodes, nedges
    b_asm = probe_bytes("assemble", probes[_capi.KERNEL_ASSEMBLE])
    b_spmv64 = probe_bytes("spmv", probes[_capi.KERNEL_SPMV64], sv=8)
    b_spmv32 = probe_bytes("spmv", probes[_capi.KERNEL_SPMV32], sv=4)
    rl = {}
    for name, b, k in (("assemble", b_asm, _capi.KERNEL_ASSEMBLE), ("spmv64", b_spmv64, _capi.KERNEL_SPMV64),
                       ("spmv32", b_spmv32, _capi.KERNEL_SPMV32)):
        rl[name] = roofline_hbm(b, probes[k], read_traffic("cfg1_" + name), name)
    # ODE/DAE form on device memory (eval_rhs!, eval_jacobian!): the assembly at
    # tstep = Inf, then the negation; assembly kernels timed by the probe
    ud = vfvm.DeviceArray.from_numpy(s.ctx, res["fp32_refine"]["U"])
    dud, jd = vfvm.DeviceArray(s.ctx, ndof), vfvm.DeviceArray(s.ctx, nnz)
    s.eval_rhs_device(ud, dud)
    s.eval_jacobian_device(ud, jd)
    t1 = time.perf_counter()
    for _ in range(10):
        s.eval_rhs_device(ud, dud)
        s.eval_jacobian_device(ud, jd)
    ode_pair_ms = 1e3 * (time.perf_counter() - t1) / 10
    with s.ctx.probe(_capi.KERNEL_RESIDUAL, _capi.KERNEL_ASSEMBLE) as pode:
        for _ in range(10):
            s.eval_rhs_device(ud, dud)
            s.eval_jacobian_device(ud, jd)
    for name, k, kind in (("ode_rhs", _capi.KERNEL_RESIDUAL, "residual"), ("ode_jacobian", _capi.KERNEL_ASSEMBLE,
                                                                           "assemble")):
        w, nl = pode[k]["work"], max(pode[k]["launches"], 1)
        b = assembly_bytes(1, w[0] / nl, w[1] / nl, w[2] / nl, transient=False)
        rl[name] = roofline_hbm(b, pode[k], None, name)
    for x in (ud, dud, jd):
        x.free()
    log(f"ODE form: eval_rhs + eval_jacobian on device memory {ode_pair_ms:.2f} ms per pair (wall clock)")
    out = dict(
        value=res["fp32_refine"]["solves_per_s"], ms_per_step=1e3 * res["fp32_refine"]["seconds"] / K,
        plain_gpu_newton_solves_per_s=res["plain_fp64"]["solves_per_s"],
        speedup_vs_plain=res["fp32_refine"]["solves_per_s"] / res["plain_fp64"]["solves_per_s"],
        max_abs_diff_fp32_vs_fp64=err32, roofline=rl["spmv32"], roofline_assemble=rl["assemble"],
        roofline_spmv64=rl["spmv64"], geometry_update_s=min(tg), system_create_s=create_s, system_create_host_setup_s=create_host_s,
        roofline_ode_rhs=rl["ode_rhs"], roofline_ode_jacobian=rl["ode_jacobian"], ode_rhs_jacobian_pair_ms=ode_pair_ms,
        config=dict(workload=f"cfg1: 3D nonlinear diffusion D(u)=1+u^2, tetrahedral {n}^3 grid, implicit Euler "
                             f"dt=1e-2, BiCGStab+ILU(0) FP32 inside FP64 Newton (plain: FP64 Krylov)",
                    nodes=n_, unknowns=ndof, nnz=nnz, steps_per_solve="one implicit-Euler step"),
        dtype="f64+f32")
    return out


def cpu_cfg1(args):
    """The oracle (CPU restatement) on a bounded sample: one implicit-Euler step of the
    same problem on an 81^3 grid (531,441 nodes), single thread."""
    import cases
    import oracle as O
    n = 81
    sys_, osys = cases.nonlinear_diffusion_3d(n)
    u0 = cases.initial_sine_3d(sys_.grid)
    c = O.control(method_linear="bicgstab_ilu0", reltol_linear=RELTOL_LINEAR, abstol_linear=0.0,
                  maxiters_linear=1000)
    t0 = time.perf_counter()
    osys.solve_step(u0, c, time=1e-2, tstep=1e-2)
    el = time.perf_counter() - t0
    return dict(value=1.0 / el, unit="solves/s", cores=1, kind="port",
                sample=f"1 implicit-Euler Newton solve of the cfg1 problem on {n}^3 nodes "
                       f"({sys_.grid.num_nodes} nodes, 1/4 of 129^3) with the oracle's BiCGStab+ILU(0) FP64, "
                       f"{el:.2f} s")


# --------------------------------------------------------------------------- cfg2
CFG2_BIAS = np.linspace(0.0, 1.0, 21)  # 0 .. 1 V in 20 continuation steps


def cfg2_control(precision):
    """Block-Jacobi BiCGStab; fixed embedding steps of 0.05 V (Δp = Δp_min = Δp_max)."""
    return vfvm.SolverControl(krylov="bicgstab", precon="blockjacobi", precision=precision,
                              reltol_linear=RELTOL_LINEAR, abstol_linear=0.0, maxiters_linear=6000,
                              inner_reltol=INNER_RELTOL, max_refinements=12, Δp=0.05, Δp_min=0.05, Δp_max=0.05)


def run_cfg2(args):
    """Bipolar drift-diffusion on 512 x 512 nodes: a bias sweep 0..1 V in 20 continuation
    steps (21 Newton solves, the first at zero bias), BiCGStab + 3x3 block Jacobi (GMRES(30)
    with block Jacobi stalls at this size: its restarted Krylov space is too small)."""
    from vfvm import problems
    n = 512
    t0 = time.perf_counter()
    s = problems.drift_diffusion_2d(n, lambda2=1e-4)
    s.handle()
    ndof, nnz, nedges = s.sizes()
    u0 = problems.dd_initial(s.grid)
    ag, sg = s.autotune()
    log(f"cfg2: {n}^2 grid, ndof={ndof} nnz={nnz} edges={nedges}, setup {time.perf_counter() - t0:.1f}s, "
        f"autotuned groups {ag}/{sg}")
    K, W = args.steps, args.warmup
    nsolve = CFG2_BIAS.size
    res = {}
    probes = {}
    for label, prec, kernels in (("plain_fp64", "fp64", (_capi.KERNEL_SPMV64,)),
                                 ("fp32_refine", "fp32", (_capi.KERNEL_SPMV32, _capi.KERNEL_ASSEMBLE))):
        ctrl = cfg2_control(prec)
        for _ in range(W):
            vfvm.solve(s, inival=u0, control=ctrl, embed=CFG2_BIAS, store_all=False)
        t1 = time.perf_counter()
        for _ in range(K):
            U = vfvm.solve(s, inival=u0, control=ctrl, embed=CFG2_BIAS, store_all=False).u[-1]
        el = time.perf_counter() - t1
        with s.ctx.probe(*kernels) as pr:  # probe pass: one more sweep, launches bracketed by events
            vfvm.solve(s, inival=u0, control=ctrl, embed=CFG2_BIAS, store_all=False)
        probes.update(pr)
        if not s.history.residual_norm <= 1e-10:
            raise RuntimeError(f"{label}: a bias step ended with ||F||_inf = {s.history.residual_norm:.3e}")
        res[label] = dict(seconds=el, solves_per_s=nsolve * K / el, U=np.ascontiguousarray(vfvm.dofs(U)),
                          newton=s.history.niter, krylov=s.history.nlin, fp64_fallbacks=s.history.fp64_fallbacks)
        log(f"{label}: {K} sweeps x {nsolve} solves in {el:.3f}s -> {nsolve * K / el:.2f} solves/s, newton "
            f"{s.history.niter}, krylov {s.history.nlin}, linear solves done in FP64 after an FP32 failure "
            f"{s.history.fp64_fallbacks} per sweep")
    diff = float(np.abs(res["fp32_refine"]["U"] - res["plain_fp64"]["U"]).max())
    b_spmv32 = probe_bytes("spmv", probes[_capi.KERNEL_SPMV32], sv=4)
    b_asm = probe_bytes("assemble", probes[_capi.KERNEL_ASSEMBLE], nspec=3, transient=False)
    rl = {}
    for name, b, k in (("spmv32", b_spmv32, _capi.KERNEL_SPMV32), ("assemble", b_asm, _capi.KERNEL_ASSEMBLE)):
        rl[name] = roofline_hbm(b, probes[k], read_traffic("cfg2_" + name), name)
    return dict(value=res["fp32_refine"]["solves_per_s"], ms_per_step=1e3 * res["fp32_refine"]["seconds"] / K,
                plain_gpu_newton_solves_per_s=res["plain_fp64"]["solves_per_s"],
                speedup_vs_plain=res["fp32_refine"]["solves_per_s"] / res["plain_fp64"]["solves_per_s"],
                max_abs_diff_fp32_vs_fp64=diff, roofline=rl["spmv32"], roofline_assemble=rl["assemble"],
                newton_iterations_per_sweep=res["fp32_refine"]["newton"],
                fp64_fallbacks_per_sweep=res["fp32_refine"]["fp64_fallbacks"],
                config=dict(workload=f"cfg2: bipolar drift-diffusion (psi, n, p), {n}x{n} nodes, lambda=1e-2, "
                                     f"SG fluxes, bias 0..1 V in 20 continuation steps, BiCGStab + 3x3 "
                                     f"block Jacobi, FP32 inside FP64 Newton (plain: FP64 Krylov); a step is "
                                     f"one sweep of {nsolve} solves", unknowns=ndof, nnz=nnz),
                dtype="f64+f32")


def cpu_cfg2(args):
    """The oracle on a bounded sample: the same sweep on a 65 x 65 grid, banded direct solves."""
    import cases
    import oracle as O
    n = 65
    s, o = cases.drift_diffusion_2d(n, lambda2=1e-4)
    u0 = cases.dd_initial(s.grid)
    U = vfvm.dofs(u0).copy()
    t0 = time.perf_counter()
    for lam in CFG2_BIAS:  # fixed continuation steps: one Newton solve per bias value
        U, _ = o.solve_step(U, O.control(), embed=float(lam))
    el = time.perf_counter() - t0
    return dict(value=CFG2_BIAS.size / el, unit="solves/s", cores=1, kind="port",
                sample=f"the cfg2 bias sweep ({CFG2_BIAS.size} solves) on {n}x{n} nodes with the oracle's "
                       f"banded direct solves, {el:.2f} s")


# --------------------------------------------------------------------------- cfg3
def cfg3_members(count=256):
    from vfvm import problems
    return [problems.batch_member(seed) for seed in range(count)]


def build_poisson(n, eps, vals, f):
    X = vfvm.linspace(0.0, 1.0, n)
    g = vfvm.simplexgrid(X, X)
    bcs = [vfvm.Dirichlet(species=1, region=r + 1, value=float(v)) for r, v in enumerate(vals)]
    src = np.full((1, g.num_nodes), f, order="F")
    ph = vfvm.Physics(flux=vfvm.LinearFlux(eps), reaction=vfvm.SinhReaction(), source=src, bcondition=bcs)
    return vfvm.System(g, ph, species=[1])


def cfg3_control(precision):
    return vfvm.SolverControl(krylov="bicgstab", precon="jacobi", precision=precision, reltol_linear=RELTOL_LINEAR,
                              abstol_linear=0.0, maxiters_linear=3000, inner_reltol=INNER_RELTOL, max_refinements=12,
                              check_every=8)


def run_cfg3(args):
    members = cfg3_members()
    t0 = time.perf_counter()
    systems = [build_poisson(*m) for m in members]
    for s in systems:
        s.handle()
    ndof = sum(s.grid.num_nodes for s in systems)
    nnz = sum(s.sizes()[1] for s in systems)
    batch = vfvm.Batch(systems)
    batch.set_graphs(not args.no_graphs)
    log(f"cfg3: {len(systems)} instances, {ndof} unknowns, nnz {nnz}, setup {time.perf_counter() - t0:.1f}s, "
        f"autotuned groups {batch.groups}")
    K, W = args.steps, args.warmup
    res = {}
    probes = {}
    # the kernel-per-step form (one launch per BiCGStab step for all members, iteration
    # blocks as graphs) and the one-launch form (each member's whole BiCGStab in one block)
    for label, prec, kernels, mblock in (("batched_fp64", "fp64", (_capi.KERNEL_SPMV64,), 0),
                                         ("batched_fp32", "fp32", (_capi.KERNEL_SPMV32, _capi.KERNEL_ASSEMBLE), 0),
                                         ("batched_fp32_one_launch", "fp32", (_capi.KERNEL_MSOLVE32,), 1024)):
        batch.set_member_solver(mblock)
        ctrl = cfg3_control(prec)
        for _ in range(W):
            batch.solve(0.0, ctrl)
        t1 = time.perf_counter()
        for _ in range(K):
            U = batch.solve(0.0, ctrl)
        el = time.perf_counter() - t1
        with systems[0].ctx.probe(*kernels) as pr:  # probe pass (launched one by one, not as graphs)
            batch.solve(0.0, ctrl)
        probes.update(pr)
        worst = max(h.residual_norm for h in batch.histories)
        if not worst <= 1e-10:
            raise RuntimeError(f"{label}: a member ended with ||F||_inf = {worst:.3e}")
        res[label] = dict(seconds=el, solves_per_s=len(systems) * K / el, U=[vfvm.dofs(u).copy() for u in U],
                          newton=sum(h.niter for h in batch.histories), krylov=sum(h.nlin for h in batch.histories))
        log(f"{label}: {K} x {len(systems)} solves in {el:.3f}s -> {res[label]['solves_per_s']:.1f} solves/s "
            f"(newton {res[label]['newton']}, krylov {res[label]['krylov']}, worst ||F|| {worst:.2e})")
    batch.set_member_solver(0)
    # plain GPU Newton: the same 256 solves one system at a time, FP64
    ctrl = cfg3_control("fp64")
    vfvm.solve(systems[0], inival=0.0, control=ctrl)
    t1 = time.perf_counter()
    Useq = []
    for s in systems:
        Useq.append(vfvm.dofs(vfvm.solve(s, inival=0.0, control=ctrl)).copy())
        if not s.history.residual_norm <= 1e-10:
            raise RuntimeError(f"sequential: ||F||_inf = {s.history.residual_norm:.3e}")
    el = time.perf_counter() - t1
    plain = len(systems) / el
    log(f"sequential plain GPU Newton: {len(systems)} solves in {el:.2f}s -> {plain:.1f} solves/s")
    diff64 = max(float(np.abs(a - b).max()) for a, b in zip(res["batched_fp64"]["U"], Useq))
    diff32 = max(float(np.abs(a - b).max()) for a, b in zip(res["batched_fp32"]["U"], Useq))
    diff1 = max(float(np.abs(a - b).max()) for a, b in zip(res["batched_fp32_one_launch"]["U"], Useq))
    p32, pasm, p64 = probes[_capi.KERNEL_SPMV32], probes[_capi.KERNEL_ASSEMBLE], probes[_capi.KERNEL_SPMV64]
    rl = roofline_hbm(probe_bytes("spmv", p32, sv=4), p32, read_traffic("cfg3_bspmv32"), "batched FP32 SpMV")
    rl["kernel"] = "batched FP32 SpMV over the active members (k_bspmv)"
    rl_asm = roofline_hbm(probe_bytes("assemble", pasm, transient=False), pasm, read_traffic("cfg3_assemble"),
                          "batched assembly")
    rl64 = roofline_hbm(probe_bytes("spmv", p64, sv=8), p64, read_traffic("cfg3_bspmv64"), "batched FP64 SpMV")
    pms = probes[_capi.KERNEL_MSOLVE32]
    rl_ms = roofline_hbm(probe_bytes("msolve", pms, sv=4), pms, read_traffic("cfg3_msolve32"),
                         "one-launch batched BiCGStab (FP32)")
    rl_ms["kernel"] = "k_msolve: each member's whole BiCGStab/Jacobi in one 1024-thread block, one launch per solve"
    return dict(value=res["batched_fp32"]["solves_per_s"], ms_per_step=1e3 * res["batched_fp32"]["seconds"] / K,
                plain_gpu_newton_solves_per_s=plain,
                speedup_vs_plain=res["batched_fp32"]["solves_per_s"] / plain,
                batched_fp64_solves_per_s=res["batched_fp64"]["solves_per_s"],
                batched_fp32_one_launch_solves_per_s=res["batched_fp32_one_launch"]["solves_per_s"],
                max_abs_diff_vs_plain=dict(batched_fp64=diff64, batched_fp32=diff32, batched_fp32_one_launch=diff1),
                roofline=rl, roofline_assemble=rl_asm, roofline_spmv64=rl64, roofline_one_launch=rl_ms,
                config=dict(workload="cfg3: 256 nonlinear-Poisson instances, N x N grids with N in "
                                     "{33,65,129,257}, seeded eps/boundary values/source, batched "
                                     "BiCGStab+Jacobi, FP32 inside FP64 Newton (plain: 256 sequential FP64 "
                                     "GPU Newton solves)", instances=len(systems), unknowns=ndof, nnz=nnz,
                            hip_graphs=not args.no_graphs),
                dtype="f64+f32")


def cpu_cfg3(args):
    """The oracle on a bounded sample: the first 24 of the 256 instances, direct solves."""
    import cases
    import oracle as O
    members = cfg3_members(24)
    t = 0.0
    for (n, eps, vals, f) in members:
        s, o = cases.nonlinear_poisson_2d(n, eps, f, tuple(vals))
        t0 = time.perf_counter()
        o.solve_step(np.zeros(s.grid.num_nodes))
        t += time.perf_counter() - t0
    return dict(value=len(members) / t, unit="solves/s", cores=1, kind="port",
                sample=f"the first {len(members)} of the 256 cfg3 instances (seeds 0..{len(members) - 1}), "
                       f"oracle Newton with banded direct solves, {t:.2f} s")


# --------------------------------------------------------------------------- cfg4
CFG4_BATCH = 512  # members per batched solve (the 1,024 held-out seeds form 2 blocks)
def run_cfg4(args):
    """Learned initial guess + batching + FP32 Krylov vs plain FP64 GPU Newton from zero,
    on the 1,024 held-out seeds of the 129 x 129 nonlinear Poisson family."""
    from concurrent.futures import ThreadPoolExecutor
    from vfvm import surrogate as S
    from vfvm.system import System
    B = CFG4_BATCH
    seeds = list(S.EVAL_SEEDS)
    K = args.steps
    # the held-out seeds in blocks of B, one batch per block (each holds its block's
    # sources and side values); step k solves block k mod (number of blocks)
    blocks = [seeds[B * k:B * (k + 1)] for k in range(len(seeds) // B)]
    nb = len(blocks)
    fams = [S.PoissonFamily(size=B) for _ in range(nb)]
    fam = fams[0]
    for f in fams:
        f.batch.set_graphs(not args.no_graphs)
        f.batch.set_member_solver(args.member_block)
    # BiCGStab preconditioned by the smoothed-aggregation AMG V-cycle (the reference's
    # SmoothedAggregationPreconBuilder, examples/DevEx003_Solvers.jl:149-156), FP32 inside
    # FP64 refinement; the FP32 solve aims at reltol_linear at once (one refinement sweep
    # per Newton step), and the solver states are read after every iteration
    ctrl32 = vfvm.SolverControl(precision="fp32", precon="amg", reltol_linear=RELTOL_LINEAR, abstol_linear=0.0,
                                maxiters_linear=3000, inner_reltol=RELTOL_LINEAR, max_refinements=12, check_every=1)
    ctx = fam.systems[0].ctx
    n = fam.grid.num_nodes
    ndof, off = fam.batch.sizes()
    assert ndof == B * n and (np.diff(off) == n).all()  # packed member order = sample-major
    # the network on a second context (its own stream): the next block's inference runs
    # while the current block's Newton solve occupies the first
    ctx2 = _capi.Context(0)
    net = S.Surrogate(System(fam.grid, fam.systems[0].physics, species=[1], context=ctx2))
    meta = net.load()
    # device-resident inputs and results (the timed region starts with every block's
    # sources and side values in HBM and leaves the guesses and solutions there)
    raws = [f.load_seeds(blk) for f, blk in zip(fams, blocks)]
    F_dev = [vfvm.DeviceArray.from_numpy(ctx, r[0]) for r in raws]
    V_dev = [vfvm.DeviceArray.from_numpy(ctx, r[1]) for r in raws]
    G_dev = [vfvm.DeviceArray(ctx, ndof) for _ in range(nb)]
    U_dev = [vfvm.DeviceArray(ctx, ndof) for _ in range(nb)]
    Z_dev = vfvm.DeviceArray.from_numpy(ctx, np.zeros(ndof))

    def infer(k):
        net.predict_sides_device(F_dev[k % nb], V_dev[k % nb], G_dev[k % nb], B)  # features, smoothing, network

    def pipeline(steps):
        """steps batched solves; inference of step k + 1 overlaps the Newton solve of step k"""
        hist = []
        infer(0)
        with ThreadPoolExecutor(max_workers=1) as ex:
            for k in range(steps):
                nxt = ex.submit(infer, k + 1) if k + 1 < steps else None
                fams[k % nb].batch.solve_device(U_dev[k % nb], G_dev[k % nb], ctrl32)
                hist.append(list(fams[k % nb].batch.histories))
                if nxt is not None:
                    nxt.result()
        return hist

    ctrl64b = vfvm.SolverControl(precision="fp64", precon="amg", reltol_linear=RELTOL_LINEAR, abstol_linear=0.0,
                                 maxiters_linear=3000, check_every=1)
    for _ in range(max(args.warmup, 0)):  # warm-up (untimed): every batch, every path
        pipeline(nb)
        for f in fams:
            f.batch.solve_device(U_dev[0], Z_dev, ctrl32)
            f.batch.solve_device(U_dev[0], Z_dev, ctrl64b)
        net.predict_sides(*raws[0])
    # timed: K steps back to back
    t1 = time.perf_counter()
    hist = pipeline(K)
    t_sur = time.perf_counter() - t1
    it_sur = [h.niter for hs in hist for h in hs]
    worst = max(h.residual_norm for hs in hist for h in hs)
    # beside it, per block: the host-buffer path, a probe pass, and the batched solves from zero
    t_sur_host, t_zero, t_zero64, it_zero, sur_err = 0.0, 0.0, 0.0, [], []
    converged = []
    kinds = ((_capi.KERNEL_MSOLVE32,) if args.member_block else (_capi.KERNEL_SPMV32, _capi.KERNEL_AMG32)) + \
        (_capi.KERNEL_ASSEMBLE,)
    probes = {k: [] for k in kinds}
    for b in range(min(K, nb)):
        f = fams[b]
        infer(b)
        f.batch.solve_device(U_dev[b], G_dev[b], ctrl32)
        g = G_dev[b].numpy().reshape(B, n)
        U = U_dev[b].numpy().reshape(B, n)
        t1 = time.perf_counter()  # the host-buffer entry points (PCIe transfers and packing included)
        gh = net.predict_sides(*raws[b])
        Uh = f.batch.solve([x.reshape(1, -1) for x in gh], ctrl32)
        t_sur_host += time.perf_counter() - t1
        worst = max(worst, max(h.residual_norm for h in f.batch.histories))
        with ctx.probe(*kinds) as pr:  # timed launch by launch
            f.batch.solve_device(U_dev[b], G_dev[b], ctrl32)
        for k, v in pr.items():
            probes[k].append(v)
        t1 = time.perf_counter()
        f.batch.solve_device(U_dev[b], Z_dev, ctrl32)
        t_zero += time.perf_counter() - t1
        it_zero += [h.niter for h in f.batch.histories]
        worst = max(worst, max(h.residual_norm for h in f.batch.histories))
        t1 = time.perf_counter()  # batching alone: FP64 Krylov from zero
        f.batch.solve_device(U_dev[b], Z_dev, ctrl64b)
        t_zero64 += time.perf_counter() - t1
        worst = max(worst, max(h.residual_norm for h in f.batch.histories))
        for m in range(B):
            sur_err.append(float(np.abs(g[m] - U[m]).max()))
            if not np.abs(vfvm.dofs(Uh[m]) - U[m]).max() <= 1e-9:
                raise RuntimeError("cfg4: host-buffer and device-resident solutions differ")
        converged += [U[m].copy() for m in range(B)]
    if not worst <= 1e-10:
        raise RuntimeError(f"cfg4: a solve ended with ||F||_inf = {worst:.3e}")
    nfill = B * min(K, nb)
    nsol = B * K
    sur_rate, zero_rate = nsol / t_sur, nfill / t_zero
    log(f"surrogate+batched+FP32 through the host-buffer entry points (PCIe and packing included, no "
        f"overlap): {nfill / t_sur_host:.1f} solves/s")
    log(f"surrogate+batched+FP32: {nsol} solves in {t_sur:.3f}s -> {sur_rate:.1f} solves/s, "
        f"mean newton {np.mean(it_sur):.2f}; batched+FP32 from zero {zero_rate:.1f} solves/s, mean newton "
        f"{np.mean(it_zero):.2f}; surrogate max-norm error mean {np.mean(sur_err):.3e} max {np.max(sur_err):.3e}")
    plain_rate = cont_rate = None
    diff = cdiff = float("nan")
    it_plain, st_cont, it_cont = [float("nan")], [float("nan")], [float("nan")]
    ncont = 0
    if not args.batched_only:
        plain_rate, it_plain, diff = cfg4_plain(args, S, fam, blocks, K, converged)
        cont_rate, ncont, st_cont, it_cont, cdiff = cfg4_continuation(S, fam, blocks, nsol, converged)
    # roofline: batched FP32 SpMV (the Krylov hot op) and the Jacobian assembly, timed live
    # inside the surrogate solves (work counted over the members still active), and the
    # network's layer GEMM
    return cfg4_result(args, S, fam, net, meta, probes, B, K, nb, nsol, t_sur, sur_rate, zero_rate, nfill, t_zero64,
                       t_sur_host, it_sur, it_zero, sur_err, worst, plain_rate, it_plain, diff, cont_rate, ncont,
                       st_cont, it_cont, cdiff)


def cfg4_plain(args, S, fam, blocks, K, converged):
    """plain FP64 GPU Newton from zero, one system at a time, on every seed the timed steps solved"""
    single = S.make_system(fam.grid, np.zeros(fam.grid.num_nodes), np.zeros(4))
    ctrl64 = vfvm.SolverControl(precision="fp64", reltol_linear=RELTOL_LINEAR, abstol_linear=0.0, maxiters_linear=3000,
                                check_every=8)
    single.handle()
    vfvm.solve(single, inival=0.0, control=ctrl64)
    t_plain, it_plain, diff = 0.0, [], 0.0
    k = 0
    for blk in blocks[:min(K, len(blocks))]:  # every distinct seed the timed steps solved
        for sd in blk:
            f, vals = S.draw(sd, fam.grf)
            single.set_source(np.asfortranarray(f.reshape(1, -1)))
            single.update_bcondition(S.bcs(vals))
            single.handle()
            t1 = time.perf_counter()
            Up = vfvm.solve(single, inival=0.0, control=ctrl64)
            t_plain += time.perf_counter() - t1
            if not single.history.residual_norm <= 1e-10:
                raise RuntimeError("plain solve did not reach ||F|| <= 1e-10")
            it_plain.append(single.history.niter)
            diff = max(diff, float(np.abs(vfvm.dofs(Up) - converged[k]).max()))
            k += 1
    plain_rate = k / t_plain
    log(f"plain FP64 GPU Newton from zero: {k} solves in {t_plain:.2f}s -> {plain_rate:.1f} solves/s, "
        f"mean newton {np.mean(it_plain):.2f}; max |U_sur - U_plain| = {diff:.2e}")
    return plain_rate, it_plain, diff


def cfg4_continuation(S, fam, blocks, nsol, converged):
    """continuation on a sample: source and side values ramped by the embedding parameter
    over [0, 1] (solve(...; embed)), plain FP64 GPU Newton at every accepted step"""
    ncont = min(32, nsol)
    from vfvm.physics import Physics, LinearFlux, SinhReaction, Dirichlet
    cph = Physics(flux=LinearFlux(S.EPS), reaction=SinhReaction(), source=np.zeros((1, fam.grid.num_nodes), order="F"),
                  bcondition=[Dirichlet(species=1, region=r + 1, value=0.0, value_lambda=0.0) for r in range(4)],
                  embed_source=True)
    cont = vfvm.System(fam.grid, cph, species=[1])
    cctrl = vfvm.SolverControl(precision="fp64", reltol_linear=RELTOL_LINEAR, abstol_linear=0.0, maxiters_linear=3000,
                               check_every=8, Δp=0.25, Δp_min=1e-2, Δp_max=0.5, Δp_grow=1.5, Δu_opt=0.5)
    t_cont, it_cont, st_cont, cdiff = 0.0, [], [], 0.0
    for k, sd in enumerate(blocks[0][:ncont]):
        f, vals = S.draw(sd, fam.grf)
        cont.set_source(np.asfortranarray(f.reshape(1, -1)))
        cont.update_bcondition([Dirichlet(species=1, region=r + 1, value=0.0, value_lambda=float(v))
                                for r, v in enumerate(vals)])
        cont.handle()
        t1 = time.perf_counter()
        sol = vfvm.solve(cont, inival=0.0, control=cctrl, embed=[0.0, 1.0])
        t_cont += time.perf_counter() - t1
        if not cont.history.residual_norm <= 1e-10:
            raise RuntimeError("continuation did not reach ||F|| <= 1e-10")
        it_cont.append(cont.history.niter)  # Newton iterations over all steps
        st_cont.append(len(sol.t) - 1)
        cdiff = max(cdiff, float(np.abs(vfvm.dofs(sol.u[-1]) - converged[k]).max()))
    cont_rate = ncont / t_cont
    log(f"continuation (plain FP64 GPU, {ncont} seeds): {cont_rate:.1f} solves/s, mean accepted steps "
        f"{np.mean(st_cont):.1f}, mean newton {np.mean(it_cont):.1f}; max |U_cont - U_sur| = {cdiff:.2e}")
    return cont_rate, ncont, st_cont, it_cont, cdiff


def cfg4_result(args, S, fam, net, meta, probes, B, K, nb, nsol, t_sur, sur_rate, zero_rate, nfill, t_zero64,
                t_sur_host, it_sur, it_zero, sur_err, worst, plain_rate, it_plain, diff, cont_rate, ncont, st_cont,
                it_cont, cdiff):
    def merged(lst):
        return dict(total_ms=sum(x["total_ms"] for x in lst), launches=sum(x["launches"] for x in lst),
                    avg_ms=sum(x["total_ms"] for x in lst) / max(sum(x["launches"] for x in lst), 1),
                    work=tuple(sum(x["work"][i] for x in lst) for i in range(3)))
    pasm = merged(probes[_capi.KERNEL_ASSEMBLE])
    extra = {}
    if _capi.KERNEL_MSOLVE32 in probes:
        # the whole FP32 Krylov solve of every member, one block per member (k_msolve):
        # algorithmic bytes from the iterations each member took
        pm = merged(probes[_capi.KERNEL_MSOLVE32])
        rl_main = roofline_hbm(probe_bytes("msolve", pm, sv=4), pm, read_traffic("cfg4_msolve32"),
                               "one-block-per-member BiCGStab + AMG (FP32)")
        rl_main["kernel"] = (f"k_msolve: each of the {B} members' right-preconditioned BiCGStab (two operator "
                             f"applications and two AMG V(1,1) cycles per iteration) in one {args.member_block}-"
                             f"thread block, one launch per linear solve of the batch")
    else:
        pspmv, pamg = merged(probes[_capi.KERNEL_SPMV32]), merged(probes[_capi.KERNEL_AMG32])
        rl_spmv = roofline_hbm(probe_bytes("spmv", pspmv, sv=4), pspmv, read_traffic("cfg4_bspmv32"),
                               "batched FP32 SpMV")
        rl_spmv["kernel"] = f"batched FP32 SpMV over the active members of {B} x 129^2 (k_bspmv)"
        extra["roofline_bspmv"] = rl_spmv
        # the AMG V-cycle's finest-level post-smoothing sweep: an SpMV of the row-scaled
        # operator plus b read once more, over the members still iterating
        rl_main = roofline_hbm(probe_bytes("amg_sweep", pamg, sv=4), pamg, read_traffic("cfg4_amg_sweep32"),
                               "AMG finest-level sweep (FP32)")
        rl_main["kernel"] = (f"AMG V-cycle finest-level Jacobi sweep over the active members of {B} x 129^2 "
                             f"(k_amg_sweep: FP16 SELL-64 values, shared column indices, FP32 vectors)")
    rl_asm = roofline_hbm(probe_bytes("assemble", pasm, transient=False), pasm, read_traffic("cfg4_assemble"),
                          "batched assembly")
    import ctypes as C
    v = C.c_double()
    rows = B * fam.grid.num_nodes  # one inference batch
    _capi.check(_capi.lib().vfvm_mpnn_time_layer(net._h, rows, 20, C.byref(v)))
    t_layer = v.value
    layer_bytes = rows * 64 * 4 * 2  # h_l read once, h_{l+1} written once (FP32, 64 wide)
    layer_gbs = layer_bytes / (t_layer * 1e-3) / 1e9
    layer_tflops = 2.0 * rows * 128 * 64 / (t_layer * 1e-3) / 1e12
    log(f"inference layer (aggregation + 128x64 MFMA GEMM) over {rows} rows: {t_layer * 1e3:.1f} us, "
        f"{layer_gbs:.0f} GB/s, {layer_tflops:.1f} TFLOP/s")
    return dict(value=sur_rate, ms_per_step=1e3 * t_sur / K, steps=K, plain_gpu_newton_solves_per_s=plain_rate,
                speedup_vs_plain=(sur_rate / plain_rate) if plain_rate else None,
                batched_fp32_from_zero_solves_per_s=zero_rate,
                batched_fp64_from_zero_solves_per_s=nfill / t_zero64,
                host_buffer_path_solves_per_s=nfill / t_sur_host,
                continuation_plain_solves_per_s=cont_rate,
                continuation=dict(seeds=ncont, mean_accepted_steps=float(np.mean(st_cont)),
                                  mean_newton_iterations=float(np.mean(it_cont)),
                                  max_abs_diff_vs_surrogate=cdiff),
                newton_iterations=dict(surrogate=float(np.mean(it_sur)), zero_batched=float(np.mean(it_zero)),
                                       plain=float(np.mean(it_plain))),
                surrogate_max_norm_error=dict(mean=float(np.mean(sur_err)), max=float(np.max(sur_err))),
                max_abs_diff_vs_plain=diff, worst_residual=worst,
                surrogate=dict(train_samples=int(meta["samples"]), epochs=int(meta["epochs"]),
                               final_train_mse=float(meta["final_loss"])),
                roofline=rl_main, **extra, roofline_assemble=rl_asm,
                roofline_mpnn_layer=dict(bound="hbm", achieved=round(layer_gbs, 1), peak=8000.0, unit="GB/s",
                                         frac=round(layer_gbs / 8000.0, 4), traffic=read_traffic("cfg4_mpnn_layer"),
                                         traffic_note="mean over the four layers of one inference; the last "
                                                      "writes u only",
                                         algorithmic_bytes=int(layer_bytes), avg_ms=round(t_layer, 5),
                                         mfma_tflops=round(layer_tflops, 2), rows=rows,
                                         kernel="k_layer_infer: graph aggregation + 128x64 BF16 MFMA GEMM + residual"),
                config=dict(workload="cfg4: 129x129 nonlinear Poisson (eps=1e-2, GRF source ell=0.1 amp=5, side "
                                     "values U[-1,1]), 1,024 held-out seeds in 2 batches of 512 (step k solves "
                                     "batch k mod 2; the inference of step k+1 runs on a second stream during "
                                     "the Newton solve of step k): MPNN initial guess "
                                     "(4 layers, width 64, BF16 MFMA) + batched BiCGStab preconditioned by "
                                     "smoothed-aggregation AMG (each member's solve in one block, one launch per "
                                     "linear solve of the batch), FP32 inside FP64 Newton, vs plain FP64 GPU Newton "
                                     "(BiCGStab/Jacobi) from zero one solve at a time; reltol_linear 1e-4 (the "
                                     "reference's default) in every form",
                            solves=nsol, unknowns_per_solve=fam.grid.num_nodes, batch=B,
                            hip_graphs=not args.no_graphs, member_solver_block=args.member_block,
                            resident="sources, side values, guesses and solutions in HBM (the *_device "
                                     "entry points); host_buffer_path_solves_per_s includes the PCIe "
                                     "transfers"),
                dtype="f64+f32+bf16")


def cpu_cfg4(args):
    """The oracle on bounded samples of the held-out seeds, Newton from zero with banded
    direct solves: one thread, then every core of the host's share (independent solves
    on a thread pool; the oracle's ctypes calls release the GIL).  `value` is the
    all-core rate."""
    from concurrent.futures import ThreadPoolExecutor
    import cases
    from vfvm import surrogate as S
    grf = S.GRF()
    seeds = list(S.EVAL_SEEDS)

    def build(sd):
        f, vals = S.draw(sd, grf)
        _, o = cases.nonlinear_poisson_2d(S.GRID_N, S.EPS, 0.0, tuple(vals))
        o.set_source(f)
        return o

    def run(o):
        o.solve_step(np.zeros(S.GRID_N * S.GRID_N))

    n1 = 8
    systems = [build(sd) for sd in seeds[:n1]]
    t0 = time.perf_counter()
    for o in systems:
        run(o)
    t1 = time.perf_counter() - t0
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))  # the host share of one GPU
    nall = 4 * cores
    systems = [build(sd) for sd in seeds[:nall]]
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=cores) as ex:
        list(ex.map(run, systems))
    tn = time.perf_counter() - t0
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return dict(value=nall / tn, unit="solves/s", cores=cores, kind="port", single_thread_value=n1 / t1,
                cpu_model=model,
                sample=f"held-out cfg4 seeds, oracle Newton from zero with banded direct solves: {nall} seeds on "
                       f"{cores} threads in {tn:.2f} s; {n1} seeds on one thread in {t1:.2f} s")


# --------------------------------------------------------------------------- copy
def run_copy(args):
    """Device-to-device copy of 1 GiB (FP32, torch's elementwise copy kernel): the
    measured HBM bandwidth reference, and the kernel of known size against which the
    rocprofv3 FETCH_SIZE / WRITE_SIZE counters are scaled."""
    import torch
    n = 256 << 20
    a = torch.ones(n, device="cuda", dtype=torch.float32)
    b = torch.empty_like(a)
    for _ in range(max(args.warmup, 1)):
        b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(args.steps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    gbs = 2 * 4 * n / (ms * 1e-3) / 1e9
    log(f"copy 1 GiB: {ms:.3f} ms -> {gbs:.0f} GB/s read+write")
    return dict(value=gbs, ms_per_step=ms, unit="GB/s", metric="HBM copy bandwidth (read + write bytes / time)",
                roofline=dict(bound="hbm", achieved=round(gbs, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                              frac=round(gbs / HBM_PEAK_GBS, 4), traffic=None, algorithmic_bytes=8 * n,
                              avg_ms=round(ms, 5)),
                config=dict(workload="copy: 1 GiB FP32 device-to-device (torch copy kernel)"), dtype="f32")


def cpu_copy(args):
    return None


# --------------------------------------------------------------------------- impedance
IMP_N = 257
IMP_OMEGAS = 1e-3 * 1.1 ** np.arange(113)  # Example150's sweep: w from 1e-3 by factors 1.1 below 50


def build_impedance(n, oracle=False):
    """Example150's problem on the unit square: C u_t - div(D grad u) + R u = 0 (C = D = R = 1),
    u = 1 on the west side (region 4, excited), u = 0 on the east side (region 2, measured)."""
    X = vfvm.linspace(0.0, 1.0, n)
    g = vfvm.simplexgrid(X, X)
    phys = dict(flux=vfvm.LinearFlux(1.0), reaction=vfvm.PowerReaction(m=1, k=1.0), storage=vfvm.IdentityStorage())
    s = vfvm.System(g, vfvm.Physics(**phys), species=[1])
    s.boundary_dirichlet(1, 4, 1.0)
    s.boundary_dirichlet(1, 2, 0.0)
    o = None
    if oracle:  # the checker's twin (the CPU baseline leg only)
        import cases
        o = cases.oracle_twin(s)
    return g, s, o, vfvm.Physics(**phys)


def run_impedance(args):
    """Impedance spectrum of Example150's problem on a 257 x 257 grid over Example150's
    113 frequencies: steady state, test function and measurement rows, then the whole
    sweep solved as one batch (complex BiCGStab, relative residual 1e-10)."""
    n = IMP_N
    g, s, _, _ = build_impedance(n)
    c = vfvm.SolverControl(reltol_linear=1e-12, abstol_linear=0.0, maxiters_linear=20000, precon="ilu0")
    U0 = vfvm.solve(s, inival=0.0, control=c)
    T = vfvm.testfunction(s, [4], [2])
    ds, dt = vfvm.measurement_derivatives(s, U0, T, 1)
    isys = vfvm.ImpedanceSystem(s, U0, 1, 4)
    for _ in range(max(args.warmup, 0)):
        isys.impedances(IMP_OMEGAS, ds, dt, reltol=1e-10)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        Z = isys.impedances(IMP_OMEGAS, ds, dt, reltol=1e-10)
    el = (time.perf_counter() - t0) / args.steps
    it = isys.iterations
    log(f"impedance: {g.num_nodes} nodes, {IMP_OMEGAS.size} frequencies in {el:.3f} s per sweep -> "
        f"{IMP_OMEGAS.size / el:.1f} frequencies/s, BiCGStab iterations {it.min()}..{it.max()}")
    return dict(value=IMP_OMEGAS.size / el, unit="frequencies/s", ms_per_step=1e3 * el,
                metric="impedance spectrum: frequencies solved per second (one sweep of 113 as a batch)",
                iterations_min=int(it.min()), iterations_max=int(it.max()),
                Z_first=[float(Z[0].real), float(Z[0].imag)], Z_last=[float(Z[-1].real), float(Z[-1].imag)],
                config=dict(workload=f"impedance: Example150 problem on a {n}x{n} grid, 113 frequencies "
                                     f"1e-3..50, reltol 1e-10", nodes=g.num_nodes), dtype="c128")


def cpu_impedance(args):
    """The oracle's form of the same sweep (scipy's sparse LU per frequency, as the
    reference's solve!) on a bounded sample: 8 of the 113 frequencies."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    n = IMP_N
    g, _, osys, _ = build_impedance(n, oracle=True)
    U0, _ = osys.solve_step(np.zeros(g.num_nodes))
    osys.assemble(U0, U0, tstep=1e30)
    J0 = osys.matrix_scipy().tocsc()
    M = sp.diags(osys.nodevolumes())
    b = np.zeros(g.num_nodes, dtype=complex)
    for f, r in zip(g.bfacenodes.T, g.bfaceregions):
        if r == 4:
            for k in f:
                b[k] += 1e30
    om = IMP_OMEGAS[::14][:8]
    t0 = time.perf_counter()
    for w in om:
        spla.spsolve((J0 + 1j * w * M).tocsc(), b)
    el = time.perf_counter() - t0
    return dict(value=om.size / el, unit="frequencies/s", cores=1, kind="port",
                sample=f"{om.size} of the 113 frequencies on the {n}x{n} grid, scipy sparse LU per frequency "
                       f"(the reference's solve! factorises per frequency), {el:.2f} s")


WORKLOADS = {"impedance": (run_impedance, cpu_impedance), "copy": (run_copy, cpu_copy), "cfg0": (run_cfg0, cpu_cfg0), "cfg1": (run_cfg1, cpu_cfg1), "cfg2": (run_cfg2, cpu_cfg2), "cfg3": (run_cfg3, cpu_cfg3), "cfg4": (run_cfg4, cpu_cfg4)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="cfg4", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batched-only", action="store_true",
                    help="cfg4: skip the plain one-at-a-time and continuation legs (profiling runs)")
    ap.add_argument("--member-block", type=int, default=1024, choices=(0, 256, 512, 1024),
                    help="cfg4: batched Krylov solves as one block of this many threads per member, one launch "
                         "per linear solve (0: one launch per BiCGStab step for all members)")
    ap.add_argument("--no-graphs", action="store_true",
                    help="batched Krylov blocks launched kernel by kernel instead of as HIP graphs "
                         "(for kernel-trace profiling)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (requested --gpus {args.gpus}); multi-GPU runs are not part "
              f"of this build", file=sys.stderr)
        sys.exit(2)
    run, cpu = WORKLOADS[args.workload]
    r = run(args)
    line = dict(metric=r.pop("metric", METRIC), value=r.pop("value"), unit=r.pop("unit", "solves/s"), n_gpus=1,
                steps=r.pop("steps", args.steps),
                warmup=args.warmup, ms_per_step=r.pop("ms_per_step"), higher_is_better=True, scaling="weak",
                vs_baseline=None, dtype=r.pop("dtype"), data="synthetic", config=r.pop("config"))
    line.update(r)
    line["cpu_baseline"] = None if args.no_cpu else cpu(args)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
