"""Host-side logic of the Python mirror (no device calls): SolverControl defaults
and the step-control fields it hands to the library."""
import sys

import pytest

import vfvm


def test_solver_control_defaults_follow_reference():
    # src/vfvm_solvercontrol.jl:36-211
    c = vfvm.SolverControl()
    assert (c.abstol, c.reltol, c.maxiters, c.tol_mono, c.damp_initial, c.damp_growth, c.max_round) == \
        (1e-10, 1e-10, 100, 1e-3, 1.0, 1.2, 1000)
    assert (c.Δt, c.Δt_min, c.Δt_max, c.Δt_grow, c.Δt_decrease) == (0.1, 1e-3, 1.0, 1.2, 0.5)
    assert (c.Δp, c.Δp_min, c.Δp_max, c.Δp_grow, c.Δp_decrease) == (1.0, 1e-3, 1.0, 1.0, 0.5)
    assert (c.Δu_opt, c.Δu_max_factor, c.num_final_steps, c.force_first_step) == (0.1, 1.2, 5, False)
    assert c.handle_exceptions is False  # src/vfvm_solvercontrol.jl:216


def test_step_control_fields():
    c = vfvm.SolverControl(Δp=0.05, Δp_min=0.01, Δp_max=0.2)
    t = c.to_step_c(True)
    e = c.to_step_c(False)
    assert (t.d, t.d_min, t.d_max, t.d_grow) == (0.1, 1e-3, 1.0, 1.2)
    assert (e.d, e.d_min, e.d_max, e.d_grow) == (0.05, 0.01, 0.2, 1.0)
    assert e.du_opt == 0.1 and e.num_final_steps == 5
    assert t.handle_exceptions == 0
    assert vfvm.SolverControl(handle_exceptions=True).to_step_c(True).handle_exceptions == 1
    # the library's defaults agree with the mirror's
    from vfvm import _capi
    for tr in (1, 0):
        d = _capi.StepControlT()
        _capi.lib().vfvm_default_stepcontrol(tr, d)
        m = vfvm.SolverControl().to_step_c(bool(tr))
        assert [getattr(d, f) for f, _ in d._fields_] == [getattr(m, f) for f, _ in m._fields_]


def test_fixed_timesteps():
    # src/vfvm_solvercontrol.jl:320-327
    c = vfvm.fixed_timesteps(vfvm.SolverControl(), 0.25, grow=1.5)
    assert (c.Δt, c.Δt_min, c.Δt_max, c.Δt_grow) == (0.25, 0.25, 0.25, 1.5)
    assert c.Δu_opt == sys.float_info.max


def test_product_and_gpu_problem_builders_do_not_load_the_oracle():
    """The package, including the benchmark configurations' problem builders
    (vfvm.problems), must not import the oracle or the test suite's helpers: the oracle
    is the checker, loaded only by tests and the CPU baselines."""
    import subprocess
    import sys as _sys
    root = __import__("os").path.dirname(__import__("os").path.dirname(__import__("os").path.abspath(__file__)))
    code = ("import sys; sys.path[:0] = ['voronoifvm.jl_amd']\n"
            "import vfvm, vfvm.surrogate, vfvm.impedance, vfvm.problems as P\n"
            "P.nonlinear_poisson_2d(9); P.drift_diffusion_2d(9); P.nonlinear_diffusion_3d(5)\n"
            "print('oracle' in sys.modules, 'mpnn_oracle' in sys.modules, 'cases' in sys.modules)\n")
    out = subprocess.run([_sys.executable, "-c", code], cwd=root, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    assert out.stdout.strip() == "False False False"


def test_benchmark_gpu_legs_import_no_test_code():
    """bench.py's GPU legs (run_*) build their problems from the package, not from the
    test suite or the oracle (those serve the cpu_baseline legs only)."""
    import ast
    import os
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    tree = ast.parse(open(os.path.join(root, "bench.py")).read())
    for fn in tree.body:
        if isinstance(fn, ast.FunctionDef) and fn.name.startswith("run_"):
            for node in ast.walk(fn):
                names = []
                if isinstance(node, ast.Import):
                    names = [a.name for a in node.names]
                elif isinstance(node, ast.ImportFrom):
                    names = [node.module or ""]
                for n in names:
                    assert n.split(".")[0] not in ("cases", "oracle", "mpnn_oracle"), (fn.name, n)


def test_update_grid_device_refuses_wrong_arrays():
    """update_grid_device checks the array's size and context before any library call"""
    import vfvm
    from vfvm.grids import linspace, simplexgrid
    g = simplexgrid(linspace(0, 1, 4), linspace(0, 1, 3))
    ctx = object()
    s = vfvm.System(g, vfvm.Physics(flux=vfvm.LinearFlux(1.0)), species=[1], context=ctx)

    class Coords:
        def __init__(self, n, c):
            self.n, self.ctx, self.ptr = n, c, 1

    with pytest.raises(ValueError):
        s.update_grid_device(Coords(2 * g.num_nodes - 1, ctx))
    with pytest.raises(ValueError):
        s.update_grid_device(Coords(2 * g.num_nodes, object()))
    assert s._h is None and s.grid is g
