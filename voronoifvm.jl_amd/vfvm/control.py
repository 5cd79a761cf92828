"""SolverControl (src/vfvm_solvercontrol.jl:13-270): the fields on the path plus
the device solver choice."""
from __future__ import annotations

from dataclasses import dataclass, fields

_KRYLOV = {"bicgstab": 0, "gmres": 1}
# "ilu0": ILU(0) in a multicolour order (few launches per triangular solve);
# "ilu0_natural": ILUZero.jl's natural-order ILU(0), level-scheduled
_PRECON = {"jacobi": 0, "blockjacobi": 1, "ilu0": 2, "amg": 3, "ilu0_natural": 4}
_PRECISION = {"fp64": 0, "fp32": 1}


@dataclass
class SolverControl:
    abstol: float = 1.0e-10
    reltol: float = 1.0e-10
    maxiters: int = 100
    tol_round: float = 1.0e-10
    tol_mono: float = 1.0e-3
    damp_initial: float = 1.0
    damp_growth: float = 1.2
    max_round: int = 1000
    updatecontrol: bool = True
    reltol_linear: float = 1.0e-4
    abstol_linear: float = 1.0e-8
    maxiters_linear: int = 100
    # device linear solver: what method_linear selects in the reference
    krylov: str = "bicgstab"
    precon: str = "jacobi"
    precision: str = "fp64"
    gmres_restart: int = 30
    max_refinements: int = 20
    inner_reltol: float = 1.0e-3
    # src/vfvm_solvercontrol.jl:123-131: the preconditioner (ILU(0) factor) of the first
    # Newton step is reused by the later ones unless this is set
    factorize_every_newtonstep: bool = False
    check_every: int = 4
    # time stepping / embedding step control (src/vfvm_solvercontrol.jl:138-211)
    Δp: float = 1.0
    Δp_max: float = 1.0
    Δp_min: float = 1.0e-3
    Δp_grow: float = 1.0
    Δp_decrease: float = 0.5
    Δt: float = 0.1
    Δt_max: float = 1.0
    Δt_min: float = 1.0e-3
    Δt_grow: float = 1.2
    Δt_decrease: float = 0.5
    Δu_opt: float = 0.1
    Δu_max_factor: float = 1.2
    force_first_step: bool = False
    # src/vfvm_solvercontrol.jl:205-216: retry failing Newton steps with smaller steps and
    # return the steps done so far (with a warning) instead of raising
    handle_exceptions: bool = False
    num_final_steps: int = 5
    verbose: bool = False

    def update(self, **kw):
        names = {f.name for f in fields(self)}
        for k, v in kw.items():
            if k in names:
                setattr(self, k, v)
        return self

    def to_c(self, is_linear=False):
        from ._capi import ControlT, lib
        c = ControlT()
        lib().vfvm_default_control(c)
        for name in ("abstol", "reltol", "tol_round", "tol_mono", "damp_initial", "damp_growth", "reltol_linear",
                     "abstol_linear", "inner_reltol"):
            setattr(c, name, float(getattr(self, name)))
        for name in ("maxiters", "max_round", "maxiters_linear", "gmres_restart", "max_refinements", "check_every"):
            setattr(c, name, int(getattr(self, name)))
        c.updatecontrol = int(bool(self.updatecontrol))
        c.reuse_precon = int(not self.factorize_every_newtonstep)
        c.is_linear = int(bool(is_linear))
        c.krylov = _KRYLOV[self.krylov]
        c.precon = _PRECON[self.precon]
        c.precision = _PRECISION[self.precision]
        return c

    def to_step_c(self, transient: bool):
        """vfvm_stepcontrol_t from the Δt fields (time stepping) or the Δp fields (embedding)."""
        from ._capi import StepControlT
        sc = StepControlT()
        pre = "Δt" if transient else "Δp"
        sc.d = float(getattr(self, pre))
        sc.d_min = float(getattr(self, pre + "_min"))
        sc.d_max = float(getattr(self, pre + "_max"))
        sc.d_grow = float(getattr(self, pre + "_grow"))
        sc.d_decrease = float(getattr(self, pre + "_decrease"))
        sc.du_opt = float(self.Δu_opt)
        sc.du_max_factor = float(self.Δu_max_factor)
        sc.num_final_steps = int(self.num_final_steps)
        sc.force_first_step = int(bool(self.force_first_step))
        sc.handle_exceptions = int(bool(self.handle_exceptions))
        return sc


def fixed_timesteps(control: SolverControl, Δt: float, grow: float = 1.0) -> SolverControl:
    """fixed_timesteps!(control, Δt; grow) (src/vfvm_solvercontrol.jl:320-327): time
    steps Δt, Δt grow, Δt grow^2, ... (Δu_opt = floatmax disables the Δu control)."""
    import sys
    control.Δt = Δt
    control.Δt_max = Δt
    control.Δt_min = Δt
    control.Δt_grow = grow
    control.Δu_opt = sys.float_info.max
    return control
