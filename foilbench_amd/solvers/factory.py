"""Solver registry (fb/solvers/factory.py:8-22): the reference's solver IDs, each backed by the
gfx950 implementation. ``form`` selects the plain (FP64 drop-in) or a precision form."""

from __future__ import annotations

_SOLVERS = ("stable-fluids", "lbm-d2q9", "pic-flip")
_FORMS = {
    "stable-fluids": ("plain", "mixed", "precond32"),
    "lbm-d2q9": ("plain", "f64", "f32", "f32-mixed", "f16-mixed"),
    "pic-flip": ("plain", "mixed"),
}


def solver_ids() -> tuple[str, ...]:
    return _SOLVERS


def solver_forms(solver_id: str) -> tuple[str, ...]:
    if solver_id not in _FORMS:
        raise KeyError(f"unknown solver {solver_id!r}; choose from {', '.join(_SOLVERS)}")
    return _FORMS[solver_id]


def create_solver(solver_id: str, form: str = "plain", **kwargs):
    forms = solver_forms(solver_id)
    if form not in forms:
        raise ValueError(f"solver {solver_id!r} has forms {forms}, not {form!r}")
    if solver_id == "stable-fluids":
        from .stable_fluids import StableFluidsSolver

        return StableFluidsSolver(form, **kwargs)
    if solver_id == "lbm-d2q9":
        from .lbm import LBMSolver

        return LBMSolver(form, **kwargs)
    from .pic_flip import PicFlipSolver

    return PicFlipSolver(form, **kwargs)
