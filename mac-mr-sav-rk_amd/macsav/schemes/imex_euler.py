"""First-order IMEX Euler and its mr-ccSAV form (reference: solver/schemes/imex_euler.py,
imex_euler_mrsav.py)."""
from __future__ import annotations

from typing import Any

import numpy as np

from ..core import Trial
from ..model import Model, stage_divergences
from ..momentum import stage_momentum
from .scalar_correction import solve_scalar_stage


class IMEXEuler:
    name = "imex_euler"

    def step(self, model: Model, state: Any, dt: float) -> Trial:
        if not np.isfinite(dt) or dt <= 0 or not np.isfinite(state.t):
            raise ValueError("Invalid step size or state time")
        mom = stage_momentum(model)
        old = model.vector(state)
        conv = mom.nonlinear(old)
        rhs = model.combine((1.0, mom.mass(old)), (dt, mom.force(state.t)), (-dt, conv))
        solved = mom.solve(rhs, mass=1.0, viscosity=model.nu * dt)
        velocity = solved.velocity
        pressure = model.combine((1 / dt, solved.pressure))
        defect = model.combine((1.0, mom.mass(velocity)), (model.nu * dt, mom.linear(velocity)),
                               (dt, mom.gradient(pressure)), (-1.0, rhs))
        residual = mom.residual(defect, rhs)
        phys, cont = stage_divergences(model, velocity)
        if (not np.isfinite([residual, phys, cont, model.max_abs(velocity), model.max_abs(pressure)]).all()
                or residual > 1e-8 or cont > 1e-8 * (1 + model.max_abs(velocity))):
            raise RuntimeError(f"IMEX Euler residual {residual}, divergence {cont}")
        stage = model.stage(pressure, residual, phys, continuity_residual=cont)
        return Trial(model.state(state.t + dt, velocity), [stage])


class IMEXEulerMRSAV:
    name = "imex_euler_mrsav"

    def __init__(self, gamma: float = 1.0) -> None:
        if not np.isfinite(gamma) or gamma < 0:
            raise ValueError("gamma must be finite and nonnegative")
        self.gamma = gamma

    def step(self, model: Model, state: Any, dt: float) -> Trial:
        if not np.isfinite(dt) or dt <= 0 or not np.isfinite([state.t, state.r]).all():
            raise ValueError("Invalid step size or scalar state")
        mom = stage_momentum(model)
        old = model.vector(state)
        conv, lifting = mom.nonlinear_with_lifting(old)
        rhs = model.combine((1.0, mom.mass(old)), (dt, mom.force(state.t)))
        solved = mom.solve_columns([rhs, conv], mass=1.0, viscosity=model.nu * dt)
        V, W = solved[0].velocity, solved[1].velocity
        alpha = mom.pair(conv, V) + lifting
        beta = mom.pair(conv, W)
        if beta < 0.0 and beta < -1e-12 * (1 + mom.cauchy_scale(conv, W)):
            raise RuntimeError("Negative Stokes response energy")
        beta = max(beta, 0.0)
        d = 1 + self.gamma * dt
        upd = solve_scalar_stage(state.r, dt, d, state.r, alpha, beta, order=1)
        r, coupling = upd.r, upd.coupling
        velocity = model.combine((1.0, V), (-dt * coupling, W))
        pressure = model.combine((1 / dt, solved[0].pressure), (-coupling, solved[1].pressure))
        scalar = d * r - state.r + dt * (mom.pair(conv, velocity) + lifting)
        scalar_residual = abs(scalar) / (1 + abs(state.r) + abs(r))
        defect = model.combine((1.0, mom.mass(velocity)), (model.nu * dt, mom.linear(velocity)),
                               (dt, mom.gradient(pressure)), (-1.0, rhs), (dt * coupling, conv))
        residual = mom.residual(defect, rhs, conv, dt * abs(coupling))
        phys, cont = stage_divergences(model, velocity)
        if (not np.isfinite([r, residual, scalar_residual, phys, cont, model.max_abs(velocity),
                             model.max_abs(pressure)]).all()
                or residual > 1e-8 or scalar_residual > 1e-8 or cont > 1e-8 * (1 + model.max_abs(velocity))):
            raise RuntimeError(f"IMEX Euler SAV residuals {residual}, {scalar_residual}, divergence {cont}")
        stage = upd.record(model.stage(pressure, residual, phys, r, scalar_residual=scalar_residual,
                                       continuity_residual=cont))
        return Trial(model.state(state.t + dt, velocity, r), [stage])
