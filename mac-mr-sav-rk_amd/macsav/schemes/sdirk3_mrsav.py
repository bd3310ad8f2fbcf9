"""SDIRK3 mr-ccSAV: two frozen scalar corrections per stage, two Stokes right-hand sides
per stage sharing one operator (reference: solver/schemes/sdirk3_mrsav.py:11-65)."""
from __future__ import annotations

from typing import Any

import numpy as np

from ..core import Trial
from ..model import Model, stage_divergences
from ..momentum import stage_momentum
from .scalar_correction import solve_scalar_stage
from .sdirk3 import C, D, embedded_velocity, known_terms


class SDIRK3MRSAV:
    name = "sdirk3_mrsav"

    def __init__(self, gamma: float = 1.0) -> None:
        if not np.isfinite(gamma) or gamma < 0:
            raise ValueError("gamma must be finite and nonnegative")
        self.gamma = gamma

    def step(self, model: Model, state: Any, dt: float) -> Trial:
        if not np.isfinite(dt) or dt <= 0:
            raise ValueError("Invalid step size")
        mom = stage_momentum(model)
        velocities = [model.vector(state)]
        linear: list[Any] = []
        scalars = [state.r]
        stages = []
        nonlinear = [mom.nonlinear_with_lifting(velocities[0])]
        forces = [mom.force(state.t + C[j] * dt) for j in range(4)]
        for i in range(4):
            conv, force, lifting = known_terms(model, nonlinear, forces, i)
            rhs = model.combine((1.0, mom.mass(velocities[-1])), (dt, force),
                                *[(-model.nu * dt * D[i, j], linear[j]) for j in range(i)])
            scalar_rhs = scalars[-1] - self.gamma * dt * sum(D[i, j] * scalars[j + 1] for j in range(i))
            solved = mom.solve_columns([rhs, conv], mass=1.0, viscosity=model.nu * dt * D[i, i])
            V, W = solved[0].velocity, solved[1].velocity
            alpha = mom.pair(conv, V) + lifting
            beta = mom.pair(conv, W)
            if beta < 0.0 and beta < -1e-12 * (1 + mom.cauchy_scale(conv, W)):
                raise RuntimeError("Negative Stokes response energy")
            beta = max(beta, 0.0)
            d = 1 + self.gamma * dt * D[i, i]
            upd = solve_scalar_stage(scalars[-1], dt, d, scalar_rhs, alpha, beta, order=3)
            r, coupling = upd.r, upd.coupling
            velocity = model.combine((1.0, V), (-dt * coupling, W))
            pressure = model.combine((1 / dt, solved[0].pressure), (-coupling, solved[1].pressure))
            scalar = d * r - scalar_rhs + dt * upd.factor * (mom.pair(conv, velocity) + lifting)
            scalar_residual = abs(scalar) / (1 + abs(scalar_rhs) + abs(r))
            linear_velocity = mom.linear(velocity)
            defect = model.combine((1.0, mom.mass(velocity)), (model.nu * dt * D[i, i], linear_velocity),
                                   (dt, mom.gradient(pressure)), (-1.0, rhs), (dt * coupling, conv))
            residual = mom.residual(defect, rhs, conv, dt * abs(coupling))
            phys, cont = stage_divergences(model, velocity)
            if (not np.isfinite([residual, scalar_residual, cont]).all() or residual > 1e-8
                    or scalar_residual > 1e-8 or cont > 1e-8 * (1 + model.max_abs(velocity))):
                raise RuntimeError(f"SAV stage residuals {residual}, {scalar_residual}, divergence {cont}")
            velocities.append(velocity)
            linear.append(linear_velocity)
            if i < 3:
                nonlinear.append(mom.nonlinear_with_lifting(velocity))
            scalars.append(r)
            stages.append(upd.record(model.stage(pressure, residual, phys, r, scalar_residual=scalar_residual,
                                                 continuity_residual=cont)))
        return Trial(model.state(state.t + dt, velocities[-1], scalars[-1]), stages,
                     embedded_velocity(model, velocities))
