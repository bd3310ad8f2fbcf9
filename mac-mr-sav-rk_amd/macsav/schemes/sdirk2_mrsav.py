"""SDIRK2 mr-ccSAV: one frozen-factor scalar correction per stage, two Stokes right-hand
sides per stage sharing one operator (reference: solver/schemes/sdirk2_mrsav.py:14-75)."""
from __future__ import annotations

from typing import Any

import numpy as np

from ..core import Trial
from ..model import Model, stage_divergences
from ..momentum import stage_momentum
from .scalar_correction import solve_scalar_stage
from .sdirk2 import DELTA, ETA


class SDIRK2MRSAV:
    name = "sdirk2_mrsav"

    def __init__(self, gamma: float = 1.0) -> None:
        if not np.isfinite(gamma) or gamma < 0:
            raise ValueError("gamma must be finite and nonnegative")
        self.gamma = gamma

    def _stage(self, model: Model, rhs: Any, scalar_rhs: float, convection: Any, dt: float,
               predictor: float, lifting_work: float = 0.0) -> tuple[Any, Any, Any]:
        mom = stage_momentum(model)
        solved = mom.solve_columns([rhs, convection], mass=1.0, viscosity=model.nu * ETA * dt)
        V, W = solved[0].velocity, solved[1].velocity
        alpha = mom.pair(convection, V) + lifting_work
        beta = mom.pair(convection, W)
        if beta < 0.0 and beta < -1e-12 * (1 + mom.cauchy_scale(convection, W)):
            raise RuntimeError("Negative Stokes response energy")
        beta = max(beta, 0.0)
        d = 1 + self.gamma * ETA * dt
        upd = solve_scalar_stage(predictor, dt, d, scalar_rhs, alpha, beta, order=2)
        r, coupling = upd.r, upd.coupling
        velocity = model.combine((1.0, V), (-dt * coupling, W))
        pressure = model.combine((1.0 / dt, solved[0].pressure), (-coupling, solved[1].pressure))
        scalar = d * r - scalar_rhs + dt * upd.factor * (mom.pair(convection, velocity) + lifting_work)
        scalar_residual = abs(scalar) / (1 + abs(scalar_rhs) + abs(r))
        linear_velocity = mom.linear(velocity)
        defect = model.combine((1.0, mom.mass(velocity)), (model.nu * ETA * dt, linear_velocity),
                               (dt, mom.gradient(pressure)), (-1.0, rhs), (dt * coupling, convection))
        residual = mom.residual(defect, rhs, convection, dt * abs(coupling))
        phys, divergence = stage_divergences(model, velocity)
        if (not np.isfinite([residual, scalar_residual, divergence]).all() or scalar_residual > 1e-8
                or residual > 1e-8 or divergence > 1e-8 * (1 + model.max_abs(velocity))):
            raise RuntimeError(f"SAV stage residuals {residual}, {scalar_residual}, divergence {divergence}")
        stage = upd.record(model.stage(pressure, residual, phys, r, scalar_residual=scalar_residual,
                                       continuity_residual=divergence))
        return velocity, stage, linear_velocity

    def step(self, model: Model, state: Any, dt: float) -> Trial:
        if not np.isfinite(dt) or dt <= 0:
            raise ValueError("Invalid step size")
        mom = stage_momentum(model)
        old = model.vector(state)
        n_old, lift_old = mom.nonlinear_with_lifting(old)
        f0 = mom.force(state.t)
        first, s1, k1 = self._stage(model, model.combine((1.0, mom.mass(old)), (ETA * dt, f0)), state.r,
                                    model.combine((ETA, n_old)), dt, state.r, ETA * lift_old)
        n1, lift1 = mom.nonlinear_with_lifting(first)
        f1 = mom.force(state.t + ETA * dt)
        rhs = model.combine((1.0, mom.mass(first)), (-model.nu * dt * (1 - 2 * ETA), k1), (-dt, f0),
                            (dt * (1 - DELTA), f1))
        second, s2, _ = self._stage(model, rhs, (1 - self.gamma * dt * (1 - 2 * ETA)) * s1.r,
                                    model.combine((-1.0, n_old), (1 - DELTA, n1)), dt, s1.r,
                                    -lift_old + (1 - DELTA) * lift1)
        embedded = model.combine((1.0 - 1.0 / ETA, old), (1.0 / ETA, first))
        return Trial(model.state(state.t + dt, second, s2.r), [s1, s2], embedded)
