"""Two-stage incremental IMEX-SDIRK2 (reference: solver/schemes/sdirk2.py:14-43).

eta = 1 - 1/sqrt(2), delta = -1/sqrt(2).  Stage 1 solves with R = U_n + eta dt f(t_n) -
eta dt N(U_n); stage 2 with R = U_1 - nu dt (1-2 eta) K U_1 - dt f(t_n) + dt(1-delta) f(t_n +
eta dt) + dt N(U_n) - dt(1-delta) N(U_1); both with mass 1 and viscosity nu eta dt.
"""
from __future__ import annotations

from typing import Any

import numpy as np

from ..core import Trial
from ..model import Model, stage_divergences
from ..momentum import stage_momentum

ETA = 1 - 1 / np.sqrt(2.0)
DELTA = -1 / np.sqrt(2.0)


class SDIRK2:
    name = "sdirk2"

    def step(self, model: Model, state: Any, dt: float) -> Trial:
        if not np.isfinite(dt) or dt <= 0:
            raise ValueError("Invalid step size")
        mom = stage_momentum(model)
        old = model.vector(state)
        n_old = mom.nonlinear(old)
        f0 = mom.force(state.t)
        visc = model.nu * ETA * dt
        first = mom.solve(model.combine((1.0, mom.mass(old)), (ETA * dt, f0), (-ETA * dt, n_old)),
                          mass=1.0, viscosity=visc)
        f1 = mom.force(state.t + ETA * dt)
        n1 = mom.nonlinear(first.velocity)
        rhs = model.combine((1.0, mom.mass(first.velocity)),
                            (-model.nu * dt * (1 - 2 * ETA), mom.linear(first.velocity)),
                            (-dt, f0), (dt * (1 - DELTA), f1), (dt, n_old), (-dt * (1 - DELTA), n1))
        second = mom.solve(rhs, mass=1.0, viscosity=visc)
        stages = []
        for solved in (first, second):
            phys, cont = stage_divergences(model, solved.velocity)
            stages.append(model.stage(model.combine((1.0 / dt, solved.pressure)), solved.residual, phys,
                                      continuity_residual=cont))
        embedded = model.combine((1.0 - 1.0 / ETA, old), (1.0 / ETA, first.velocity))
        return Trial(model.state(state.t + dt, second.velocity), stages, embedded)
