"""Four-stage, third-order incremental IMEX-SDIRK (reference: solver/schemes/sdirk3.py:8-57).

The Butcher-type tables D (implicit), DHAT (explicit) and nodes C are the scheme's
definition and are kept identical; the embedded second-order velocity is
2/3 U_n - 128/99 U_1 + 50/33 U_2 + 1/9 U_3.
"""
from __future__ import annotations

from typing import Any

import numpy as np

from ..core import Trial
from ..model import Model, stage_divergences
from ..momentum import stage_momentum

D = np.array([[1 / 4, 0, 0, 0], [3 / 10, 1 / 4, 0, 0],
              [-31 / 220, 1 / 11, 1 / 4, 0], [31 / 198, 5 / 132, -4 / 9, 1 / 4]])
DHAT = np.array([[1 / 4, 0, 0, 0], [-33 / 50, 121 / 100, 0, 0],
                 [239 / 400, -71 / 100, 5 / 16, 0],
                 [-1 / 48, -17 / 66, 235 / 528, -1 / 6]])
C = np.array([0.0, 1 / 4, 4 / 5, 1.0, 1.0])


def embedded_velocity(model: Model, velocities: list[Any]) -> Any:
    return model.combine((2 / 3, velocities[0]), (-128 / 99, velocities[1]),
                         (50 / 33, velocities[2]), (1 / 9, velocities[3]))


def known_terms(model: Model, nonlinear: list[tuple[Any, float]], forces: list[Any], i: int
                ) -> tuple[Any, Any, float]:
    """Explicit convection, force and lifting pairing of stage i (0-based), DHAT-weighted."""
    conv = model.combine(*[(DHAT[i, j], nonlinear[j][0]) for j in range(i + 1)])
    lifting = float(sum(DHAT[i, j] * nonlinear[j][1] for j in range(i + 1)))
    force = model.combine(*[(DHAT[i, j], forces[j]) for j in range(i + 1)])
    return conv, force, lifting


class SDIRK3:
    name = "sdirk3"

    def step(self, model: Model, state: Any, dt: float) -> Trial:
        if not np.isfinite(dt) or dt <= 0:
            raise ValueError("Invalid step size")
        mom = stage_momentum(model)
        velocities = [model.vector(state)]
        linear: list[Any] = []
        stages = []
        nonlinear = [mom.nonlinear_with_lifting(velocities[0])]
        forces = [mom.force(state.t + C[j] * dt) for j in range(4)]
        for i in range(4):
            conv, force, _ = known_terms(model, nonlinear, forces, i)
            rhs = model.combine((1.0, mom.mass(velocities[-1])), (dt, force), (-dt, conv),
                                *[(-model.nu * dt * D[i, j], linear[j]) for j in range(i)])
            solved = mom.solve(rhs, mass=1.0, viscosity=model.nu * dt * D[i, i])
            velocities.append(solved.velocity)
            if i < 3:
                linear.append(mom.linear(solved.velocity))
                nonlinear.append(mom.nonlinear_with_lifting(solved.velocity))
            phys, cont = stage_divergences(model, solved.velocity)
            stages.append(model.stage(model.combine((1 / dt, solved.pressure)), solved.residual, phys,
                                      continuity_residual=cont))
        return Trial(model.state(state.t + dt, velocities[-1]), stages, embedded_velocity(model, velocities))
