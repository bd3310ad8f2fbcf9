"""`MACNavierStokes`: the MAC model the schemes drive, with an injectable Stokes backend.

Same seam members and semantics as the reference's solver/mac_ns.py:12-107; the
default backend is the GPU stage solve `GpuTensorStokes` (the reference defaults to a
sparse-direct solve).  Any object with `solve(rhs, *, mass, viscosity)` can be
injected through `backend=` or by assigning `model.backend`.
"""
from __future__ import annotations

from collections.abc import Callable, Sequence
from typing import Any

import numpy as np

from .core import Stage, State
from .mac.grid import Array, MACGrid
from .mac.kernels import convection, lid_load
from .mac.operators import MACOperators
from .mac.stokes import StokesBackend, StokesResult
from .model import SolveResult


class MACNavierStokes:
    def __init__(self, grid: MACGrid, nu: float, force: Callable[[float], Array] | None = None,
                 backend: StokesBackend | None = None, cache_size: int = 4, lid: float = 0.0) -> None:
        if not np.isfinite(nu) or nu <= 0:
            raise ValueError("nu must be positive and finite")
        if not np.isfinite(lid):
            raise ValueError("lid must be finite")
        self.grid = grid
        self.nu = nu
        self.lid = float(lid)
        self.ops = MACOperators(grid)
        if backend is None:
            from .mac.gpu_stokes import GpuTensorStokes
            backend = GpuTensorStokes(self.ops)
        self.backend = backend
        self.cache_size = cache_size
        body = force if force is not None else (lambda t: np.zeros(grid.size))
        if self.lid != 0.0:
            # A moving top wall enters as its viscous boundary load, nu * 2 lid / hy^2 on
            # the top row of u faces, added to the body force.
            wall = nu * lid_load(grid, self.lid)
            self.force = lambda t: body(t) + wall
        else:
            self.force = body

    # ------------------------------------------------------------- fields
    def nonlinear(self, velocity: Array) -> Array:
        u, v = self.grid.unpack(velocity)
        cu, cv = convection(u, v, self.grid.hx, self.grid.hy)
        return self.grid.pack(cu, cv)

    def nonlinear_with_lifting(self, velocity: Any) -> tuple[Any, float]:
        # No boundary lifting pairing on the MAC box: zero pairing, as the reference's MAC
        # model.  A moving lid adds no advective flux (the wall's normal velocity is zero);
        # its viscous load is carried by the force.
        return self.nonlinear(velocity), 0.0

    def vector(self, state: State) -> Array:
        return self.grid.pack(state.u, state.v)

    def state(self, t: float, velocity: Array, r: float = 0.0) -> State:
        u, v = self.grid.unpack(np.asarray(velocity))
        return State(t, u, v, r)

    # ---------------------------------------------------------------- seam
    def combine(self, *terms: tuple[float, Array]) -> Array:
        """sum of coefficient*field, accumulated left to right into a fresh array."""
        total = np.zeros_like(terms[0][1], dtype=float)
        for coefficient, field in terms:
            total += coefficient * field
        return total

    def apply_K(self, velocity: Array) -> Array:
        return self.ops.K @ velocity

    def apply_G(self, pressure: Array) -> Array:
        return self.ops.G @ np.asarray(pressure).reshape(-1)

    def apply_D(self, velocity: Array) -> Array:
        return self.ops.D @ velocity

    def physical_divergence_inf(self, velocity: Array) -> float:
        return self.max_abs(self.apply_D(velocity))

    def stage_divergences(self, velocity: Array) -> tuple[float, float]:
        d = self.max_abs(self.apply_D(velocity))
        return d, d

    def inner(self, left: Array, right: Array) -> float:
        return self.grid.inner(left, right)

    def max_abs(self, field: Array) -> float:
        return float(np.max(np.abs(field)))

    def solve(self, rhs: Array, *, mass: float, viscosity: float) -> StokesResult:
        return self.backend.solve(rhs, mass=mass, viscosity=viscosity)

    def solve_columns(self, columns: Sequence[Any], *, mass: float, viscosity: float) -> list[SolveResult]:
        """All columns in one backend call (one shared operator), split afterwards."""
        stacked = np.column_stack([np.asarray(c, dtype=float) for c in columns])
        out = self.backend.solve(stacked, mass=mass, viscosity=viscosity)
        return [StokesResult(out.velocity[:, k], out.pressure[:, :, k], out.residual, out.divergence_inf)
                for k in range(len(columns))]

    def stage(self, pressure: Array, residual: float, divergence_inf: float, r: float = 0.0,
              candidates: Sequence[float] = (), root_residuals: Sequence[float] = (),
              scalar_residual: float = 0.0, *, continuity_residual: float = float("nan")) -> Stage:
        return Stage(pressure, residual, divergence_inf, r, list(candidates), list(root_residuals),
                     scalar_residual, continuity_residual)

    # --------------------------------------------------------- diagnostics
    def collocated_velocity(self, velocity: Array) -> tuple[Array, Array]:
        u, v = self.grid.unpack(velocity)
        return 0.5 * (u[:, :-1] + u[:, 1:]), 0.5 * (v[:-1, :] + v[1:, :])

    def angular_momentum(self, velocity: Array) -> float:
        cu, cv = self.collocated_velocity(velocity)
        x, y = self.grid.coordinates("p")
        return float(self.grid.area * np.sum((x - 0.5 * self.grid.lx) * cv - (y - 0.5 * self.grid.ly) * cu))

    def diagnostics(self, state: State) -> dict[str, float]:
        velocity = self.vector(state)
        kinetic = 0.5 * self.grid.inner(velocity, velocity)
        h1 = self.grid.inner(velocity, self.ops.K @ velocity)
        return {"kinetic": kinetic, "modified_energy": kinetic + 0.5 * (state.r - 1) ** 2,
                "divergence_inf": float(np.max(np.abs(self.ops.D @ velocity))),
                "h1_seminorm_squared": h1, "r": state.r,
                "angular_momentum": self.angular_momentum(velocity),
                "power": self.grid.inner(self.force(state.t), velocity),
                "dissipation": self.nu * h1}
