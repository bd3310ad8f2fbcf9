"""The MAC Navier-Stokes model: the model seam over a field-algebra object.

`MACNavierStokes(grid, nu, force, backend, cache_size, lid)` keeps the reference's
constructor and seam (solver/mac_ns.py:12-107, protocol solver/model.py:39-106): the
schemes reach the discretisation only through `combine`, `apply_K/G/D`, `nonlinear`,
`inner`, `max_abs`, `solve`, `solve_columns`, `stage` and `diagnostics`.  Here every
field operation is delegated to a `fields` object, one per storage:

  HostFields    NumPy arrays; K, D, G as the CSR matrices of `MACOperators`
  DeviceFields  torch device tensors and libmacstokes.so's kernels (mac_ns_device.py)

so the model itself holds no arithmetic and both storages share the seam, the
moving-lid load, the backend injection and the diagnostics.  The default backend is
the GPU stage solve `GpuTensorStokes` (the reference defaults to sparse LU); any object
with `solve(rhs, *, mass, viscosity)` can be injected through `backend=` or by
assigning `model.backend`.

The moving lid (no counterpart in the reference's MAC model, which has no-slip walls
only): a top wall sliding at `lid` enters as the viscous load nu * 2 lid / hy^2 on the
top row of u faces, added to the body force (`mac.kernels.lid_load`); convection and the
Stokes operator are unchanged.
"""
from __future__ import annotations

from collections.abc import Callable, Sequence
from typing import Any

import numpy as np

from .core import Stage, State
from .mac.grid import Array, MACGrid
from .mac.kernels import lid_load, packed_convection
from .mac.operators import MACOperators
from .mac.stokes import StokesResult


def angular_weights(grid: MACGrid) -> Array:
    """Packed weights w with <w, U> (plain dot) = the angular momentum about the centre,
    area * sum_cells ((x - xc) v_c - (y - yc) u_c), u_c and v_c the cell averages of the
    two faces that bound each cell: an interior u face in row j sits in two cells of
    that row, so it carries -area (y_j - yc); a v face in column i carries area (x_i - xc)."""
    yu = (np.arange(grid.ny) + 0.5) * grid.hy - 0.5 * grid.ly
    xv = (np.arange(grid.nx) + 0.5) * grid.hx - 0.5 * grid.lx
    wu = np.repeat(-grid.area * yu, grid.nx - 1)
    wv = np.tile(grid.area * xv, grid.ny - 1)
    return np.concatenate([wu, wv])


class HostFields:
    """Field algebra on NumPy arrays."""

    def __init__(self, grid: MACGrid, ops: MACOperators) -> None:
        self.grid, self.ops = grid, ops
        self._angular = angular_weights(grid)

    def field(self, x: Any) -> Array:
        return np.asarray(x, dtype=float)

    def lincomb(self, terms: Sequence[tuple[float, Any]]) -> Array:
        # Accumulated term by term, left to right, into a fresh array.
        out = np.zeros(np.shape(terms[0][1]))
        for c, f in terms:
            out += c * f
        return out

    def K(self, u: Array) -> Array:
        return self.ops.K @ u

    def G(self, p: Any) -> Array:
        return self.ops.G @ np.ravel(p)

    def D(self, u: Array) -> Array:
        return self.ops.D @ u

    def convection(self, u: Array) -> Array:
        return packed_convection(self.grid, u)

    def dot(self, a: Any, b: Any) -> float:
        return float(np.dot(np.ravel(a), np.ravel(b)))

    def maxabs(self, a: Any) -> float:
        return float(np.max(np.abs(a)))

    def finite(self, a: Any) -> bool:
        return bool(np.isfinite(a).all())

    def stack(self, columns: Sequence[Any]) -> Array:
        return np.column_stack([self.field(c) for c in columns])

    def split(self, out: StokesResult, k: int) -> list[StokesResult]:
        return [StokesResult(out.velocity[:, c], out.pressure[:, :, c], out.residual, out.divergence_inf)
                for c in range(k)]

    def unpack_state(self, velocity: Any) -> tuple[Any, Any]:
        return self.grid.unpack(self.field(velocity))

    def pack_state(self, state: State) -> Array:
        return self.grid.pack(state.u, state.v)

    def diagnostic_sums(self, u: Array, f: Array) -> tuple[float, float, float, float, float]:
        """<u,u>, <u,Ku>, max|Du|, <w_ang,u>, <f,u> as plain sums (no area factor on the
        dots except the angular one, whose weights carry it)."""
        return (self.dot(u, u), self.dot(u, self.K(u)), self.maxabs(self.D(u)), self.dot(self._angular, u),
                self.dot(f, u))


class MACNavierStokes:
    """The MAC model (see the module docstring)."""

    def __init__(self, grid: MACGrid, nu: float, force: Callable[[float], Any] | None = None,
                 backend: Any = None, cache_size: int = 4, lid: float = 0.0, *, fields: Any = None) -> None:
        if not np.isfinite(nu) or nu <= 0:
            raise ValueError("nu must be positive and finite")
        if not np.isfinite(lid):
            raise ValueError("lid must be finite")
        self.grid, self.nu, self.lid, self.cache_size = grid, nu, float(lid), cache_size
        self.ops = MACOperators(grid)
        self.fields = fields if fields is not None else HostFields(grid, self.ops)
        if backend is None:
            from .mac.gpu_stokes import GpuTensorStokes
            backend = GpuTensorStokes(self.ops)
        self.backend = backend
        self._body = force
        self._wall = self.fields.field(nu * lid_load(grid, self.lid)) if self.lid != 0.0 else None
        self._zero = None

    # The force seen by the schemes: body force plus the lid's wall load.
    def force(self, t: float) -> Any:
        if self._body is None:
            if self._wall is not None:
                return self._wall
            if self._zero is None:
                self._zero = self.fields.field(np.zeros(self.grid.size))
            return self._zero
        body = self.fields.field(self._body(t))
        return body if self._wall is None else self.fields.lincomb([(1.0, body), (1.0, self._wall)])

    # ------------------------------------------------------------ states
    def vector(self, state: State) -> Any:
        return self.fields.pack_state(state)

    def state(self, t: float, velocity: Any, r: float = 0.0) -> State:
        u, v = self.fields.unpack_state(velocity)
        return State(t, u, v, r)

    # -------------------------------------------------------------- seam
    def nonlinear(self, velocity: Any) -> Any:
        return self.fields.convection(self.fields.field(velocity))

    def nonlinear_with_lifting(self, velocity: Any) -> tuple[Any, float]:
        # No lifting pairing on the MAC box (walls at rest, or a lid carried by the force:
        # the wall's normal velocity is zero, so it adds no advective flux).
        return self.nonlinear(velocity), 0.0

    def combine(self, *terms: tuple[float, Any]) -> Any:
        return self.fields.lincomb(terms)

    def apply_K(self, velocity: Any) -> Any:
        return self.fields.K(self.fields.field(velocity))

    def apply_G(self, pressure: Any) -> Any:
        return self.fields.G(pressure)

    def apply_D(self, velocity: Any) -> Any:
        return self.fields.D(self.fields.field(velocity))

    def physical_divergence_inf(self, velocity: Any) -> float:
        return self.max_abs(self.apply_D(velocity))

    def stage_divergences(self, velocity: Any) -> tuple[float, float]:
        # One discrete divergence serves as both the physical and the algebraic one.
        d = self.physical_divergence_inf(velocity)
        return d, d

    def inner(self, left: Any, right: Any) -> float:
        return self.grid.area * self.fields.dot(left, right)

    def max_abs(self, field: Any) -> float:
        return self.fields.maxabs(field)

    def all_finite(self, field: Any) -> bool:
        return self.fields.finite(field)

    def solve(self, rhs: Any, *, mass: float, viscosity: float) -> Any:
        return self.backend.solve(self.fields.field(rhs), mass=mass, viscosity=viscosity)

    def solve_columns(self, columns: Sequence[Any], *, mass: float, viscosity: float) -> list[Any]:
        """All columns in one backend call (one operator, shared transforms)."""
        out = self.backend.solve(self.fields.stack(columns), mass=mass, viscosity=viscosity)
        return self.fields.split(out, len(columns))

    def stage(self, pressure: Any, residual: float, divergence_inf: float, r: float = 0.0,
              candidates: Sequence[float] = (), root_residuals: Sequence[float] = (),
              scalar_residual: float = 0.0, *, continuity_residual: float = float("nan")) -> Stage:
        return Stage(pressure, residual, divergence_inf, r, list(candidates), list(root_residuals),
                     scalar_residual, continuity_residual)

    # The drivers mark every trial's end: a backend that keeps solves for its warm start
    # (the same-stage history) keeps an accepted step's and forgets a rejected one's.
    def commit_trial(self) -> None:
        mark = getattr(self.backend, "commit_history", None)
        if mark is not None:
            mark()

    def discard_trial(self) -> None:
        mark = getattr(self.backend, "discard_history", None)
        if mark is not None:
            mark()

    # ------------------------------------------------------- diagnostics
    def angular_momentum(self, velocity: Any) -> float:
        return float(self.fields.dot(angular_weights(self.grid), velocity))

    def diagnostics(self, state: State) -> dict[str, float]:
        """kinetic 1/2|u|^2, modified energy (+ (r-1)^2/2), divergence max|Du|, H1
        seminorm <u, Ku>, r, angular momentum, power <f, u>, dissipation nu <u, Ku>
        (the reference's keys, solver/mac_ns.py:95-107)."""
        u = self.vector(state)
        uu, uku, div, ang, fu = self.fields.diagnostic_sums(u, self.force(state.t))
        a = self.grid.area
        kinetic, h1 = 0.5 * a * uu, a * uku
        return {"kinetic": kinetic, "modified_energy": kinetic + 0.5 * (state.r - 1.0) ** 2,
                "divergence_inf": div, "h1_seminorm_squared": h1, "r": state.r, "angular_momentum": ang,
                "power": a * fu, "dissipation": self.nu * h1}
