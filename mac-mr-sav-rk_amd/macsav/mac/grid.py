"""Uniform no-slip MAC grid: face arrays, packed interior velocities, inner product.

Same public surface as the reference's `MACGrid` (solver/mac/grid.py:8-71): u lives on
(ny, nx+1) faces, v on (ny+1, nx), pressure on (ny, nx) cells, first axis y; the packed
free vector is [u[:, 1:-1] row-major ; v[1:-1, :] row-major] and the velocity inner
product is hx*hy*dot(a, b).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
from numpy.typing import NDArray

Array = NDArray[np.float64]


@dataclass(frozen=True)
class MACGrid:
    nx: int
    ny: int
    lx: float = 1.0
    ly: float = 1.0

    def __post_init__(self) -> None:
        for n in (self.nx, self.ny):
            if isinstance(n, bool) or not isinstance(n, int):
                raise ValueError("nx, ny must be integers")
        if min(self.nx, self.ny) < 3 or not np.isfinite([self.lx, self.ly]).all() or min(self.lx, self.ly) <= 0:
            raise ValueError("Require nx,ny >= 3 and positive finite domain lengths")

    hx = property(lambda self: self.lx / self.nx)
    hy = property(lambda self: self.ly / self.ny)
    nu = property(lambda self: (self.nx - 1) * self.ny)
    nv = property(lambda self: self.nx * (self.ny - 1))
    size = property(lambda self: self.nu + self.nv)
    np = property(lambda self: self.nx * self.ny)
    area = property(lambda self: self.hx * self.hy)

    def coordinates(self, field: str) -> tuple[Array, Array]:
        """Mesh of the sample points of 'u', 'v', 'p' or 'vertex' (x, y arrays)."""
        cells_x = (np.arange(self.nx, dtype=float) + 0.5) * self.hx
        cells_y = (np.arange(self.ny, dtype=float) + 0.5) * self.hy
        nodes_x = np.arange(self.nx + 1, dtype=float) * self.hx
        nodes_y = np.arange(self.ny + 1, dtype=float) * self.hy
        axes = {"u": (nodes_x, cells_y), "v": (cells_x, nodes_y), "p": (cells_x, cells_y),
                "vertex": (nodes_x, nodes_y)}
        if field not in axes:
            raise ValueError(f"Unknown field {field}")
        return tuple(np.meshgrid(*axes[field]))  # type: ignore[return-value]

    def pack(self, u: Array, v: Array) -> Array:
        if u.shape != (self.ny, self.nx + 1) or v.shape != (self.ny + 1, self.nx):
            raise ValueError("Incorrect MAC array shapes")
        if np.any(u[:, [0, -1]] != 0) or np.any(v[[0, -1], :] != 0):
            raise ValueError("Normal boundary velocities must be zero")
        return np.concatenate((u[:, 1:-1].ravel(), v[1:-1, :].ravel()))

    def unpack(self, velocity: Array) -> tuple[Array, Array]:
        if velocity.shape != (self.size,):
            raise ValueError("Incorrect free-vector shape")
        u = np.zeros((self.ny, self.nx + 1))
        v = np.zeros((self.ny + 1, self.nx))
        u[:, 1:-1] = velocity[: self.nu].reshape(self.ny, self.nx - 1)
        v[1:-1, :] = velocity[self.nu:].reshape(self.ny - 1, self.nx)
        return u, v

    def inner(self, a: Array, b: Array) -> float:
        return float(self.area * np.dot(a, b))

    def norm(self, a: Array) -> float:
        return float(np.sqrt(self.inner(a, a)))
