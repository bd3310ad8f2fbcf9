"""Host-side MAC convection: skew-symmetric dual-cell centred fluxes.

Same discrete operator as the reference's Numba kernel (solver/mac/kernels.py:6-35):
for every interior face, centred transport velocities on the dual cell, the
conservative flux difference, minus half the dual-cell divergence times the local
velocity, with odd ghosts (-u) across the tangential walls.  The transport velocity
through a wall is its zero normal velocity, so the advective flux there vanishes for any
ghost value: a moving lid changes nothing here and enters only through `lid_load`.
Written here over whole
face arrays with explicit ghost rows/columns; the operation order per face is the
loop's, so results are bitwise those of the loop.
"""
from __future__ import annotations

import numpy as np

from .grid import Array


def convection(u: Array, v: Array, hx: float, hy: float) -> tuple[Array, Array]:
    ny, nx = u.shape[0], u.shape[1] - 1
    cu = np.zeros_like(u)
    cv = np.zeros_like(v)
    # u faces (j, i), i = 1..nx-1; ghost rows above/below are -u (odd extension).
    uc = u[:, 1:nx]
    ug = np.vstack([-uc[:1], uc, -uc[-1:]])          # rows j-1 .. j+1 with ghosts
    e = 0.5 * (uc + u[:, 2:])
    w = 0.5 * (uc + u[:, :nx - 1])
    n = 0.5 * (v[1:, :nx - 1] + v[1:, 1:nx])
    s = 0.5 * (v[:-1, :nx - 1] + v[:-1, 1:nx])
    flux = (e * 0.5 * (uc + u[:, 2:]) - w * 0.5 * (uc + u[:, :nx - 1])) / hx
    flux = flux + (n * 0.5 * (uc + ug[2:]) - s * 0.5 * (uc + ug[:-2])) / hy
    cu[:, 1:nx] = flux - 0.5 * uc * ((e - w) / hx + (n - s) / hy)
    # v faces (j, i), j = 1..ny-1; ghost columns left/right are -v.
    vc = v[1:ny, :]
    vg = np.hstack([-vc[:, :1], vc, -vc[:, -1:]])
    e = 0.5 * (u[:ny - 1, 1:] + u[1:ny, 1:])
    w = 0.5 * (u[:ny - 1, :nx] + u[1:ny, :nx])
    n = 0.5 * (vc + v[2:, :])
    s = 0.5 * (vc + v[:ny - 1, :])
    flux = (e * 0.5 * (vc + vg[:, 2:]) - w * 0.5 * (vc + vg[:, :-2])) / hx
    flux = flux + (n * 0.5 * (vc + v[2:, :]) - s * 0.5 * (vc + v[:ny - 1, :])) / hy
    cv[1:ny, :] = flux - 0.5 * vc * ((e - w) / hx + (n - s) / hy)
    return cu, cv


def packed_convection(grid, velocity: Array) -> Array:
    """N(U) on the packed free-face vector."""
    u, v = grid.unpack(velocity)
    cu, cv = convection(u, v, grid.hx, grid.hy)
    return grid.pack(cu, cv)


def lid_load(grid, lid: float) -> Array:
    """The viscous boundary load per unit viscosity of a top wall moving tangentially at
    `lid` (no counterpart in the reference's MAC model, which has no-slip walls only).
    The wall value lid is imposed by the ghost 2 lid - u above the top row of u faces,
    so there -Laplacian_ghost u = K u - 2 lid / hy^2: the load is 2 lid / hy^2 on those
    faces and zero elsewhere."""
    load = np.zeros(grid.size)
    if lid != 0.0:
        top = (grid.ny - 1) * (grid.nx - 1)
        load[top:top + grid.nx - 1] = 2.0 * lid / grid.hy ** 2
    return load
