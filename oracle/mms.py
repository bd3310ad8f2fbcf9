"""Manufactured no-slip fields for driver and scheme tests (test oracle).

TEST INFRASTRUCTURE ONLY.  Restates the reference's test-owned mathematical fields
(tests/references/mac.py:14-53): the vertex-curl initial velocity of
psi = a sin^2(pi x/lx) sin^2(pi y/ly) e^-t, and the body force that makes
u = curl psi, p = a cos(pi x/lx) cos(pi y/ly) e^-t an exact solution of the
incompressible Navier-Stokes equations with viscosity nu (time derivative, viscous
term, pressure gradient and convection (u.grad)u, derived symbolically with sympy).
"""
from __future__ import annotations

from functools import lru_cache

import numpy as np
import sympy as sy


def initial_velocity(grid, amplitude: float = 0.1, t: float = 0.0) -> np.ndarray:
    x, y = grid.coordinates("vertex")
    psi = amplitude * np.sin(np.pi * x / grid.lx) ** 2 * np.sin(np.pi * y / grid.ly) ** 2 * np.exp(-t)
    psi[[0, -1], :] = 0.0
    psi[:, [0, -1]] = 0.0
    return grid.pack(np.diff(psi, axis=0) / grid.hy, -np.diff(psi, axis=1) / grid.hx)


@lru_cache(maxsize=16)
def _force_terms(lx: float, ly: float, nu: float):
    x, y, t = sy.symbols("x y t")
    psi = sy.sin(sy.pi * x / lx) ** 2 * sy.sin(sy.pi * y / ly) ** 2 * sy.exp(-t)
    u, v = sy.diff(psi, y), -sy.diff(psi, x)
    p = sy.cos(sy.pi * x / lx) * sy.cos(sy.pi * y / ly) * sy.exp(-t)
    out = []
    for q, axis in ((u, x), (v, y)):
        linear = sy.diff(q, t) - nu * (sy.diff(q, x, 2) + sy.diff(q, y, 2)) + sy.diff(p, axis)
        advect = u * sy.diff(q, x) + v * sy.diff(q, y)
        out.append((sy.lambdify((x, y, t), linear, "numpy"), sy.lambdify((x, y, t), advect, "numpy")))
    return out


def forcing(grid, nu: float, amplitude: float, t: float) -> np.ndarray:
    """f = a (u_t - nu Lap u + grad p) + a^2 (u.grad)u on the interior faces."""
    (lu, au), (lv, av) = _force_terms(grid.lx, grid.ly, nu)
    xu, yu = grid.coordinates("u")
    xv, yv = grid.coordinates("v")
    fu = amplitude * lu(xu, yu, t) + amplitude ** 2 * au(xu, yu, t)
    fv = amplitude * lv(xv, yv, t) + amplitude ** 2 * av(xv, yv, t)
    fu[:, [0, -1]] = 0.0
    fv[[0, -1], :] = 0.0
    return grid.pack(fu, fv)
