"""Seeded, exactly divergence-free initial velocities for the stage-solve workloads.

psi at the (ny+1, nx+1) vertices is
    sin^2(pi x/lx) sin^2(pi y/ly) * sum_{k,l=1..modes} a_kl sin(k pi x/lx) sin(l pi y/ly),
with a_kl ~ N(0, std=(k^2+l^2)^-1) drawn from numpy.random.default_rng(seed) as one
(modes, modes) array (k the first axis).  Velocities are the vertex curl
u = diff(psi, axis=0)/hy on (ny, nx+1) and v = -diff(psi, axis=1)/hx on (ny+1, nx)
(the construction of the reference's tests/references/mac.py:14-21), which the
discrete D maps to zero exactly; the field is scaled so max(|u|, |v|) = amplitude.
"""
from __future__ import annotations

import numpy as np

from .mac.grid import Array, MACGrid


def stream_function_velocity(grid: MACGrid, seed: int, modes: int = 8, amplitude: float = 1.0) -> Array:
    x = np.arange(grid.nx + 1) * grid.hx / grid.lx
    y = np.arange(grid.ny + 1) * grid.hy / grid.ly
    k = np.arange(1, modes + 1)
    coeff = np.random.default_rng(seed).normal(size=(modes, modes)) / (k[:, None] ** 2 + k[None, :] ** 2)
    sx = np.sin(np.pi * np.outer(k, x))          # (modes, nx+1), index k
    sy = np.sin(np.pi * np.outer(k, y))          # (modes, ny+1), index l
    series = sy.T @ coeff.T @ sx                  # psi[j, i] = sum_kl a_kl sin(k pi x_i) sin(l pi y_j)
    psi = np.outer(np.sin(np.pi * y) ** 2, np.sin(np.pi * x) ** 2) * series
    psi[[0, -1], :] = 0.0
    psi[:, [0, -1]] = 0.0
    u = np.diff(psi, axis=0) / grid.hy
    v = -np.diff(psi, axis=1) / grid.hx
    scale = max(np.abs(u).max(), np.abs(v).max())
    if scale > 0:
        u, v = u * (amplitude / scale), v * (amplitude / scale)
    return grid.pack(u, v)
