"""The Stokes-backend plugin seam: result type and the protocols a backend satisfies.

Mirrors solver/mac/stokes.py:11-40 of the reference: a backend is any object with
`solve(rhs, *, mass, viscosity) -> StokesResult` solving
(mass I + viscosity K) u + G p = rhs, D u = 0, mean(p) = 0.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Any, Protocol

from .grid import Array


class StokesGrid(Protocol):
    @property
    def size(self) -> int: ...
    @property
    def np(self) -> int: ...
    @property
    def nx(self) -> int: ...
    @property
    def ny(self) -> int: ...


class StokesOperators(Protocol):
    @property
    def grid(self) -> StokesGrid: ...
    @property
    def K(self) -> Any: ...
    @property
    def G(self) -> Any: ...
    @property
    def D(self) -> Any: ...


@dataclass
class StokesResult:
    velocity: Array
    pressure: Array
    residual: float
    divergence_inf: float


class StokesBackend(Protocol):
    def solve(self, rhs: Array, *, mass: float, viscosity: float) -> StokesResult: ...
