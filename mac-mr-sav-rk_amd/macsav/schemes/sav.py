"""The scalar-auxiliary-variable update of one mr-SAV stage.

With V = L^-1 R and W = L^-1 B the two Stokes solutions of a stage (R the momentum
right-hand side, B the explicit convection load, L the stage's Stokes operator), the
stage couples the velocity and the scalar r through the factor q = P_{p-1}(z):

    L U - tau q B r = R - tau q B            =>  U = V - tau q (1 - r) W
    tau q <B, U> + d r = s                    (s the scalar right-hand side)

Eliminating U gives the scalar alone,
    r = (s - tau q alpha + tau^2 q^2 beta) / (d + tau^2 q^2 beta),
    alpha = <B, V>, beta = <B, W> >= 0.
q depends on the scalar through z, so the update is a fixed number of substitutions
z -> r(z) starting from the predictor, with q evaluated at the incoming z
(reference: solver/schemes/scalar_correction.py:8-52, docs/numerics.md:51-79): one
substitution for orders 1 and 2, two for order 3, the factor of the last one kept as
the frozen factor.  No root finding.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass(frozen=True)
class SavUpdate:
    predictor: float
    values: tuple[float, ...]   # the scalar after each substitution; the last is r
    factor: float               # q at the last substitution's input

    @property
    def r(self) -> float:
        return self.values[-1]

    @property
    def coupling(self) -> float:
        """q (1 - r): the weight of W in U = V - tau q (1 - r) W."""
        return self.factor * (1.0 - self.r)


def frozen_factor(order: int, z: float) -> float:
    """P_{order-1}(z) = 1 + z + ... + z^(order-1), by Horner."""
    q = 1.0
    for _ in range(order - 1):
        q = 1.0 + z * q
    return q


def sav_update(predictor: float, tau: float, d: float, s: float, alpha: float, beta: float, *,
               order: int, count: int) -> SavUpdate:
    if order not in (1, 2, 3) or count < 1:
        raise ValueError("SAV order must be 1, 2 or 3 with at least one substitution")
    if not np.isfinite([predictor, tau, d, s, alpha, beta]).all() or tau <= 0 or d <= 0 or beta < 0:
        raise ValueError("Invalid frozen SAV scalar inputs")
    z, q, seq = float(predictor), 1.0, []
    for _ in range(count):
        q = frozen_factor(order, z)
        response = tau * tau * q * q * beta
        denominator = d + response
        if not np.isfinite(denominator) or denominator <= 0:
            raise RuntimeError("Invalid frozen SAV scalar denominator")
        z = (s - tau * q * alpha + response) / denominator
        if not np.isfinite(z):
            raise RuntimeError("Nonfinite frozen SAV correction")
        seq.append(float(z))
    return SavUpdate(float(predictor), tuple(seq), float(q))
