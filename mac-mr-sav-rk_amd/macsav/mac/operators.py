"""Sparse MAC operators on the packed free-face vector (host side, scipy CSR).

K is the block-diagonal positive vector Laplacian: 5-point, diagonal 3/h^2 on the
tangential half-cell wall rows, 2/h^2 otherwise; D maps face fluxes to cell
divergence; G = -D^T.  Same matrices as the reference's `MACOperators`
(solver/mac/operators.py:6-29); kept for API compatibility and for the host-side
scheme algebra, while the stage solve itself runs matrix-free on the GPU.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from .grid import MACGrid


def second_difference(n: int, spacing: float, half_cell_wall: bool) -> sp.csr_matrix:
    diag = np.full(n, 2.0)
    if half_cell_wall:
        diag[0] = diag[-1] = 3.0
    return sp.diags([-np.ones(n - 1), diag, -np.ones(n - 1)], [-1, 0, 1], format="csr") / spacing**2


class MACOperators:
    def __init__(self, grid: MACGrid) -> None:
        self.grid = grid
        nx, ny = grid.nx, grid.ny
        k_u = (sp.kron(sp.eye(ny), second_difference(nx - 1, grid.hx, False))
               + sp.kron(second_difference(ny, grid.hy, True), sp.eye(nx - 1)))
        k_v = (sp.kron(sp.eye(ny - 1), second_difference(nx, grid.hx, True))
               + sp.kron(second_difference(ny - 1, grid.hy, False), sp.eye(nx)))
        self.K = sp.block_diag([k_u, k_v], format="csr")
        jj, ii = np.divmod(np.arange(grid.np), nx)
        entries = (
            (ii < nx - 1, jj * (nx - 1) + ii, 1.0 / grid.hx),
            (ii > 0, jj * (nx - 1) + ii - 1, -1.0 / grid.hx),
            (jj < ny - 1, grid.nu + jj * nx + ii, 1.0 / grid.hy),
            (jj > 0, grid.nu + (jj - 1) * nx + ii, -1.0 / grid.hy),
        )
        rows = np.concatenate([np.flatnonzero(m) for m, _, _ in entries])
        cols = np.concatenate([c[m] for m, c, _ in entries])
        vals = np.concatenate([np.full(int(m.sum()), v) for m, _, v in entries])
        self.D = sp.csr_matrix((vals, (rows, cols)), shape=(grid.np, grid.size))
        self.G = -self.D.T.tocsr()
