"""The coefficient tables that define every stage scheme, and the scheme metadata.

All schemes here are incremental IMEX diagonally-implicit schemes: stage i starts from
the previous stage's velocity U_i (U_0 the step's initial velocity) and solves

    (I + nu dt a_ii K) U_{i+1} + G q = U_i + dt sum_{j<=i} ah_ij (f_j - N_j)
                                       - nu dt sum_{j<i} a_ij K U_{j+1},     D U_{i+1} = 0,

with f_j = f(t + c_j dt) and N_j = N(U_j).  In the mr-SAV form the convection sum
sum_j ah_ij N_j is taken out of the right-hand side and becomes the stage's second
Stokes column, scaled by the scalar auxiliary variable's coupling.

The tables are the schemes' definitions and are the reference's:
  SDIRK2 (solver/schemes/sdirk2.py:14-43): eta = 1 - 1/sqrt 2, delta = -1/sqrt 2; the two
      stages of that file, written in the incremental form above, have
      a = [[eta, 0], [1 - 2 eta, eta]], ah = [[eta, 0], [-1, 1 - delta]], c = [0, eta],
      and its embedded first-order velocity (1 - 1/eta) U_0 + (1/eta) U_1.
  SDIRK3 (solver/schemes/sdirk3.py:8-19): the 4x4 tables D, DHAT and nodes C, and the
      embedded second-order velocity 2/3 U_0 - 128/99 U_1 + 50/33 U_2 + 1/9 U_3.
  IMEX Euler (solver/schemes/imex_euler.py, imex_euler_mrsav.py): one stage, a = ah = 1.
The metadata rows (order, stages, roots, estimator order, pressure semantics, scalar
corrections per stage) are solver/schemes/registry.py:14-33's.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass(frozen=True)
class SchemeInfo:
    order: int
    n_stages: int
    n_roots: int
    estimator_order: int | None
    pressure_semantics: str = "incremental_stage_multiplier_mean_zero"
    n_corrections: int = 0


@dataclass(frozen=True)
class Tableau:
    """implicit[i, j] = a_ij (lower triangular), explicit[i, j] = ah_ij (j <= i), nodes
    c_i; `embedded` holds the weights of the embedded velocity over [U_0, ..., U_s]
    (None: no embedded pair).  `sav_order` is the order p of the frozen factor
    P_{p-1}(z) = 1 + z + ... + z^{p-1} (0: no scalar variable), `corrections` the number
    of frozen-factor updates per stage.  `own_gate`: the plain stage forms its momentum
    defect and gates it itself instead of reporting the backend's residual."""
    implicit: np.ndarray
    explicit: np.ndarray
    nodes: np.ndarray
    embedded: np.ndarray | None
    sav_order: int = 0
    corrections: int = 0
    own_gate: bool = False

    @property
    def stages(self) -> int:
        return self.implicit.shape[0]


_ETA = 1.0 - 1.0 / np.sqrt(2.0)
_DELTA = -1.0 / np.sqrt(2.0)
_SDIRK2 = dict(implicit=np.array([[_ETA, 0.0], [1.0 - 2.0 * _ETA, _ETA]]),
               explicit=np.array([[_ETA, 0.0], [-1.0, 1.0 - _DELTA]]),
               nodes=np.array([0.0, _ETA]),
               embedded=np.array([1.0 - 1.0 / _ETA, 1.0 / _ETA, 0.0]))
_SDIRK3 = dict(implicit=np.array([[1 / 4, 0, 0, 0], [3 / 10, 1 / 4, 0, 0],
                                  [-31 / 220, 1 / 11, 1 / 4, 0], [31 / 198, 5 / 132, -4 / 9, 1 / 4]]),
               explicit=np.array([[1 / 4, 0, 0, 0], [-33 / 50, 121 / 100, 0, 0],
                                  [239 / 400, -71 / 100, 5 / 16, 0],
                                  [-1 / 48, -17 / 66, 235 / 528, -1 / 6]]),
               nodes=np.array([0.0, 1 / 4, 4 / 5, 1.0]),
               embedded=np.array([2 / 3, -128 / 99, 50 / 33, 1 / 9, 0.0]))
_EULER = dict(implicit=np.array([[1.0]]), explicit=np.array([[1.0]]), nodes=np.array([0.0]), embedded=None)

TABLEAUX: dict[str, Tableau] = {
    "sdirk2": Tableau(**_SDIRK2),
    "sdirk2_mrsav": Tableau(**_SDIRK2, sav_order=2, corrections=1),
    "sdirk3": Tableau(**_SDIRK3),
    "sdirk3_mrsav": Tableau(**_SDIRK3, sav_order=3, corrections=2),
    "imex_euler": Tableau(**_EULER, own_gate=True),
    "imex_euler_mrsav": Tableau(**_EULER, sav_order=1, corrections=1),
}

INFO: dict[str, SchemeInfo] = {
    "sdirk2": SchemeInfo(2, 2, 3, 2),
    "sdirk2_mrsav": SchemeInfo(2, 2, 3, 2, n_corrections=1),
    "sdirk3": SchemeInfo(3, 4, 5, 3),
    "sdirk3_mrsav": SchemeInfo(3, 4, 5, 3, n_corrections=2),
    "imex_euler": SchemeInfo(1, 1, 1, None, "euler_step_multiplier_mean_zero"),
    "imex_euler_mrsav": SchemeInfo(1, 1, 1, None, "euler_step_multiplier_mean_zero", n_corrections=1),
}
