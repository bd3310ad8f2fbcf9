"""Time schemes: one table-driven stage engine, selected by name.

`make_scheme(name, gamma)`, `scheme_info(name)`, `SCHEMES` and `DEFAULT_SCHEMES` are the
reference's registry interface (solver/schemes/registry.py:24-53); the scheme classes
keep the reference's names for callers that construct them directly.
"""
from __future__ import annotations

from .engine import StageEngine
from .sav import SavUpdate, frozen_factor, sav_update
from .tableau import INFO, TABLEAUX, SchemeInfo, Tableau

SCHEMES = tuple(INFO)
DEFAULT_SCHEMES = ("sdirk2", "sdirk2_mrsav", "sdirk3", "sdirk3_mrsav")


def scheme_info(name: str) -> SchemeInfo:
    if name not in INFO:
        raise ValueError(f"Unknown scheme: {name}")
    return INFO[name]


def make_scheme(name: str, gamma: float = 1.0) -> StageEngine:
    scheme_info(name)
    return StageEngine(name, TABLEAUX[name], gamma)


def _named(name: str) -> type:
    def __init__(self: StageEngine, gamma: float = 1.0) -> None:
        StageEngine.__init__(self, name, TABLEAUX[name], gamma)
    return type(name.upper().replace("_MRSAV", "MRSAV").replace("IMEX_EULER", "IMEXEuler"),
                (StageEngine,), {"__init__": __init__})


SDIRK2, SDIRK2MRSAV, SDIRK3, SDIRK3MRSAV, IMEXEuler, IMEXEulerMRSAV = (
    _named(n) for n in ("sdirk2", "sdirk2_mrsav", "sdirk3", "sdirk3_mrsav", "imex_euler", "imex_euler_mrsav"))

__all__ = ["StageEngine", "SavUpdate", "frozen_factor", "sav_update", "SchemeInfo", "Tableau", "TABLEAUX",
           "SCHEMES", "DEFAULT_SCHEMES", "scheme_info", "make_scheme", "SDIRK2", "SDIRK2MRSAV", "SDIRK3",
           "SDIRK3MRSAV", "IMEXEuler", "IMEXEulerMRSAV"]
