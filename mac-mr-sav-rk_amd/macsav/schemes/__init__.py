"""Time schemes driving the stage solve (reference: solver/schemes/__init__.py)."""
from .imex_euler import IMEXEuler, IMEXEulerMRSAV
from .registry import DEFAULT_SCHEMES, SCHEMES, SchemeInfo, make_scheme, scheme_info

__all__ = ["IMEXEuler", "IMEXEulerMRSAV", "DEFAULT_SCHEMES", "SCHEMES", "SchemeInfo", "make_scheme",
           "scheme_info"]
