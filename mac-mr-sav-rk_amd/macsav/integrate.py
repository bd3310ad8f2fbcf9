"""Import path of the reference's fixed-step driver (solver/integrate.py); the drivers
are in `drivers.py`."""
from .drivers import Result, integrate, step_sizes

__all__ = ["Result", "integrate", "step_sizes"]
