"""Import path of the reference's adaptive driver (solver/adaptivity); the drivers are
in `drivers.py`."""
from .drivers import (AdaptiveContinuation, AdaptiveResult, IController, PIController, ScalarErrorControl,
                      embedded_error, integrate_adaptive)

__all__ = ["AdaptiveContinuation", "AdaptiveResult", "IController", "PIController", "ScalarErrorControl",
           "embedded_error", "integrate_adaptive"]
