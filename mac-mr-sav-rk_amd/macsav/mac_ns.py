"""Import path of the reference's MAC model (solver/mac_ns.py); the model is in
`mac_model.py`."""
from .mac_model import HostFields, MACNavierStokes, angular_weights

__all__ = ["HostFields", "MACNavierStokes", "angular_weights"]
