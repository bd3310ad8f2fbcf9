"""MI355X-native MAC shifted-Stokes stage solve behind the reference's Stokes-backend seam.

Host-side mirror of the reference API (MACGrid, MACOperators, MACNavierStokes,
make_scheme, integrate, integrate_adaptive, save_result) with `GpuTensorStokes` -- the
gfx950 HIP backend -- as the stage solver, `DeviceMACNavierStokes` keeping the fields in
HBM, and `GpuBatchedStokes` for many independent cavities.
"""
from .mac.grid import MACGrid
from .mac.batch_stokes import GpuBatchedStokes
from .mac.gpu_stokes import GpuTensorStokes
from .mac.operators import MACOperators
from .mac.stokes import StokesBackend, StokesResult
from .mac_ns import MACNavierStokes
from .mac_ns_device import DeviceMACNavierStokes
from .schemes import make_scheme
from .integrate import integrate
from .adaptivity import IController, PIController, integrate_adaptive
from .records import load_result, save_result

__all__ = ["MACGrid", "MACOperators", "GpuTensorStokes", "GpuBatchedStokes", "StokesBackend", "StokesResult",
           "MACNavierStokes", "DeviceMACNavierStokes", "make_scheme", "integrate", "IController", "PIController",
           "integrate_adaptive", "save_result", "load_result"]
