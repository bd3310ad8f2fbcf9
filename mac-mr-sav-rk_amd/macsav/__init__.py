"""MI355X-native MAC shifted-Stokes stage solve behind the reference's Stokes-backend seam.

Host-side mirror of the reference API (MACGrid, MACOperators, MACNavierStokes,
make_scheme, integrate) with `GpuTensorStokes` -- the gfx950 HIP backend -- as the
stage solver.
"""
from .mac.grid import MACGrid
from .mac.batch_stokes import GpuBatchedStokes
from .mac.gpu_stokes import GpuTensorStokes
from .mac.operators import MACOperators
from .mac.stokes import StokesBackend, StokesResult

__all__ = ["MACGrid", "MACOperators", "GpuTensorStokes", "GpuBatchedStokes", "StokesBackend", "StokesResult"]
