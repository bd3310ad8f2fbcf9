"""`DeviceMACNavierStokes`: the MAC model with its fields resident on the GPU.

Same seam members and semantics as `MACNavierStokes` (the reference's
solver/mac_ns.py:12-107), so the unchanged schemes and drivers run it.  Velocities and
pressures are torch device tensors, every field operation between stage solves runs in
libmacstokes.so's model kernels, and the stage solves hand the backend device tensors
(zero-copy).  The schemes' scalar logic (SAV corrections, gates) stays on the host as
FP64 Python floats, as in the reference.

- `nonlinear`: the skew-symmetric convection (bitwise the reference loop);
- `combine`: left-to-right sum of coefficient * field (bitwise the reference);
- `apply_K`, `apply_G`, `apply_D`: the reference's CSR products, matrix-free;
- `inner`, `max_abs`: FP64 fixed-order reductions (hx hy sum a b, max |a|).
"""
from __future__ import annotations

import ctypes
from collections.abc import Callable, Sequence
from typing import Any

import numpy as np

from . import _native
from .core import State
from .mac.grid import MACGrid
from .mac.stokes import StokesResult
from .mac_ns import MACNavierStokes
from .model import SolveResult


class DeviceMACNavierStokes(MACNavierStokes):
    def __init__(self, grid: MACGrid, nu: float, force: Callable[[float], Any] | None = None,
                 backend: Any = None, device: int = 0, lid: float = 0.0) -> None:
        import torch
        from .mac.kernels import lid_load
        if backend is None:
            from .mac.gpu_stokes import GpuTensorStokes
            backend = GpuTensorStokes(grid, device=device)
        super().__init__(grid, nu, force=None, backend=backend, lid=lid)
        self.device = torch.device(f"cuda:{device}")
        self._lib = _native.load()
        wall = torch.as_tensor(nu * lid_load(grid, self.lid), device=self.device)
        if force is None:
            self.force = lambda t: wall
        else:
            self.force = lambda t: self.combine((1.0, force(t)), (1.0, wall))

    # ----------------------------------------------------------------- plumbing
    def _check(self, code: int) -> None:
        _native.raise_for(code, self._lib.mac_model_last_error().decode())

    def _dev(self, x: Any) -> Any:
        import torch
        return torch.as_tensor(x, dtype=torch.float64, device=self.device).contiguous()

    def _empty(self, n: int) -> Any:
        import torch
        return torch.empty(n, dtype=torch.float64, device=self.device)

    def _sync_torch(self) -> None:
        import torch
        torch.cuda.current_stream(self.device).synchronize()

    # ------------------------------------------------------------------ fields
    def nonlinear(self, velocity: Any) -> Any:
        g = self.grid
        vel = self._dev(velocity)
        out = self._empty(g.size)
        self._check(self._lib.mac_convection(g.nx, g.ny, float(g.lx), float(g.ly), vel.data_ptr(), g.size, 1,
                                             out.data_ptr(), g.size))
        return out

    def vector(self, state: State) -> Any:
        return state.u

    def state(self, t: float, velocity: Any, r: float = 0.0) -> State:
        return State(t, self._dev(velocity).reshape(-1), None, r)

    def host_velocity(self, state: State) -> np.ndarray:
        return state.u.cpu().numpy()

    # -------------------------------------------------------------------- seam
    def combine(self, *terms: tuple[float, Any]) -> Any:
        if len(terms) > 8:
            head = self.combine(*terms[:8])
            return self.combine((1.0, head), *terms[8:])
        fields = [self._dev(f) for _, f in terms]
        coeff = (ctypes.c_double * len(terms))(*[float(c) for c, _ in terms])
        ptrs = (ctypes.c_void_p * len(terms))(*[f.data_ptr() for f in fields])
        out = self._empty(fields[0].numel())
        self._check(self._lib.mac_combine(out.numel(), len(terms), coeff, ptrs, out.data_ptr()))
        return out

    def apply_K(self, velocity: Any) -> Any:
        g = self.grid
        vel = self._dev(velocity)
        out = self._empty(g.size)
        self._check(self._lib.mac_apply_K(g.nx, g.ny, float(g.lx), float(g.ly), vel.data_ptr(), out.data_ptr()))
        return out

    def apply_G(self, pressure: Any) -> Any:
        g = self.grid
        p = self._dev(pressure).reshape(-1)
        out = self._empty(g.size)
        self._check(self._lib.mac_grid_gradient(g.nx, g.ny, float(g.lx), float(g.ly), p.data_ptr(),
                                                out.data_ptr()))
        return out

    def apply_D(self, velocity: Any) -> Any:
        g = self.grid
        vel = self._dev(velocity)
        out = self._empty(g.np)
        self._check(self._lib.mac_grid_divergence(g.nx, g.ny, float(g.lx), float(g.ly), vel.data_ptr(),
                                                  out.data_ptr()))
        return out

    def inner(self, left: Any, right: Any) -> float:
        a, b = self._dev(left).reshape(-1), self._dev(right).reshape(-1)
        result = ctypes.c_double()
        self._check(self._lib.mac_dot(a.numel(), a.data_ptr(), b.data_ptr(), ctypes.byref(result)))
        return float(self.grid.area * result.value)

    def max_abs(self, field: Any) -> float:
        a = self._dev(field).reshape(-1)
        result = ctypes.c_double()
        self._check(self._lib.mac_max_abs(a.numel(), a.data_ptr(), ctypes.byref(result)))
        return float(result.value)

    def all_finite(self, field: Any) -> bool:
        return bool(np.isfinite(self.max_abs(field)))

    def solve(self, rhs: Any, *, mass: float, viscosity: float) -> StokesResult:
        self._sync_torch()
        return self.backend.solve(self._dev(rhs), mass=mass, viscosity=viscosity)

    def solve_columns(self, columns: Sequence[Any], *, mass: float, viscosity: float) -> list[SolveResult]:
        import torch
        # Columns stacked as rows of a (k, size) tensor: the backend's column-major input,
        # so no transpose copy; the results come back the same way and each column is a
        # contiguous row view.
        stacked = torch.stack([self._dev(c).reshape(-1) for c in columns], dim=0).t()
        out = self.backend.solve(stacked, mass=mass, viscosity=viscosity)
        velocity, pressure = out.velocity.t(), out.pressure.permute(2, 0, 1)
        return [StokesResult(velocity[k], pressure[k], out.residual, out.divergence_inf)
                for k in range(len(columns))]

    # ------------------------------------------------------------- diagnostics
    def collocated_velocity(self, velocity: Any) -> tuple[Any, Any]:
        g = self.grid
        vel = self._dev(velocity)
        zc = vel.new_zeros((g.ny, 1))
        zr = vel.new_zeros((1, g.nx))
        u = torch_cat_cols(zc, vel[:g.nu].reshape(g.ny, g.nx - 1), zc)
        v = torch_cat_rows(zr, vel[g.nu:].reshape(g.ny - 1, g.nx), zr)
        return 0.5 * (u[:, :-1] + u[:, 1:]), 0.5 * (v[:-1, :] + v[1:, :])

    def angular_momentum(self, velocity: Any) -> float:
        cu, cv = self.collocated_velocity(velocity)
        x, y = self.grid.coordinates("p")
        xs, ys = self._dev(x - 0.5 * self.grid.lx), self._dev(y - 0.5 * self.grid.ly)
        return float(self.grid.area * (xs * cv - ys * cu).sum().item())

    def diagnostics(self, state: State) -> dict[str, float]:
        velocity = self.vector(state)
        kinetic = 0.5 * self.inner(velocity, velocity)
        h1 = self.inner(velocity, self.apply_K(velocity))
        return {"kinetic": kinetic, "modified_energy": kinetic + 0.5 * (state.r - 1) ** 2,
                "divergence_inf": self.max_abs(self.apply_D(velocity)),
                "h1_seminorm_squared": h1, "r": state.r,
                "angular_momentum": self.angular_momentum(velocity),
                "power": self.inner(self.force(state.t), velocity),
                "dissipation": self.nu * h1}


def torch_cat_cols(*parts: Any) -> Any:
    import torch
    return torch.cat(parts, dim=1)


def torch_cat_rows(*parts: Any) -> Any:
    import torch
    return torch.cat(parts, dim=0)
