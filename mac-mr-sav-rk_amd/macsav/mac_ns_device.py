"""`DeviceMACNavierStokes`: the MAC model with its fields resident on the GPU.

The same model as `MACNavierStokes` over `DeviceFields`: velocities and pressures are
torch device tensors, and every field operation runs in libmacstokes.so's model
kernels on the null stream (torch's default stream):
- convection: the skew-symmetric dual-cell form, bitwise the reference loop
  (solver/mac/kernels.py:6-35);
- combinations: sum of coefficient * field, left to right, bitwise the host model;
- K, G = -D^T, D: the reference's CSR products in row order, matrix-free;
- dot products and max-norms: FP64 fixed-order two-level reductions.
The stage solves take the device tensors in place (no copies).

Reductions can be queued instead of read one by one: `queue_reduction` writes the
result into a numbered device slot without waiting, `read_reductions` fetches the
slots in one transfer.  The stage engine uses this to make one host read per stage
(the SAV pairings), and the diagnostics of a state are one read as well.
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
from .mac_model import MACNavierStokes, angular_weights

_MODE = {"dot": 0, "maxabs": 1}


class DeviceFields:
    """Field algebra on torch device tensors through the library's model kernels."""

    SLOTS = 256       # slots [0, QUEUE) for queued reductions, then diagnostics, then one scratch
    QUEUE = 240
    _DIAG = 248
    _ONE = 255

    def __init__(self, grid: MACGrid, device: Any) -> None:
        import torch
        self.grid, self.device = grid, device
        self.lib = _native.load()
        self.slots = torch.zeros(self.SLOTS, dtype=torch.float64, device=device)
        self.scales = np.ones(self.SLOTS)
        self._angular = self.field(angular_weights(grid))
        g = grid
        self._geom = (g.nx, g.ny, float(g.lx), float(g.ly))

    def _check(self, code: int) -> None:
        _native.raise_for(code, self.lib.mac_model_last_error().decode())

    def field(self, x: Any) -> Any:
        import torch
        return torch.as_tensor(x, dtype=torch.float64, device=self.device).contiguous()

    def _empty(self, n: int) -> Any:
        import torch
        return torch.empty(n, dtype=torch.float64, device=self.device)

    def lincomb(self, terms: Sequence[tuple[float, Any]]) -> Any:
        terms = list(terms)
        if len(terms) > 8:   # the kernel takes up to eight terms; fold the rest in order
            return self.lincomb([(1.0, self.lincomb(terms[:8]))] + terms[8:])
        fields = [self.field(f).reshape(-1) for _, f in terms]
        coeff = (ctypes.c_double * len(terms))(*[float(c) for c, _ in terms])
        ptrs = (ctypes.c_void_p * len(terms))(*[f.data_ptr() for f in fields])
        out = self._empty(fields[0].numel())
        self._check(self.lib.mac_combine(out.numel(), len(terms), coeff, ptrs, out.data_ptr()))
        return out

    def K(self, u: Any) -> Any:
        out = self._empty(self.grid.size)
        self._check(self.lib.mac_apply_K(*self._geom, u.data_ptr(), out.data_ptr()))
        return out

    def G(self, p: Any) -> Any:
        p = self.field(p).reshape(-1)
        out = self._empty(self.grid.size)
        self._check(self.lib.mac_grid_gradient(*self._geom, p.data_ptr(), out.data_ptr()))
        return out

    def D(self, u: Any) -> Any:
        out = self._empty(self.grid.np)
        self._check(self.lib.mac_grid_divergence(*self._geom, u.data_ptr(), out.data_ptr()))
        return out

    def convection(self, u: Any) -> Any:
        g = self.grid
        out = self._empty(g.size)
        self._check(self.lib.mac_convection(*self._geom, u.data_ptr(), g.size, 1, out.data_ptr(), g.size))
        return out

    # ---- reductions: queued into slots, read in one transfer
    def queue(self, slot: int, kind: str, a: Any, b: Any = None, scale: float = 1.0) -> None:
        if not 0 <= slot < self.SLOTS:
            raise ValueError("reduction slot out of range")
        a = self.field(a).reshape(-1)
        bp = None
        if b is not None:
            b = self.field(b).reshape(-1)
            bp = b.data_ptr()
        self.scales[slot] = scale
        out = self.slots.data_ptr() + 8 * slot
        self._check(self.lib.mac_reduce_into(a.numel(), a.data_ptr(), bp, _MODE[kind], out))

    def read(self, count: int, first: int = 0) -> np.ndarray:
        stop = first + count
        return self.slots[first:stop].cpu().numpy() * self.scales[first:stop]

    def dot(self, a: Any, b: Any) -> float:
        self.queue(self._ONE, "dot", a, b)
        return float(self.read(1, self._ONE)[0])

    def maxabs(self, a: Any) -> float:
        self.queue(self._ONE, "maxabs", a)
        return float(self.read(1, self._ONE)[0])

    def finite(self, a: Any) -> bool:
        return bool(np.isfinite(self.maxabs(a)))

    def stack(self, columns: Sequence[Any]) -> Any:
        import torch
        # Rows of a (k, size) tensor: the backend's column-major input, no transpose copy.
        return torch.stack([self.field(c).reshape(-1) for c in columns], dim=0).t()

    def split(self, out: StokesResult, k: int) -> list[StokesResult]:
        velocity, pressure = out.velocity.t(), out.pressure.permute(2, 0, 1)
        return [StokesResult(velocity[c], pressure[c], out.residual, out.divergence_inf) for c in range(k)]

    def unpack_state(self, velocity: Any) -> tuple[Any, Any]:
        return self.field(velocity).reshape(-1), None

    def pack_state(self, state: State) -> Any:
        return state.u

    def diagnostic_sums(self, u: Any, f: Any) -> tuple[float, ...]:
        u = self.field(u).reshape(-1)
        d = self._DIAG
        self.queue(d, "dot", u, u)
        self.queue(d + 1, "dot", u, self.K(u))
        self.queue(d + 2, "maxabs", self.D(u))
        self.queue(d + 3, "dot", self._angular, u)
        self.queue(d + 4, "dot", f, u)
        return tuple(float(v) for v in self.read(5, d))


class DeviceMACNavierStokes(MACNavierStokes):
    """The MAC model with device-resident fields (see the module docstring)."""

    def __init__(self, grid: MACGrid, nu: float, force: Callable[[float], Any] | None = None,
                 backend: Any = None, device: int = 0, lid: float = 0.0) -> None:
        import torch
        self.device = torch.device(f"cuda:{device}")
        if backend is None:
            from .mac.gpu_stokes import GpuTensorStokes
            backend = GpuTensorStokes(grid, device=device)
        super().__init__(grid, nu, force=force, backend=backend, lid=lid, fields=DeviceFields(grid, self.device))
        self._queued = 0

    def host_velocity(self, state: State) -> np.ndarray:
        return state.u.cpu().numpy()

    # ---- deferred reductions for the stage engine (slots 0.. of this model)
    def reset_reductions(self) -> None:
        self._queued = 0

    def queue_reduction(self, slot: int, kind: str, a: Any, b: Any = None, scale: float = 1.0) -> None:
        if slot >= DeviceFields.QUEUE:
            raise RuntimeError("too many queued reductions in one step")
        self.fields.queue(slot, kind, a, b, scale)
        self._queued = max(self._queued, slot + 1)

    def read_reductions(self, count: int) -> np.ndarray:
        return self.fields.read(count)
