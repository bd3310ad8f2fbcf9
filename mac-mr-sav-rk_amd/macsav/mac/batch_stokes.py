"""`GpuBatchedStokes`: the stage solves of many independent MAC cavities in one call.

The reference advances independent cavities by calling `TensorStokes.solve`
(solver/mac/tensor_stokes.py:245-364) once per cavity.  Here every cavity of a batch
(its own grid, lengths, (mass, viscosity) and right-hand-side columns) is solved by one
library call, `mac_stokes_solve_batched`, whose transform passes and vector steps each
cover every cavity in one grouped launch.

`precision="fp64"` (default) gives every cavity the results of `GpuTensorStokes` on that
cavity alone, bitwise.  `precision="mixed"` runs the mixed form per cavity: FP64 defect
correction to `refine_tolerance` on the residual and max|D u|, around inner Schur-CG solves
(to `inner_tolerance`) whose transforms are split bf16 with FP32 accumulation; cavities
that have met the bar drop out of the remaining sweeps on the device.

`solve(rhs_list, mass=..., viscosity=...)` returns one `StokesResult` per cavity; the
zero-copy form `solve_packed` works on one flat device buffer in the batch layout given
by `layout()`.  `autotune` times the launch shapes (graph replay or direct launches,
CG iterations queued per host check; for the mixed form also the stage graph with its CG
iterations per round) on a real right-hand side and keeps the fastest.

Launch shapes `(use_graph, n)`: `(0, n)` direct launches and `(1, n)` one captured CG
iteration replayed, with n iterations queued per host read of the convergence flags;
`(2, n)` (mixed form, the default there) the stage graph: one captured hipGraph launch per
stage, whose outer tests, inner set-ups, n CG iterations per round and updates are switched
by device flags, with one host read per launch.  Results do not depend on the shape.
"""
from __future__ import annotations

import ctypes
from time import perf_counter
from typing import Any, Sequence

import numpy as np

from .. import _native
from .stokes import StokesResult


def _is_torch(x: Any) -> bool:
    return type(x).__module__.startswith("torch")


def _as_list(v: Any, n: int, name: str) -> np.ndarray:
    a = np.asarray(v, dtype=float)
    if a.ndim == 0:
        a = np.full(n, float(a))
    if a.shape != (n,):
        raise ValueError(f"{name} must be a scalar or one value per cavity")
    return np.ascontiguousarray(a)


class GpuBatchedStokes:
    """Batched stage solve over cavities of uneven sizes on one GPU (FP64 or mixed form)."""

    LAUNCH_SHAPES = ((1, 1), (1, 2), (1, 3), (0, 1))
    STAGE_SHAPES = ((2, 2), (2, 4), (2, 8))

    def __init__(self, grids: Sequence[Any], ncols: int = 2, tolerance: float = 1e-9,
                 krylov_tolerance: float = 1e-12, max_iterations: int = 100, device: int = 0,
                 precision: str = "fp64", refine_tolerance: float = 1e-10, inner_tolerance: float = 1e-3,
                 max_sweeps: int = 30) -> None:
        if (tolerance <= 0 or krylov_tolerance <= 0 or not np.isfinite([tolerance, krylov_tolerance]).all()
                or max_iterations < 1 or precision not in ("fp64", "mixed")):
            raise ValueError("Invalid solver controls")
        grids = [g if hasattr(g, "nx") else g.grid for g in grids]
        if not grids:
            raise ValueError("Require at least one cavity")
        start = perf_counter()
        self.grids = grids
        self.ncols = int(ncols)
        self.tolerance = tolerance
        self.krylov_tolerance = krylov_tolerance
        self.max_iterations = max_iterations
        self.precision = precision
        lib = _native.load()
        opts = _native.default_opts()
        opts.tolerance = tolerance
        opts.krylov_tolerance = krylov_tolerance
        opts.max_iterations = max_iterations
        opts.device = device
        if precision == "mixed":
            opts.precision = _native.MAC_PRECISION_MIXED
            opts.refine_tolerance = refine_tolerance
            opts.inner_tolerance = inner_tolerance
            opts.max_sweeps = max_sweeps
        n = len(grids)
        nx = (ctypes.c_int32 * n)(*[g.nx for g in grids])
        ny = (ctypes.c_int32 * n)(*[g.ny for g in grids])
        lx = (ctypes.c_double * n)(*[float(g.lx) for g in grids])
        ly = (ctypes.c_double * n)(*[float(g.ly) for g in grids])
        handle = lib.mac_batch_create(n, nx, ny, lx, ly, self.ncols, ctypes.byref(opts))
        if not handle:
            message = lib.mac_batch_create_error().decode()
            if message.startswith(("Require", "Invalid solver controls")):
                raise ValueError(message)
            raise RuntimeError(message)
        self._lib, self._h = lib, handle
        nseg = n * self.ncols
        vo, po = (ctypes.c_int64 * (nseg + 1))(), (ctypes.c_int64 * (nseg + 1))()
        _native.raise_for(lib.mac_batch_layout(handle, vo, po), "mac_batch_layout failed")
        self.velocity_offsets = np.frombuffer(vo, dtype=np.int64).copy()
        self.pressure_offsets = np.frombuffer(po, dtype=np.int64).copy()
        self.launch = (1, 1)
        if precision == "mixed":
            self.set_launch(2, 4)
        self.setup_seconds = perf_counter() - start
        self.solve_seconds = 0.0
        self.solves = 0
        self.right_hand_sides = 0
        self.total_iterations = 0
        self.last_iterations: list[list[int]] = []
        self.last_sweeps = 0

    def __del__(self) -> None:
        h = getattr(self, "_h", None)
        if h:
            self._lib.mac_batch_destroy(h)
            self._h = None

    # ------------------------------------------------------------------ layout
    @property
    def total_faces(self) -> int:
        return int(self.velocity_offsets[-1])

    @property
    def total_cells(self) -> int:
        return int(self.pressure_offsets[-1])

    def layout(self) -> tuple[np.ndarray, np.ndarray]:
        """First element of every (cavity, column) segment in the velocity and pressure
        buffers, cavity-major and column-minor, with the totals last."""
        return self.velocity_offsets, self.pressure_offsets

    def pack(self, rhs_list: Sequence[Any]) -> Any:
        """Per-cavity right-hand sides ((size,) or (size, ncols)) -> one flat buffer."""
        if len(rhs_list) != len(self.grids):
            raise ValueError("Require one right-hand side per cavity")
        parts = []
        for g, r in zip(self.grids, rhs_list):
            shape = tuple(r.shape)
            if shape not in ((g.size, self.ncols),) + (((g.size,),) if self.ncols == 1 else ()):
                raise ValueError("Invalid Stokes RHS")
            parts.append(r.reshape(g.size, self.ncols).T.reshape(-1) if not _is_torch(r)
                         else r.reshape(g.size, self.ncols).t().reshape(-1))
        if _is_torch(rhs_list[0]):
            import torch
            return torch.cat([p.to(torch.float64) for p in parts]).contiguous()
        return np.ascontiguousarray(np.concatenate(parts).astype(float))

    # ------------------------------------------------------------------- solve
    def set_launch(self, use_graph: int, iters_per_check: int) -> None:
        _native.raise_for(self._lib.mac_batch_set_launch(self._h, int(use_graph), int(iters_per_check)),
                          "Invalid launch shape")
        self.launch = (int(use_graph), int(iters_per_check))

    def set_history(self, period: int, steps: int = 3) -> None:
        """Mixed form: start every batched solve from the extrapolation in time of the
        same stage of the last `steps` steps (`period` solves per step; 0 switches it off).
        Results are unchanged to the bar; only the outer sweeps' starting point moves."""
        code = self._lib.mac_batch_set_history(self._h, int(period), int(steps))
        _native.raise_for(code, self._lib.mac_batch_last_error(self._h).decode())

    def commit_history(self) -> None:
        """Keep the solves since the last mark (an accepted step)."""
        _native.raise_for(self._lib.mac_batch_history_mark(self._h, 0), "history mark failed")

    def discard_history(self) -> None:
        """Forget the solves since the last mark (a rejected trial)."""
        _native.raise_for(self._lib.mac_batch_history_mark(self._h, 1), "history mark failed")

    def solve_packed(self, rhs: Any, *, mass: Any, viscosity: Any, velocity: Any = None,
                     pressure: Any = None) -> dict:
        """Solve on flat buffers in the batch layout (torch device tensors are used in
        place; NumPy arrays are copied).  Returns per-cavity residual / divergence_inf
        and per-segment iterations."""
        n = len(self.grids)
        m = _as_list(mass, n, "mass")
        v = _as_list(viscosity, n, "viscosity")
        if not (np.isfinite(m).all() and np.isfinite(v).all()) or (np.minimum(m, v) < 0).any() or \
                ((m + v) <= 0).any():
            raise ValueError("Require mass,viscosity >= 0 and nonzero total")
        if tuple(rhs.shape) != (self.total_faces,):
            raise ValueError("Invalid Stokes RHS")
        for buf, length in ((velocity, self.total_faces), (pressure, self.total_cells)):
            if buf is not None and (tuple(buf.shape) != (length,) or str(buf.dtype).split(".")[-1] != "float64"
                                    or not (buf.is_contiguous() if _is_torch(buf) else buf.flags.c_contiguous)):
                raise ValueError("Output buffers must be contiguous float64 of the batch length")
        on_device = _is_torch(rhs) and rhs.is_cuda
        if on_device:
            import torch
            rhs = rhs.to(torch.float64).contiguous()
            if velocity is None:
                velocity = torch.empty_like(rhs)
            if pressure is None:
                pressure = torch.empty(self.total_cells, dtype=torch.float64, device=rhs.device)
            torch.cuda.current_stream(rhs.device).synchronize()
            ptrs = (rhs.data_ptr(), velocity.data_ptr(), pressure.data_ptr())
        else:
            rhs = np.ascontiguousarray(np.asarray(rhs, dtype=float))
            velocity = np.empty(self.total_faces) if velocity is None else velocity
            pressure = np.empty(self.total_cells) if pressure is None else pressure
            ptrs = (rhs.ctypes.data, velocity.ctypes.data, pressure.ctypes.data)
        nseg = n * self.ncols
        its = (ctypes.c_int32 * nseg)()
        res = (ctypes.c_double * n)()
        div = (ctypes.c_double * n)()
        start = perf_counter()
        code = self._lib.mac_stokes_solve_batched(self._h, ptrs[0], m.ctypes.data, v.ctypes.data, ptrs[1],
                                                  ptrs[2], its, res, div, 1 if on_device else 0)
        elapsed = perf_counter() - start
        iterations = [list(its[i * self.ncols:(i + 1) * self.ncols]) for i in range(n)]
        if code in (_native.MAC_OK, _native.MAC_NOT_CONVERGED, _native.MAC_GATE_FAILED):
            self.solves += 1
            self.right_hand_sides += nseg
            self.total_iterations += int(sum(its))
            self.last_iterations = iterations
            self.last_sweeps = int(self._lib.mac_batch_last_sweeps(self._h))
            self.solve_seconds += elapsed
        _native.raise_for(code, self._lib.mac_batch_last_error(self._h).decode())
        return {"velocity": velocity, "pressure": pressure, "iterations": iterations,
                "residual": list(res), "divergence_inf": list(div), "seconds": elapsed}

    def solve(self, rhs_list: Sequence[Any], *, mass: Any, viscosity: Any) -> list[StokesResult]:
        """One `StokesResult` per cavity (velocity (size[, ncols]), pressure
        (ny, nx[, ncols]))."""
        flat = self.pack(rhs_list)
        out = self.solve_packed(flat, mass=mass, viscosity=viscosity)
        if _is_torch(flat):
            import torch
            stack = torch.stack
        else:
            stack = np.stack
        results = []
        for i, g in enumerate(self.grids):
            s0 = i * self.ncols
            vseg = [out["velocity"][self.velocity_offsets[s0 + c]:self.velocity_offsets[s0 + c] + g.size]
                    for c in range(self.ncols)]
            pseg = [out["pressure"][self.pressure_offsets[s0 + c]:self.pressure_offsets[s0 + c] + g.np]
                    .reshape(g.ny, g.nx) for c in range(self.ncols)]
            single = len(tuple(rhs_list[i].shape)) == 1
            vel = vseg[0] if single else stack(vseg, 1)
            pre = pseg[0] if single else stack(pseg, 2)
            results.append(StokesResult(vel, pre, out["residual"][i], out["divergence_inf"][i]))
        return results

    STAGE_PREFERENCE = 1.02  # a stage-graph shape within 2% of the fastest is kept

    def autotune(self, rhs: Any, *, mass: Any, viscosity: Any, repeats: int = 2) -> dict:
        """Time every launch shape on this right-hand side (packed layout) and keep the
        fastest, or, in the mixed form, the fastest stage-graph shape when it is within 2%
        of that (one launch and one host read per stage); returns {shape: seconds}."""
        timings = {}
        shapes = self.LAUNCH_SHAPES + (self.STAGE_SHAPES if self.precision == "mixed" else ())
        for shape in shapes:
            self.set_launch(*shape)
            self.solve_packed(rhs, mass=mass, viscosity=viscosity)  # warm (graph capture)
            best = float("inf")
            for _ in range(repeats):
                best = min(best, self.solve_packed(rhs, mass=mass, viscosity=viscosity)["seconds"])
            timings[shape] = best
        best = min(timings, key=timings.get)
        stage = [k for k in timings if k[0] == 2]
        if stage:
            fastest_stage = min(stage, key=timings.get)
            if timings[fastest_stage] <= self.STAGE_PREFERENCE * timings[best]:
                best = fastest_stage
        self.set_launch(*best)
        return timings
