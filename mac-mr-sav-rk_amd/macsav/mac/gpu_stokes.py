"""`GpuTensorStokes`: the MI355X stage-solve backend behind the `StokesBackend` seam.

Drop-in for the reference's `TensorStokes` (solver/mac/tensor_stokes.py:118-364): same
constructor controls, same `solve(rhs, *, mass, viscosity) -> StokesResult`, same
validation errors (ValueError), same failure modes (RuntimeError on Schur-CG
non-convergence and on the residual/divergence gate) and the same counters
(`setup_seconds`, `solve_seconds`, `solves`, `right_hand_sides`, `total_iterations`,
`last_iterations`).  The work runs in libmacstokes.so on the GPU; this class only
validates, moves buffers and maps status codes.

Inputs may be NumPy arrays (copied to the device and back) or torch device tensors
(zero-copy; results come back as torch tensors on the same device).

`precision="mixed"` runs the transforms as bf16 MFMA GEMMs inside an FP64
defect-correction loop.  `fp64_refine=n` (FP64 only) follows the reference solve with
up to n FP64 defect-correction sweeps until the residual and max|D u| are both
<= `refine_tolerance`: at large grids the reference algorithm's own rounding floor for
max|D u| lies above 1e-10 (about 2e-10 at 1024^2), and one sweep removes it.
`matrix_free_operators` and `reuse_transform_buffers` are accepted for drop-in
compatibility and validated; the GPU solve is always matrix-free and always reuses
its buffers, so they change nothing.  Warm starts: `set_warm_start`.
"""
from __future__ import annotations

import ctypes
import math
from time import perf_counter
from typing import Any

import numpy as np

from .. import _native
from .grid import MACGrid
from .stokes import StokesResult

_PRECISIONS = {"fp64": _native.MAC_PRECISION_FP64, "mixed": _native.MAC_PRECISION_MIXED}


def _is_torch(x: Any) -> bool:
    return type(x).__module__.startswith("torch")


class GpuTensorStokes:
    """Shift-reusable MAC Stokes solve on one GPU (DST/DCT MFMA GEMMs + Schur-CG)."""

    def __init__(self, operators: Any, tolerance: float = 1e-9, krylov_tolerance: float = 1e-12,
                 max_iterations: int = 100, matrix_free_operators: bool = True,
                 reuse_transform_buffers: bool = True, *, precision: str = "fp64",
                 refine_tolerance: float = 1e-10, inner_tolerance: float | None = None, max_sweeps: int = 30,
                 mixed_split: int = 15, fp64_refine: int = 0, device: int = 0) -> None:
        if (tolerance <= 0 or krylov_tolerance <= 0 or not np.isfinite([tolerance, krylov_tolerance]).all()
                or max_iterations < 1):
            raise ValueError("Invalid solver controls")
        if not isinstance(matrix_free_operators, bool):
            raise ValueError("matrix_free_operators must be a bool")
        if not isinstance(reuse_transform_buffers, bool):
            raise ValueError("reuse_transform_buffers must be a bool")
        if precision not in _PRECISIONS:
            raise ValueError(f"precision must be one of {sorted(_PRECISIONS)}")
        start = perf_counter()
        grid = operators if isinstance(operators, MACGrid) else operators.grid
        self.ops = None if isinstance(operators, MACGrid) else operators
        self.grid = grid
        self.tolerance = tolerance
        self.krylov_tolerance = krylov_tolerance
        self.max_iterations = max_iterations
        self.matrix_free_operators = matrix_free_operators
        self.reuse_transform_buffers = reuse_transform_buffers
        self.precision = precision
        lib = _native.load()
        opts = _native.default_opts()
        opts.tolerance = tolerance
        opts.krylov_tolerance = krylov_tolerance
        opts.max_iterations = max_iterations
        opts.precision = _PRECISIONS[precision]
        opts.refine_tolerance = refine_tolerance
        # The mixed form's inner tolerance is the cap of its per-sweep tolerance (the sweep's
        # defect sets it below that): 1e-2, where one inner CG iteration from a history start
        # usually reaches the bar; refined FP64 keeps 1e-3 for its sweeps.
        if inner_tolerance is None:
            inner_tolerance = 1e-2 if precision == "mixed" else 1e-3
        opts.inner_tolerance = inner_tolerance
        opts.max_sweeps = max_sweeps
        opts.mixed_split = mixed_split
        opts.fp64_refine = fp64_refine
        opts.device = device
        handle = lib.mac_stokes_create(grid.nx, grid.ny, float(grid.lx), float(grid.ly), ctypes.byref(opts))
        if not handle:
            message = lib.mac_stokes_create_error().decode()
            if message.startswith("Invalid solver controls") or message.startswith("Require"):
                raise ValueError(message)
            raise RuntimeError(f"GPU Stokes setup failed: {message}")
        self._lib = lib
        self._handle = handle
        self._stream: int | None = None  # the stream handed to the library (torch's current one)
        self.device = device
        self.setup_seconds = perf_counter() - start
        self.solve_seconds = 0.0
        self.solves = 0
        self.right_hand_sides = 0
        self.total_iterations = 0
        self.last_iterations: tuple[int, ...] = ()
        self.sweeps = 0
        self.last_sweeps = 0
        self.warm_starts = 0
        self.network_starts = 0
        self.last_report: dict[str, float] = {}
        self.warm_start = None
        self._outs: dict[int, tuple[Any, Any]] = {}

    def __del__(self) -> None:
        handle = getattr(self, "_handle", None)
        if handle:
            self._lib.mac_stokes_destroy(handle)
            self._handle = None

    # ------------------------------------------------------------------ solve
    def solve(self, rhs: Any, *, mass: float, viscosity: float,
              pressure_guess: Any = None) -> StokesResult:
        mass, viscosity = float(mass), float(viscosity)
        if (not (math.isfinite(mass) and math.isfinite(viscosity)) or min(mass, viscosity) < 0
                or mass + viscosity <= 0):
            raise ValueError("Require mass,viscosity >= 0 and nonzero total")
        grid = self.grid
        if _is_torch(rhs):
            return self._solve_torch(rhs, float(mass), float(viscosity), pressure_guess)
        rhs = np.asarray(rhs)
        if rhs.ndim not in (1, 2) or rhs.shape[0] != grid.size or not np.isfinite(rhs).all():
            raise ValueError("Invalid Stokes RHS")
        two_d = rhs.ndim == 2
        cols = np.ascontiguousarray((rhs if two_d else rhs[:, None]).T, dtype=np.float64)
        k = cols.shape[0]
        vel = np.empty((k, grid.size))
        pres = np.empty((k, grid.np))
        guess_ptr, ld_guess = None, 0
        if pressure_guess is not None:
            g = np.ascontiguousarray(np.asarray(pressure_guess, dtype=np.float64).reshape(grid.np, k).T)
            guess_ptr, ld_guess = g.ctypes.data, grid.np
        report = self._call(cols.ctypes.data, grid.size, k, mass, viscosity, vel.ctypes.data, pres.ctypes.data,
                            guess_ptr, ld_guess, device=False)
        velocity = vel.T.copy() if two_d else vel[0].copy()
        pressure = (pres.T.reshape(grid.ny, grid.nx, k).copy() if two_d
                    else pres[0].reshape(grid.ny, grid.nx).copy())
        return StokesResult(velocity, pressure, report["residual"], report["divergence_inf"])

    def _solve_torch(self, rhs: Any, mass: float, viscosity: float, pressure_guess: Any) -> StokesResult:
        import torch

        grid = self.grid
        if rhs.ndim not in (1, 2) or rhs.shape[0] != grid.size or rhs.device.type != "cuda":
            raise ValueError("Invalid Stokes RHS")
        # Non-finite values are rejected by the library (MAC_INVALID_ARGUMENT -> ValueError).
        two_d = rhs.ndim == 2
        cols = (rhs if two_d else rhs[:, None]).t().contiguous().to(torch.float64)
        k = cols.shape[0]
        vel = torch.empty((k, grid.size), dtype=torch.float64, device=rhs.device)
        pres = torch.empty((k, grid.np), dtype=torch.float64, device=rhs.device)
        guess_ptr, ld_guess, guess = None, 0, None
        if pressure_guess is not None:
            guess = torch.as_tensor(pressure_guess, device=rhs.device).to(torch.float64).reshape(grid.np, k)
            guess = guess.t().contiguous()
            guess_ptr, ld_guess = guess.data_ptr(), grid.np
        # The solve runs on the caller's current stream, ordered after the work that made
        # the right-hand side: no synchronisation here, and the outputs are stream-ordered.
        stream = torch.cuda.current_stream(rhs.device).cuda_stream
        if stream != self._stream:
            _native.raise_for(self._lib.mac_stokes_set_stream(self._handle, stream, 1),
                              self._lib.mac_stokes_last_error(self._handle).decode())
            self._stream = stream
        report = self._call(cols.data_ptr(), grid.size, k, mass, viscosity, vel.data_ptr(), pres.data_ptr(),
                            guess_ptr, ld_guess, device=True)
        velocity = vel.t() if two_d else vel[0]
        pressure = pres.t().reshape(grid.ny, grid.nx, k) if two_d else pres[0].reshape(grid.ny, grid.nx)
        return StokesResult(velocity, pressure, report["residual"], report["divergence_inf"])

    def _call(self, rhs_ptr: int, ld: int, k: int, mass: float, viscosity: float, vel_ptr: int, pres_ptr: int,
              guess_ptr: int | None, ld_guess: int, *, device: bool) -> dict[str, float]:
        grid = self.grid
        # The result structures are reused from call to call (one per column count).
        out = self._outs.get(k)
        if out is None:
            out = self._outs[k] = ((ctypes.c_int32 * k)(), _native.SolveReport())
        iters, report = out
        start = perf_counter()
        code = self._lib.mac_stokes_solve(self._handle, rhs_ptr, ld, k, mass, viscosity, vel_ptr, grid.size,
                                          pres_ptr, grid.np, guess_ptr, ld_guess, iters, ctypes.byref(report),
                                          1 if device else 0)
        elapsed = perf_counter() - start
        message = self._lib.mac_stokes_last_error(self._handle).decode() if code != 0 else ""
        if code == _native.MAC_NOT_CONVERGED or code in (_native.MAC_INVALID_ARGUMENT, _native.MAC_HIP_ERROR):
            _native.raise_for(code, message)
        self.solve_seconds += elapsed
        self.solves += 1
        self.right_hand_sides += k
        self.total_iterations += int(sum(iters))
        self.last_iterations = tuple(int(i) for i in iters)
        self.sweeps += int(report.sweeps)
        self.last_sweeps = int(report.sweeps)
        self.warm_starts += int(report.warm_started > 0)
        self.network_starts += int(report.warm_started == 2)
        self.last_report = {"residual": report.residual, "divergence_inf": report.divergence_inf,
                            "velocity_max": report.velocity_max, "seconds": report.seconds,
                            "sweeps": int(report.sweeps), "total_iterations": int(report.total_iterations),
                            "warm_started": bool(report.warm_started), "network": report.warm_started == 2,
                            "first_residual": report.first_residual,
                            "first_divergence": report.first_divergence}
        _native.raise_for(code, message)
        return self.last_report

    # ---------------------------------------------------- learned warm start
    def set_warm_start(self, net: Any | None, period: int | None = None,
                       history_steps: int | None = None, network_depth: int | None = None) -> None:
        """Starting pressure when no `pressure_guess` is passed: a trained `WarmStartNet`
        (its correction of a base guess), "previous" (the previous solve's pressures),
        "history" (the same stage of the last three steps extrapolated quadratically in
        time, with `period` stage solves per step; `history_steps` sets how many steps
        the extrapolation reaches back, 1..6, default 6), or None (zero, as the
        reference).  With a network and a period, the network corrects the extrapolation
        only while fewer than `network_depth` steps are recorded (default 3).  `period` also
        makes the network's base the history extrapolation.  FP64 starts the Schur CG
        from the Galerkin-scaled guess; mixed takes it as the initial outer iterate.
        Either way the solve runs to the same tolerance, so results are unchanged to
        tolerance."""
        if network_depth is not None:
            _native.raise_for(self._lib.mac_stokes_set_network_depth(self._handle, int(network_depth)),
                              "network_depth must be in 1..7")
        if history_steps is not None:
            _native.raise_for(self._lib.mac_stokes_set_history_steps(self._handle, int(history_steps)),
                              "history_steps must be in 1..6")
        if period is not None:
            _native.raise_for(self._lib.mac_stokes_set_warm_history(self._handle, int(period)),
                              "period must be in 0..16")
        elif net == "history":
            raise ValueError('warm start "history" needs period= (stage solves per step)')
        self.warm_start = net
        if net is None:
            mode, handle, alpha = _native.MAC_WARM_OFF, None, 1.0
        elif isinstance(net, str):
            if net not in ("previous", "history"):
                raise ValueError('warm start must be a WarmStartNet, "previous", "history" or None')
            mode = _native.MAC_WARM_PREVIOUS if net == "previous" else _native.MAC_WARM_HISTORY
            handle, alpha = None, 1.0
        else:
            mode, handle, alpha = _native.MAC_WARM_NETWORK, net.handle, float(net.alpha)
        code = self._lib.mac_stokes_set_warm_start(self._handle, mode, handle, alpha)
        _native.raise_for(code, self._lib.mac_stokes_last_error(self._handle).decode())

    def reset_history(self) -> None:
        """Forget the previous solve's pressures (they feed the warm start)."""
        self._lib.mac_stokes_reset_history(self._handle)

    def commit_history(self) -> None:
        """The solves since the last mark belong to an accepted step: keep them."""
        _native.raise_for(self._lib.mac_stokes_history_mark(self._handle, 0), "history mark failed")

    def discard_history(self) -> None:
        """The solves since the last mark belong to a rejected or failed trial: forget
        them, so later guesses extrapolate accepted steps only."""
        _native.raise_for(self._lib.mac_stokes_history_mark(self._handle, 1), "history mark failed")

    def last_schur_rhs(self, ncols: int) -> Any:
        """The last FP64 solve's Schur right-hand sides b, (np, ncols) device tensor."""
        import torch
        out = torch.empty((ncols, self.grid.np), dtype=torch.float64, device=f"cuda:{self.device}")
        code = self._lib.mac_stokes_last_schur_rhs(self._handle, out.data_ptr(), self.grid.np, ncols)
        _native.raise_for(code, self._lib.mac_stokes_last_error(self._handle).decode())
        return out.t()

    # ------------------------------------------------------------ measurement
    def set_profiling(self, on: bool) -> None:
        """Time every transform-GEMM launch with HIP events on the solver's stream."""
        self._lib.mac_stokes_set_profiling(self._handle, 1 if on else 0)

    def kernel_profile(self) -> dict[str, float]:
        prof = _native.KernelProfile()
        code = self._lib.mac_stokes_get_profile(self._handle, ctypes.byref(prof))
        _native.raise_for(code, self._lib.mac_stokes_last_error(self._handle).decode())
        return {"gemm_ms": prof.gemm_ms, "gemm_flops": prof.gemm_flops, "gemm_bytes": prof.gemm_bytes,
                "gemm_launches": int(prof.gemm_launches), "gemm_transform_flops": prof.gemm_transform_flops,
                "fft_ms": prof.fft_ms, "fft_bytes": prof.fft_bytes, "fft_launches": int(prof.fft_launches),
                "fft_transform_flops": prof.fft_transform_flops}

    # ----------------------------------------------- single operators (device)
    def apply_divergence(self, velocity: Any) -> Any:
        """D on packed velocity columns (torch device tensor, (size,) or (size, k))."""
        return self._op(velocity, self.grid.size, self.grid.np, "mac_apply_divergence")

    def apply_divergence_transpose(self, pressure: Any) -> Any:
        """D^T on pressure columns (torch device tensor, (np,) or (np, k))."""
        return self._op(pressure, self.grid.np, self.grid.size, "mac_apply_divergence_transpose")

    def velocity_inverse(self, values: Any, mass: float, viscosity: float) -> Any:
        """(mass I + viscosity K)^-1 on packed velocity columns (torch device tensor)."""
        return self._op(values, self.grid.size, self.grid.size, "mac_velocity_inverse", mass, viscosity)

    def pressure_poisson_inverse(self, values: Any) -> Any:
        """(D D^T)^+ on pressure columns, mean removed (torch device tensor)."""
        return self._op(values, self.grid.np, self.grid.np, "mac_pressure_poisson_inverse")

    def _op(self, x: Any, n_in: int, n_out: int, name: str, *scalars: float) -> Any:
        import torch

        if not _is_torch(x) or x.device.type != "cuda" or x.shape[0] != n_in or x.ndim not in (1, 2):
            raise ValueError("expected a device tensor of matching length")
        two_d = x.ndim == 2
        cols = (x if two_d else x[:, None]).t().contiguous().to(torch.float64)
        k = cols.shape[0]
        out = torch.empty((k, n_out), dtype=torch.float64, device=x.device)
        torch.cuda.current_stream(x.device).synchronize()
        fn = getattr(self._lib, name)
        if scalars:
            code = fn(self._handle, cols.data_ptr(), n_in, k, *scalars, out.data_ptr(), n_out)
        else:
            code = fn(self._handle, cols.data_ptr(), n_in, k, out.data_ptr(), n_out)
        _native.raise_for(code, self._lib.mac_stokes_last_error(self._handle).decode())
        return out.t() if two_d else out[0]
