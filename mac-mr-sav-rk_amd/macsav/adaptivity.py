"""Adaptive step control for the embedded SDIRK pairs: I and PI controllers and the
driver `integrate_adaptive`.

Restates the reference's solver/adaptivity/controller.py (I/PI controllers :15-72,
scalar control :74-90, error estimate :112-123, driver :126-398) for natural-node runs:
the same trial / accept / reject sequence, error norm and step proposals, so a run
with this driver and the GPU backend makes the reference's step decisions.  Strict
snapshot partitions are not restated (ValueError), and neither are the trial / advance
observers beyond `on_accept` and `progress`.

Error estimate (controller.py:112-123): E = ||high - low|| / (atol + rtol max(||old||,
||high||)) with the model's inner-product norm (or `error_norm`); accept when E <= 1.
A rejected trial retries with step * min(I-factor(E), 0.8) and never changes the
accepted state or the PI history.  After an acceptance the next step is step * factor
(PI from the current and previous accepted errors), capped at max_step.  The terminal
step reaches T exactly and may be below min_step.
"""
from __future__ import annotations

from collections.abc import Callable
from dataclasses import dataclass, field, replace
from time import perf_counter, process_time
from typing import Any, Literal

import numpy as np

from .core import stage_diagnostics
from .integrate import Result


@dataclass(frozen=True)
class IController:
    atol: float = 1e-8
    rtol: float = 1e-4
    safety: float = 0.9
    min_factor: float = 0.2
    max_factor: float = 2.0
    min_step: float = 1e-10
    max_step: float = float("inf")
    max_rejections: int = 12
    max_steps: int = 1_000_000
    estimator_order: int = 3

    def __post_init__(self) -> None:
        if not np.isfinite([self.atol, self.rtol, self.safety, self.min_factor, self.max_factor,
                            self.min_step]).all():
            raise ValueError("Controller parameters must be finite")
        if self.atol < 0 or self.rtol < 0 or self.atol + self.rtol == 0:
            raise ValueError("At least one tolerance must be positive")
        if not (0 < self.safety < 1 and 0 < self.min_factor < 1 < self.max_factor):
            raise ValueError("Invalid I-controller factors")
        if self.min_step <= 0 or np.isnan(self.max_step) or self.max_step < self.min_step:
            raise ValueError("Invalid step bounds")
        if self.max_rejections < 0 or self.max_steps <= 0:
            raise ValueError("Invalid trial limits")
        if type(self.estimator_order) is not int or self.estimator_order < 2:
            raise ValueError("Invalid estimator order")

    def accepted_factor(self, error: float, previous_error: float | None) -> float:
        """I control: safety * E^(-1/order), clipped; a zero error takes the maximum."""
        if error == 0:
            return self.max_factor
        return float(np.clip(self.safety * error ** (-1 / self.estimator_order), self.min_factor,
                             self.max_factor))


@dataclass(frozen=True)
class PIController(IController):
    proportional: float = 0.7
    integral: float = 0.4
    error_floor: float = 1e-14

    def __post_init__(self) -> None:
        super().__post_init__()
        if (not np.isfinite([self.proportional, self.integral, self.error_floor]).all()
                or self.proportional <= self.integral or self.integral < 0 or not 0 < self.error_floor < 1):
            raise ValueError("Invalid PI-controller parameters")

    def accepted_factor(self, error: float, previous_error: float | None) -> float:
        """PI control from the current and previous accepted errors (floored); the first
        accepted step uses the I rule."""
        if previous_error is None:
            return super().accepted_factor(error, previous_error)
        now = max(error, self.error_floor)
        before = max(previous_error, self.error_floor)
        factor = (self.safety * now ** (-self.proportional / self.estimator_order)
                  * before ** (self.integral / self.estimator_order))
        return float(np.clip(factor, self.min_factor, self.max_factor))


@dataclass(frozen=True)
class ScalarErrorControl:
    """Local defect of the auxiliary scalar against r' = -gamma r, over a fixed atol."""
    atol: float
    gamma: float

    def __post_init__(self) -> None:
        if not np.isfinite([self.atol, self.gamma]).all() or self.atol <= 0 or self.gamma < 0:
            raise ValueError("Invalid scalar error controls")

    def defect(self, old_r: float, trial_r: float, step: float) -> float:
        return float(trial_r - np.exp(-self.gamma * step) * old_r)


@dataclass(frozen=True)
class AdaptiveContinuation:
    """Restart data at a natural node; the accepted PI history carries over."""
    step: float
    previous_error: float | None = None
    previous_u: float | None = None
    previous_r: float | None = None
    accepted_steps: int = 0
    attempted_steps: int = 0


@dataclass
class AdaptiveResult(Result):
    attempts: list[dict[str, Any]] = field(default_factory=list)
    accepted_steps: int = 0
    attempted_steps: int = 0
    continuation: AdaptiveContinuation | None = None


def embedded_error(model: Any, old: Any, high: Any, low: Any, controller: IController,
                   error_norm: Callable[[Any], float] | None = None) -> float:
    """||high - low|| / (atol + rtol max(||old||, ||high||)); inf when not finite."""
    difference = model.combine((1.0, high), (-1.0, low))
    if error_norm is not None:
        squares = [error_norm(difference) ** 2, error_norm(old) ** 2, error_norm(high) ** 2]
    else:
        squares = [model.inner(difference, difference), model.inner(old, old), model.inner(high, high)]
    if not np.isfinite(squares).all() or min(squares) < 0:
        return float("inf")
    scale = controller.atol + controller.rtol * max(np.sqrt(squares[1]), np.sqrt(squares[2]))
    return float(np.sqrt(squares[0]) / scale)


def _finite(model: Any, field: Any) -> bool:
    check = getattr(model, "all_finite", None)
    return bool(check(field)) if check is not None else bool(np.isfinite(field).all())


def integrate_adaptive(model: Any, scheme: Any, initial: Any, T: float, initial_step: float, *,
                       controller: IController = IController(), snapshots: list[float] | None = None,
                       error_norm: Callable[[Any], float] | None = None, strict_snapshots: bool = False,
                       progress: Callable[[int, float, float], None] | None = None,
                       on_accept: Callable[[Any], None] | None = None,
                       scalar_control: ScalarErrorControl | None = None,
                       min_step_policy: Literal["fail", "force"] = "fail", retain_history: bool = True,
                       continuation: AdaptiveContinuation | None = None) -> AdaptiveResult:
    """Advance `initial` by T with embedded error control (see the module docstring)."""
    if strict_snapshots:
        raise ValueError("Strict snapshot partitions are not supported by this driver")
    if min_step_policy not in ("fail", "force"):
        raise ValueError("Unknown min_step_policy")
    if not np.isfinite([T, initial_step, initial.t]).all() or T <= 0 or initial_step <= 0:
        raise ValueError("T and initial_step must be finite and positive")
    if initial_step < controller.min_step:
        raise ValueError("initial_step is below min_step")
    end = initial.t + T
    requests = [] if snapshots is None else list(snapshots)
    if any(not np.isfinite(t) or t < initial.t or t > end for t in requests):
        raise ValueError("Snapshot time outside integration interval")
    if continuation is not None:
        if not np.isfinite(continuation.step) or not controller.min_step <= continuation.step <= controller.max_step:
            raise ValueError("Invalid continuation step")
        past = [v for v in (continuation.previous_error, continuation.previous_u, continuation.previous_r)
                if v is not None]
        if (any(not np.isfinite(v) or v < 0 for v in past) or continuation.accepted_steps < 0
                or continuation.attempted_steps < continuation.accepted_steps):
            raise ValueError("Invalid continuation history")
    best = [(abs(t - initial.t), initial) for t in requests]
    result = AdaptiveResult(initial, [initial.t], [model.diagnostics(initial)])
    saved = continuation or AdaptiveContinuation(min(initial_step, controller.max_step))
    result.continuation = saved
    step = saved.step
    previous_error, previous_u, previous_r = saved.previous_error, saved.previous_u, saved.previous_r
    scalar_controller = replace(controller, estimator_order=2)
    rejections = 0
    start, cpu_start = perf_counter(), process_time()
    while result.final.t < end:
        if saved.accepted_steps + result.accepted_steps >= controller.max_steps:
            result.status, result.error = "failed", "Maximum accepted steps exceeded"
            break
        remaining = end - result.final.t
        step = min(step, remaining)
        terminal = step == remaining
        force_floor = min_step_policy == "force" and step == controller.min_step
        if step < controller.min_step and not terminal:
            result.status, result.error = "failed", "Step below min_step"
            break
        reason = ""
        error = error_u = float("inf")
        error_r = scalar_defect = float("nan")
        trial = diagnostic = None
        try:
            trial = scheme.step(model, result.final, step)
            if trial.embedded_velocity is None:
                raise ValueError("Scheme does not provide an embedded velocity")
            high = model.vector(trial.state)
            if terminal:
                trial.state = model.state(end, high, trial.state.r)
            if not _finite(model, high) or not _finite(model, trial.embedded_velocity) or \
                    not np.isfinite(trial.state.r):
                raise FloatingPointError("Nonfinite trial state")
            error_u = embedded_error(model, model.vector(result.final), high, trial.embedded_velocity,
                                     controller, error_norm)
            error = error_u
            if scalar_control is not None:
                scalar_defect = scalar_control.defect(result.final.r, trial.state.r, step)
                error_r = abs(scalar_defect) / scalar_control.atol
                error = max(error_u, error_r) if np.isfinite(error_r) else float("inf")
            if not np.isfinite(error):
                reason = "nonfinite_error"
            elif error > 1.0 and not force_floor:
                reason = "error_exceeded"
            else:
                diagnostic = model.diagnostics(trial.state)
                if not all(np.isfinite(v) for v in diagnostic.values()):
                    raise FloatingPointError("Nonfinite diagnostic")
        except (FloatingPointError, RuntimeError, np.linalg.LinAlgError) as exc:
            reason = f"{type(exc).__name__}: {exc}"
        if reason:
            # A rejection depends on the current trial only; the accepted PI history stays.
            factor = min(IController.accepted_factor(controller, error_u, None), 0.8)
            if scalar_control is not None:
                factor = min(factor, IController.accepted_factor(
                    scalar_controller, error_r if np.isfinite(error_r) else float("inf"), None))
        else:
            factor = controller.accepted_factor(error_u, previous_u)
            if scalar_control is not None:
                factor = min(factor, scalar_controller.accepted_factor(error_r, previous_r))
        result.attempts.append({"t": float(result.final.t), "step": float(step), "error": float(error),
                                "accepted": not reason, "error_u": float(error_u), "error_r": float(error_r),
                                "scalar_defect": float(scalar_defect), "reason": reason,
                                "factor": float(factor), "boundary_partition": False,
                                "output_truncated": False,
                                "previous_accepted_error": float("nan") if previous_error is None else previous_error,
                                "previous_accepted_error_u": float("nan") if previous_u is None else previous_u,
                                "previous_accepted_error_r": float("nan") if previous_r is None else previous_r,
                                "controller": type(controller).__name__,
                                "elapsed_seconds": perf_counter() - start,
                                "cpu_seconds": process_time() - cpu_start})
        result.attempted_steps += 1
        if not retain_history:
            result.attempts[:] = result.attempts[-1:]
        if reason:
            rejections += 1
            if rejections > controller.max_rejections:
                result.status, result.error = "failed", reason or "Maximum rejections exceeded"
                break
            proposed = step * factor
            if proposed < controller.min_step:
                if min_step_policy == "force" and step > controller.min_step and not terminal:
                    step = controller.min_step
                    continue
                result.status, result.error = "failed", reason
                break
            step = proposed
            continue
        rejections = 0
        previous_error, previous_u = error, error_u
        if scalar_control is not None:
            previous_r = error_r
        result.final = trial.state
        result.accepted_steps += 1
        result.times.append(trial.state.t)
        result.diagnostics.append(diagnostic)
        result.final_stages = trial.stages
        result.stages.append([stage_diagnostics(s) for s in trial.stages])
        if not retain_history:
            result.times[:] = [initial.t, trial.state.t]
            result.diagnostics[:] = result.diagnostics[-1:]
            result.stages[:] = result.stages[-1:]
        for k, request in enumerate(requests):
            distance = abs(request - trial.state.t)
            if distance < best[k][0]:
                best[k] = (distance, trial.state)
        step = min(step * factor, controller.max_step)
        if min_step_policy == "force":
            step = max(step, controller.min_step)
        result.continuation = AdaptiveContinuation(step, previous_error, previous_u, previous_r,
                                                   saved.accepted_steps + result.accepted_steps,
                                                   saved.attempted_steps + result.attempted_steps)
        if on_accept is not None:
            try:
                on_accept(trial.state)
            except Exception as exc:  # noqa: BLE001 - reported like the reference
                result.status = "failed"
                result.error = f"Accepted-state observer failed: {type(exc).__name__}: {exc}"
                break
        if progress is not None:
            progress(result.accepted_steps, trial.state.t, perf_counter() - start)
    for request, (_, state) in zip(requests, best):
        result.snapshot_requests.append(request)
        result.snapshot_times.append(state.t)
        result.snapshots.append(state)
    result.seconds = perf_counter() - start
    return result
