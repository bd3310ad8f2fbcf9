"""Drivers: fixed / prescribed steps and embedded-error adaptive steps.

Interface of the reference (solver/integrate.py:11-100, solver/adaptivity/controller.py:
14-398): `integrate`, `step_sizes`, `Result`; `integrate_adaptive`, `IController`,
`PIController`, `ScalarErrorControl`, `AdaptiveContinuation`, `AdaptiveResult`.  Both
drivers share one accept path (`_Run.accept`): validate the trial, take the model's
diagnostics, append the node, keep snapshot candidates, run the accepted-state
observers on a clock that excludes their time.  Failures never raise out of a run:
the run ends with status "failed", the accepted prefix and the message.

Adaptive control, from the embedded pair (docs/numerics.md; controller.py:112-123):
the error of a trial of step h is
    E = ||U - U_emb|| / (atol + rtol max(||U_n||, ||U||))
in the model's inner-product norm (or `error_norm`).  The trial is accepted when
E <= 1 (or, with `min_step_policy="force"`, at the step floor); otherwise it is
rejected and the state is unchanged.  The next step is h f with
    I:   f = clip(safety E^(-1/k), min_factor, max_factor)          (f = max_factor at E = 0)
    PI:  f = clip(safety E^(-kp/k) E_prev^(ki/k), ...), E floored, E_prev the last accepted
where k is the estimator order; a rejection uses min(I factor, 0.8) and leaves the
accepted history alone.  With `ScalarErrorControl` the SAV scalar's local defect
against r' = -gamma r, over its own atol, is a second error of order 2; the trial
needs both <= 1 and the step takes the smaller proposal.  Strict snapshots cut the
steps so that the requested times are nodes: a step that would leave a sliver below
min_step (or below half the proposal) before the output is replaced by four (or two)
equal steps.
"""
from __future__ import annotations

from collections.abc import Callable
from dataclasses import dataclass, field, replace
from time import perf_counter, process_time
from typing import Any, Literal

import numpy as np

from .core import History, Stage, State, stage_diagnostics
from .mac.grid import Array
from .schedule import nearest_node_indices, schedule_nodes

_EPS = np.finfo(float).eps


@dataclass
class Result:
    final: State
    times: list[float]
    diagnostics: list[dict[str, float]]
    stages: list[list[dict]] = field(default_factory=list)
    final_stages: list[Stage] = field(default_factory=list)
    snapshot_requests: list[float] = field(default_factory=list)
    snapshot_times: list[float] = field(default_factory=list)
    snapshots: list[State] = field(default_factory=list)
    status: str = "complete"
    error: str = ""
    seconds: float = 0.0
    observer_cpu_seconds: float = 0.0
    observer_wall_seconds: float = 0.0


def step_sizes(T: float, *, dt: float | None = None, steps: list[float] | None = None) -> Array:
    """A step sequence that sums to T: `steps` as given (checked), or dt repeated, the
    last step shortened unless T/dt is whole to 1e-12."""
    if not np.isfinite(T) or T <= 0:
        raise ValueError("T must be positive")
    if (dt is None) == (steps is None):
        raise ValueError("Specify exactly one of dt or steps")
    if steps is not None:
        seq = np.asarray(steps, dtype=float)
        if seq.ndim != 1 or seq.size == 0 or not np.isfinite(seq).all() or (seq <= 0).any():
            raise ValueError("Steps must be finite and positive")
        if not np.isclose(seq.sum(), T, rtol=1e-12, atol=1e-14 * T):
            raise ValueError("Prescribed steps must sum to T")
        return seq
    if not np.isfinite(dt) or dt <= 0:
        raise ValueError("dt must be positive")
    n_float = T / dt
    n = round(n_float)
    if n > 0 and abs(n_float - n) <= 1e-12 * max(1.0, n_float):
        return np.full(n, float(dt))
    whole = int(np.floor(n_float))
    return np.append(np.full(whole, float(dt)), T - whole * dt)


# --------------------------------------------------------------- shared run
class _Run:
    """Clock and accepted-node bookkeeping of one driver call."""

    def __init__(self, model: Any, result: Result, retain: bool = True) -> None:
        self.model, self.result, self.retain = model, result, retain
        self.t_wall, self.t_cpu = perf_counter(), process_time()

    def wall(self) -> float:
        return perf_counter() - self.t_wall - self.result.observer_wall_seconds

    def cpu(self) -> float:
        return process_time() - self.t_cpu - self.result.observer_cpu_seconds

    def observe(self, fn: Callable[..., None] | None, *args: Any) -> BaseException | None:
        """Runs an observer off the clock; returns what it raised, if anything."""
        if fn is None:
            return None
        w, c = perf_counter(), process_time()
        try:
            fn(*args)
            return None
        except Exception as exc:  # noqa: BLE001 - observers may raise anything
            return exc
        finally:
            self.result.observer_wall_seconds += perf_counter() - w
            self.result.observer_cpu_seconds += process_time() - c

    def diagnose(self, trial: Any) -> dict[str, float]:
        """Finite-state check and diagnostics of a trial (FloatingPointError if not finite)."""
        vector = self.model.vector(trial.state)
        if not _all_finite(self.model, vector) or not np.isfinite(trial.state.r):
            raise FloatingPointError("Nonfinite trial state")
        diag = self.model.diagnostics(trial.state)
        if not np.isfinite(list(diag.values())).all():
            raise FloatingPointError("Nonfinite diagnostic")
        return diag

    def accept(self, trial: Any, diag: dict[str, float]) -> None:
        res = self.result
        res.final = trial.state
        res.final_stages = trial.stages
        res.times.append(trial.state.t)
        res.diagnostics.append(diag)
        res.stages.append([stage_diagnostics(s) for s in trial.stages])
        if not self.retain:
            res.times[:] = [res.times[0], trial.state.t]
            del res.diagnostics[:-1], res.stages[:-1]

    def finish(self, picks: list[tuple[float, State]]) -> Result:
        for request, state in picks:
            self.result.snapshot_requests.append(request)
            self.result.snapshot_times.append(state.t)
            self.result.snapshots.append(state)
        self.result.seconds = self.wall()
        return self.result


def _mark(model: Any, accepted: bool) -> None:
    """Tells a model that keeps per-trial state (the backend's warm-start history) how the
    trial ended."""
    hook = getattr(model, "commit_trial" if accepted else "discard_trial", None)
    if hook is not None:
        hook()


def _all_finite(model: Any, field_: Any) -> bool:
    check = getattr(model, "all_finite", None)
    return bool(check(field_)) if check is not None else bool(np.isfinite(field_).all())


# -------------------------------------------------------------- fixed steps
def integrate(model: Any, scheme: Any, initial: State, steps: list[float] | Array,
              snapshots: list[float] | None = None,
              progress: Callable[[int, float, float], None] | None = None, *,
              on_accept: Callable[[State], None] | None = None) -> Result:
    """Advance `initial` through the given steps.  Snapshot requests map to the nearest
    node (the earlier on a tie).  Any failure ends the run with status "failed"."""
    seq = np.asarray(steps, dtype=float)
    if seq.ndim != 1 or seq.size == 0 or not np.isfinite(seq).all() or (seq <= 0).any():
        raise ValueError("Invalid step sequence")
    nodes = schedule_nodes(initial.t, seq)
    requests = list(snapshots or [])
    if any(not np.isfinite(t) or not initial.t <= t <= nodes[-1] + 1e-12 for t in requests):
        raise ValueError("Snapshot time outside integration interval")
    wanted = nearest_node_indices(nodes, requests)
    at_node: dict[int, State] = {0: initial}
    history = History(initial)
    run = _Run(model, Result(initial, [initial.t], [model.diagnostics(initial)]))
    for k, h in enumerate(seq):
        try:
            trial = scheme.step(model, history.state, float(h))
            diag = run.diagnose(trial)
            _mark(model, True)
            run.accept(trial, diag)
            history.state, history.accepted_steps = trial.state, history.accepted_steps + 1
            if k + 1 in wanted:
                at_node[k + 1] = trial.state
            failure = run.observe(on_accept, trial.state)
            if failure is not None:
                raise RuntimeError(f"Accepted-state observer failed: {type(failure).__name__}: {failure}")
            if progress is not None:
                progress(k + 1, trial.state.t, run.wall())
        except Exception as exc:  # noqa: BLE001 - a failed step ends the run
            run.result.status, run.result.error = "failed", f"{type(exc).__name__}: {exc}"
            break
    return run.finish([(r, at_node[i]) for r, i in zip(requests, wanted) if i in at_node])


# ------------------------------------------------------------ controllers
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
        if min(self.atol, self.rtol) < 0 or self.atol + self.rtol == 0:
            raise ValueError("At least one tolerance must be positive")
        if not (0 < self.safety < 1 and 0 < self.min_factor < 1 < self.max_factor):
            raise ValueError("Invalid I-controller factors")
        if self.min_step <= 0 or not self.max_step >= self.min_step:
            raise ValueError("Invalid step bounds")
        if self.max_rejections < 0 or self.max_steps <= 0:
            raise ValueError("Invalid trial limits")
        if type(self.estimator_order) is not int or self.estimator_order < 2:
            raise ValueError("Invalid estimator order")

    def _clip(self, f: float) -> float:
        return float(min(max(f, self.min_factor), self.max_factor))

    def i_factor(self, error: float) -> float:
        if error == 0:
            return self.max_factor
        return self._clip(self.safety * error ** (-1.0 / self.estimator_order))

    def accepted_factor(self, error: float, previous_error: float | None) -> float:
        return self.i_factor(error)


@dataclass(frozen=True)
class PIController(IController):
    proportional: float = 0.7
    integral: float = 0.4
    error_floor: float = 1e-14

    def __post_init__(self) -> None:
        super().__post_init__()
        if (not np.isfinite([self.proportional, self.integral, self.error_floor]).all()
                or not self.proportional > self.integral >= 0 or not 0 < self.error_floor < 1):
            raise ValueError("Invalid PI-controller parameters")

    def accepted_factor(self, error: float, previous_error: float | None) -> float:
        if previous_error is None:
            return self.i_factor(error)
        k = self.estimator_order
        now, before = max(error, self.error_floor), max(previous_error, self.error_floor)
        return self._clip(self.safety * now ** (-self.proportional / k) * before ** (self.integral / k))


@dataclass(frozen=True)
class ScalarErrorControl:
    """Local defect of the scalar against r' = -gamma r (exact: r e^(-gamma h)), scaled by
    a fixed atol.  It controls the first-order auxiliary update, not |r| itself."""
    atol: float
    gamma: float

    def __post_init__(self) -> None:
        if not np.isfinite([self.atol, self.gamma]).all() or self.atol <= 0 or self.gamma < 0:
            raise ValueError("Invalid scalar error controls")

    def defect(self, old_r: float, trial_r: float, step: float) -> float:
        return float(trial_r - np.exp(-self.gamma * step) * old_r)


@dataclass(frozen=True)
class AdaptiveContinuation:
    """Restart data at a natural node; the accepted PI history is carried over."""
    step: float
    previous_error: float | None = None
    previous_u: float | None = None
    previous_r: float | None = None
    accepted_steps: int = 0
    attempted_steps: int = 0


@dataclass
class AdaptiveResult(Result):
    """A run plus one record per trial, rejected ones included."""
    attempts: list[dict[str, Any]] = field(default_factory=list)
    accepted_steps: int = 0
    attempted_steps: int = 0
    continuation: AdaptiveContinuation | None = None


def embedded_error(model: Any, old: Any, high: Any, low: Any, controller: IController,
                   error_norm: Callable[[Any], float] | None = None) -> float:
    """E = ||high - low|| / (atol + rtol max(||old||, ||high||)); inf unless finite."""
    norm2 = (lambda x: error_norm(x) ** 2) if error_norm is not None else (lambda x: model.inner(x, x))
    d2, o2, h2 = norm2(model.combine((1.0, high), (-1.0, low))), norm2(old), norm2(high)
    if not np.isfinite([d2, o2, h2]).all() or min(d2, o2, h2) < 0:
        return float("inf")
    return float(np.sqrt(d2) / (controller.atol + controller.rtol * max(np.sqrt(o2), np.sqrt(h2))))


class _Partition:
    """Step planning against strict output times: the next output, and the equal
    steps (two or four) that replace a sliver in front of it."""

    def __init__(self, outputs: list[float], start: float, min_step: float) -> None:
        self.outputs, self.min_step = outputs, min_step
        self.next = 1 if outputs and outputs[0] == start else 0
        self.step: float | None = None
        self.left = 0

    def pending(self) -> bool:
        return self.next < len(self.outputs)

    def plan(self, t: float, proposal: float, tol: float) -> tuple[str, float, bool, bool]:
        """('skip'|'cross'|'step', step, partitioned, truncated)."""
        gap = self.outputs[self.next] - t
        if abs(gap) <= tol:
            self.next += 1
            return "skip", proposal, False, False
        if gap < 0:
            return "cross", proposal, False, False
        if self.left:
            return "step", min(self.step, gap), True, False
        truncated = proposal >= gap
        sliver = gap - proposal
        for parts, limit in ((4, self.min_step), (2, 0.5 * proposal)):
            if 0 < sliver < limit and gap / parts >= self.min_step:
                self.step, self.left = gap / parts, parts
                return "step", self.step, True, truncated
        return "step", min(proposal, gap), False, truncated

    def drop(self) -> None:
        self.step, self.left = None, 0

    def reached(self, t: float) -> None:
        if self.pending() and abs(t - self.outputs[self.next]) <= 8 * _EPS * max(1.0, abs(t)):
            self.next += 1


def integrate_adaptive(model: Any, scheme: Any, initial: Any, T: float, initial_step: float, *,
                       controller: IController = IController(), snapshots: list[float] | None = None,
                       error_norm: Callable[[Any], float] | None = None, strict_snapshots: bool = False,
                       progress: Callable[[int, float, float], None] | None = None,
                       on_accept: Callable[[Any], None] | None = None,
                       scalar_control: ScalarErrorControl | None = None,
                       min_step_policy: Literal["fail", "force"] = "fail", retain_history: bool = True,
                       continuation: AdaptiveContinuation | None = None,
                       on_attempt: Callable[[dict[str, Any]], None] | None = None,
                       on_advance: Callable[[AdaptiveResult], None] | None = None) -> AdaptiveResult:
    """Advance `initial` by the duration T under embedded error control.  Snapshot
    requests are absolute times (nearest accepted node, the earlier on a tie); the
    terminal step lands on initial.t + T exactly and may be below min_step."""
    if min_step_policy not in ("fail", "force"):
        raise ValueError("Unknown min_step_policy")
    if not np.isfinite([T, initial_step, initial.t]).all() or T <= 0 or initial_step <= 0:
        raise ValueError("T and initial_step must be finite and positive")
    if initial_step < controller.min_step:
        raise ValueError("initial_step is below min_step")
    end = initial.t + T
    requests = list(snapshots or [])
    if any(not np.isfinite(t) or not initial.t <= t <= end for t in requests):
        raise ValueError("Snapshot time outside integration interval")
    if strict_snapshots and requests != sorted(set(requests)):
        raise ValueError("Strict snapshot times must be distinct and increasing")
    if continuation is not None:
        if strict_snapshots:
            raise ValueError("Continuation supports natural nodes, not strict snapshot partitions")
        if not np.isfinite(continuation.step) or not controller.min_step <= continuation.step <= controller.max_step:
            raise ValueError("Invalid continuation step")
        past = [v for v in (continuation.previous_error, continuation.previous_u, continuation.previous_r)
                if v is not None]
        if (not np.isfinite(past).all() or any(v < 0 for v in past) or continuation.accepted_steps < 0
                or continuation.attempted_steps < continuation.accepted_steps):
            raise ValueError("Invalid continuation history")
    origin = continuation or AdaptiveContinuation(min(initial_step, controller.max_step))
    result = AdaptiveResult(initial, [initial.t], [model.diagnostics(initial)], continuation=origin)
    run = _Run(model, result, retain_history)
    best = [(abs(t - initial.t), initial) for t in requests]
    part = _Partition(requests if strict_snapshots else [], initial.t, controller.min_step)
    scalar_ctl = replace(controller, estimator_order=2)
    force = min_step_policy == "force"
    h = origin.step
    prev_e, prev_u, prev_r = origin.previous_error, origin.previous_u, origin.previous_r
    rejections = 0

    def fail(message: str) -> None:
        result.status, result.error = "failed", message

    while result.final.t < end:
        if origin.accepted_steps + result.accepted_steps >= controller.max_steps:
            fail("Maximum accepted steps exceeded")
            break
        t = result.final.t
        remaining = end - t
        h = min(h, remaining)
        partitioned = truncated = False
        if part.pending():
            action, h, partitioned, truncated = part.plan(t, h, 8 * _EPS * max(abs(end), abs(t), 1.0))
            if action == "skip":
                continue
            if action == "cross":
                fail("Strict snapshot time was crossed")
                break
        terminal = h == remaining
        at_floor = force and h == controller.min_step
        if h < controller.min_step and not terminal:
            fail("Step below min_step")
            break

        # ---- one trial
        reason, trial, diag = "", None, None
        e = e_u = float("inf")
        e_r = defect_r = float("nan")
        try:
            trial = scheme.step(model, result.final, h)
            if trial.embedded_velocity is None:
                raise ValueError("Scheme does not provide an embedded velocity")
            high = model.vector(trial.state)
            if terminal:
                trial.state = model.state(end, high, trial.state.r)
            if (not _all_finite(model, high) or not _all_finite(model, trial.embedded_velocity)
                    or not np.isfinite(trial.state.r)):
                raise FloatingPointError("Nonfinite trial state")
            e_u = e = embedded_error(model, model.vector(result.final), high, trial.embedded_velocity,
                                     controller, error_norm)
            if scalar_control is not None:
                defect_r = scalar_control.defect(result.final.r, trial.state.r, h)
                e_r = abs(defect_r) / scalar_control.atol
                e = max(e_u, e_r) if np.isfinite(e_r) else float("inf")
            if not np.isfinite(e):
                reason = "nonfinite_error"
            elif e > 1.0 and not at_floor:
                reason = "error_exceeded"
            else:
                diag = model.diagnostics(trial.state)
                if not np.isfinite(list(diag.values())).all():
                    raise FloatingPointError("Nonfinite diagnostic")
        except (FloatingPointError, RuntimeError, np.linalg.LinAlgError) as exc:
            reason = f"{type(exc).__name__}: {exc}"

        # ---- step factor: rejections use the I rule on this trial alone
        if reason:
            f = min(controller.i_factor(e_u), 0.8)
            if scalar_control is not None:
                f = min(f, scalar_ctl.i_factor(e_r if np.isfinite(e_r) else float("inf")))
        else:
            f = controller.accepted_factor(e_u, prev_u)
            if scalar_control is not None:
                f = min(f, scalar_ctl.accepted_factor(e_r, prev_r))
        nan = float("nan")
        result.attempts.append({
            "t": float(t), "step": float(h), "error": float(e), "accepted": not reason, "error_u": float(e_u),
            "error_r": float(e_r), "scalar_defect": float(defect_r), "reason": reason, "factor": float(f),
            "boundary_partition": partitioned, "output_truncated": truncated,
            "previous_accepted_error": nan if prev_e is None else prev_e,
            "previous_accepted_error_u": nan if prev_u is None else prev_u,
            "previous_accepted_error_r": nan if prev_r is None else prev_r,
            "controller": type(controller).__name__, "elapsed_seconds": run.wall(), "cpu_seconds": run.cpu()})
        result.attempted_steps += 1
        if not retain_history:
            del result.attempts[:-1]
        failure = run.observe(on_attempt, result.attempts[-1])
        if failure is not None:
            fail(f"Trial observer failed: {type(failure).__name__}: {failure}")
            break

        if reason:
            _mark(model, False)
            if partitioned:
                part.drop()
            rejections += 1
            if rejections > controller.max_rejections:
                fail(reason)
                break
            if h * f < controller.min_step:
                if force and h > controller.min_step and not terminal:
                    h = controller.min_step
                    continue
                fail(reason)
                break
            h *= f
            continue

        # ---- accept
        _mark(model, True)
        rejections = 0
        prev_e, prev_u = e, e_u
        if scalar_control is not None:
            prev_r = e_r
        run.accept(trial, diag)
        result.accepted_steps += 1
        part.reached(trial.state.t)
        for k, request in enumerate(requests):
            if abs(request - trial.state.t) < best[k][0]:
                best[k] = (abs(request - trial.state.t), trial.state)
        if partitioned:
            part.left -= 1
            if part.left:
                h = part.step
            else:
                part.drop()
                h = min(h * f, controller.max_step)
        else:
            h = min(h * f, controller.max_step)
        if force and not part.left:
            h = max(h, controller.min_step)
        result.continuation = AdaptiveContinuation(h, prev_e, prev_u, prev_r,
                                                   origin.accepted_steps + result.accepted_steps,
                                                   origin.attempted_steps + result.attempted_steps)
        failure = run.observe(on_accept, trial.state)
        if failure is not None:
            fail(f"Accepted-state observer failed: {type(failure).__name__}: {failure}")
            break
        if progress is not None:
            progress(result.accepted_steps, trial.state.t, run.wall())
        failure = run.observe(on_advance, result)
        if failure is not None:
            fail(f"Advance observer failed: {type(failure).__name__}: {failure}")
            break
    return run.finish([(r, s) for r, (_, s) in zip(requests, best)])
