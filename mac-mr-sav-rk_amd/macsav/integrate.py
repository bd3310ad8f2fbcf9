"""Fixed / prescribed-step driver with failure prefixes and snapshots
(reference: solver/integrate.py:11-100)."""
from __future__ import annotations

from collections.abc import Callable
from dataclasses import dataclass, field
from time import perf_counter, process_time

import numpy as np

from .core import History, Scheme, Stage, State, stage_diagnostics
from .mac.grid import Array
from .model import Model
from .schedule import nearest_node_indices, schedule_nodes


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
    """A step sequence summing to T: explicit `steps`, or `dt` repeated (with a short last
    step unless T/dt is an integer to 1e-12)."""
    if not np.isfinite(T) or T <= 0:
        raise ValueError("T must be positive")
    if (dt is None) == (steps is None):
        raise ValueError("Specify exactly one of dt or steps")
    if steps is not None:
        seq = np.asarray(steps, dtype=float)
        if seq.ndim != 1 or seq.size == 0 or not np.isfinite(seq).all() or np.any(seq <= 0):
            raise ValueError("Steps must be finite and positive")
        if not np.isclose(np.sum(seq), T, rtol=1e-12, atol=1e-14 * T):
            raise ValueError("Prescribed steps must sum to T")
        return seq
    assert dt is not None
    if not np.isfinite(dt) or dt <= 0:
        raise ValueError("dt must be positive")
    ratio = T / dt
    whole = round(ratio)
    if whole > 0 and abs(ratio - whole) <= 1e-12 * max(1, ratio):
        return np.full(whole, dt)
    count = int(np.floor(ratio))
    return np.r_[np.full(count, dt), T - count * dt]


def integrate(model: Model, scheme: Scheme, initial: State, steps: list[float] | Array,
              snapshots: list[float] | None = None,
              progress: Callable[[int, float, float], None] | None = None, *,
              on_accept: Callable[[State], None] | None = None) -> Result:
    """Advance `initial` through `steps`; any exception ends the run with status 'failed'
    and keeps the accepted prefix."""
    seq = np.asarray(steps, dtype=float)
    if seq.ndim != 1 or seq.size == 0 or np.any(seq <= 0) or not np.isfinite(seq).all():
        raise ValueError("Invalid step sequence")
    nodes = schedule_nodes(initial.t, seq)
    requests = [] if snapshots is None else list(snapshots)
    if any(not np.isfinite(t) or t < initial.t or t > nodes[-1] + 1e-12 for t in requests):
        raise ValueError("Snapshot time outside integration interval")
    wanted = nearest_node_indices(nodes, requests)
    kept: dict[int, State] = {0: initial} if 0 in wanted else {}
    history = History(initial)
    result = Result(initial, [initial.t], [model.diagnostics(initial)])
    start = perf_counter()
    for k, dt in enumerate(seq):
        try:
            trial = scheme.step(model, history.state, float(dt))
            finite = getattr(model, "all_finite", None)
            field_ok = finite(model.vector(trial.state)) if finite else np.isfinite(model.vector(trial.state)).all()
            if not field_ok or not np.isfinite(trial.state.r):
                raise FloatingPointError("Nonfinite trial state")
            diag = model.diagnostics(trial.state)
            if not all(np.isfinite(v) for v in diag.values()):
                raise FloatingPointError("Nonfinite diagnostic")
            history.state = trial.state
            history.accepted_steps += 1
            result.final = trial.state
            result.times.append(trial.state.t)
            result.diagnostics.append(diag)
            result.final_stages = trial.stages
            result.stages.append([stage_diagnostics(s) for s in trial.stages])
            if k + 1 in wanted:
                kept[k + 1] = trial.state
            if on_accept is not None:
                wall, cpu = perf_counter(), process_time()
                try:
                    on_accept(trial.state)
                except Exception as error:
                    raise RuntimeError(f"Accepted-state observer failed: {type(error).__name__}: {error}") from error
                finally:
                    result.observer_wall_seconds += perf_counter() - wall
                    result.observer_cpu_seconds += process_time() - cpu
            if progress is not None:
                progress(k + 1, trial.state.t, perf_counter() - start - result.observer_wall_seconds)
        except Exception as error:
            result.status = "failed"
            result.error = f"{type(error).__name__}: {error}"
            break
    for request, index in zip(requests, wanted):
        if index in kept:
            result.snapshot_requests.append(request)
            result.snapshot_times.append(kept[index].t)
            result.snapshots.append(kept[index])
    result.seconds = perf_counter() - start - result.observer_wall_seconds
    return result
