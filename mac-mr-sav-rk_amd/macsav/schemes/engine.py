"""One stage engine for every scheme: the stage loop driven by a `Tableau`.

`StageEngine.step(model, state, dt) -> Trial` is the scheme interface of the reference
(`scheme.step`, solver/core.py:50-54) for the tables in `tableau.py`.  The engine talks
to the model only through the model seam (combine, apply_K/G/D, nonlinear, force,
solve / solve_columns, inner, max_abs, stage), so the same code drives the host MAC
model, the device-resident one and any test model.

Scalar quantities (the SAV pairings, the stage residuals and divergences) go through a
*meter*.  On a host model the meter evaluates them at once.  A model that offers
`queue_reduction` / `read_reductions` (the device MAC model) gets a deferred meter:
the reductions are queued on the device and read back in one transfer, so a stage
makes exactly one host read of its own, the SAV pairings the scalar update needs.  A
stage's gate (residual, scalar residual and continuity against 1e-8) is evaluated when
its values arrive, at the next stage's read or at the end of the step; if a later
solve fails first, the pending gate is evaluated before that failure is reported, so
the first failing check is the one raised, as when every check is immediate.

Stage record per stage (`model.stage`): the stage pressure (q/dt, or for mr-SAV
q_V/dt - q (1 - r) q_W), the momentum residual, the physical and continuity
divergences, and for mr-SAV the scalar r, its residual and the substitution history.
"""
from __future__ import annotations

from collections.abc import Callable
from typing import Any

import numpy as np

from ..core import Stage, Trial
from .sav import SavUpdate, sav_update
from .tableau import Tableau

GATE = 1e-8   # scheme-level bound on the momentum, scalar and continuity residuals


# ------------------------------------------------------------------ meters
class _ImmediateMeter:
    """Reductions evaluated as they are asked for (host models)."""

    def __init__(self, model: Any) -> None:
        self.model = model

    def dot(self, a: Any, b: Any) -> float:
        return self.model.inner(a, b)

    def maxabs(self, a: Any) -> float:
        return self.model.max_abs(a)

    def divergences(self, velocity: Any) -> tuple[float, float]:
        both = getattr(self.model, "stage_divergences", None)
        if both is not None:
            return both(velocity)
        return (self.model.physical_divergence_inf(velocity),
                self.model.max_abs(self.model.apply_D(velocity)))

    def read(self, *values: float) -> tuple[float, ...]:
        return values

    def later(self, values: tuple[float, ...], check: Callable[..., None]) -> None:
        check(*values)

    def settle(self) -> None:
        pass


class _DeferredMeter:
    """Reductions queued on the device into numbered slots; `read` fetches every slot
    queued so far in one transfer and runs the checks waiting on them."""

    def __init__(self, model: Any) -> None:
        self.model = model
        self.count = 0
        self.pending: list[tuple[tuple[int, ...], Callable[..., None]]] = []
        self.model.reset_reductions()

    def _queue(self, kind: str, a: Any, b: Any = None, scale: float = 1.0) -> int:
        slot = self.count
        self.model.queue_reduction(slot, kind, a, b, scale)
        self.count += 1
        return slot

    def dot(self, a: Any, b: Any) -> int:
        return self._queue("dot", a, b, self.model.grid.area)

    def maxabs(self, a: Any) -> int:
        return self._queue("maxabs", a)

    def divergences(self, velocity: Any) -> tuple[int, int]:
        slot = self._queue("maxabs", self.model.apply_D(velocity))
        return slot, slot

    def _fetch(self) -> np.ndarray:
        return self.model.read_reductions(self.count)

    def read(self, *slots: int) -> tuple[float, ...]:
        values = self._fetch()
        self._run_pending(values)
        return tuple(float(values[s]) for s in slots)

    def later(self, slots: tuple[int, ...], check: Callable[..., None]) -> None:
        self.pending.append((slots, check))

    def _run_pending(self, values: np.ndarray) -> None:
        waiting, self.pending = self.pending, []
        for slots, check in waiting:
            check(*(float(values[s]) for s in slots))

    def settle(self) -> None:
        if self.pending:
            self._run_pending(self._fetch())


def meter_for(model: Any) -> Any:
    if hasattr(model, "queue_reduction") and hasattr(model, "read_reductions"):
        return _DeferredMeter(model)
    return _ImmediateMeter(model)


def _finite(*values: float) -> bool:
    return bool(np.isfinite(values).all())


# ------------------------------------------------------------------ engine
class StageEngine:
    def __init__(self, name: str, tableau: Tableau, gamma: float = 1.0) -> None:
        if tableau.sav_order and (not np.isfinite(gamma) or gamma < 0):
            raise ValueError("gamma must be finite and nonnegative")
        self.name = name
        self.tableau = tableau
        self.gamma = float(gamma)

    def step(self, model: Any, state: Any, dt: float) -> Trial:
        tab = self.tableau
        if not np.isfinite(dt) or dt <= 0 or not np.isfinite(state.t):
            raise ValueError("Invalid step size or state time")
        if tab.sav_order and not np.isfinite(state.r):
            raise ValueError("Invalid scalar state")
        meter = meter_for(model)
        try:
            trial = _Step(self, model, meter, state, float(dt)).run()
            meter.settle()
        except Exception:
            meter.settle()   # a failed earlier gate is the error to report
            raise
        return trial


class _Step:
    """The data of one step in progress: stage velocities U_0..U_s, their K U, their
    explicit loads (N(U_j), lifting pairing) and forces, and the scalars r_0..r_s."""

    def __init__(self, engine: StageEngine, model: Any, meter: Any, state: Any, dt: float) -> None:
        self.e, self.m, self.meter, self.dt = engine, model, meter, dt
        self.tab = engine.tableau
        self.t0 = state.t
        self.velocity = [model.vector(state)]
        self.kvel: list[Any] = []
        self.loads = [model.nonlinear_with_lifting(self.velocity[0])]
        self.forces: list[Any] = []
        self.scalars = [float(state.r)]
        self.records: list[Stage] = []

    def run(self) -> Trial:
        tab, m, dt = self.tab, self.m, self.dt
        for i in range(tab.stages):
            self.forces.append(m.force(self.t0 + tab.nodes[i] * dt))
            weights = tab.explicit[i, :i + 1]
            force = m.combine(*zip(weights, self.forces))
            conv = m.combine(*zip(weights, [n for n, _ in self.loads]))
            lifting = float(sum(w * l for w, (_, l) in zip(weights, self.loads)))
            kappa = m.nu * dt * tab.implicit[i, i]
            history = [(-m.nu * dt * tab.implicit[i, j], self.kvel[j]) for j in range(i)]
            if tab.sav_order:
                rhs = m.combine((1.0, self.velocity[i]), (dt, force), *history)
                self._sav_stage(i, rhs, conv, lifting, kappa)
            else:
                rhs = m.combine((1.0, self.velocity[i]), (dt, force), (-dt, conv), *history)
                self._plain_stage(i, rhs, kappa)
            if i + 1 < tab.stages:
                self.loads.append(m.nonlinear_with_lifting(self.velocity[-1]))
        final = self.velocity[-1]
        r = self.scalars[-1] if tab.sav_order else 0.0
        embedded = None
        if tab.embedded is not None:
            embedded = m.combine(*[(w, u) for w, u in zip(tab.embedded, self.velocity) if w != 0.0])
        return Trial(m.state(self.t0 + dt, final, r), self.records, embedded)

    # -------------------------------------------------------------- plain
    def _plain_stage(self, i: int, rhs: Any, kappa: float) -> None:
        m, dt, meter = self.m, self.dt, self.meter
        solved = m.solve(rhs, mass=1.0, viscosity=kappa)
        u = solved.velocity
        self.velocity.append(u)
        pressure = m.combine((1.0 / dt, solved.pressure))
        later_rows = self.tab.implicit[i + 1:, i]
        ku = m.apply_K(u) if (self.tab.own_gate or np.any(later_rows != 0.0)) else None
        self.kvel.append(ku)
        record = m.stage(pressure, solved.residual, float("nan"), continuity_residual=float("nan"))
        self.records.append(record)
        phys, cont = meter.divergences(u)
        if not self.tab.own_gate:
            def fill(phys_v: float, cont_v: float) -> None:
                record.divergence_inf, record.continuity_residual = phys_v, cont_v
            meter.later((phys, cont), fill)
            return
        defect = m.combine((1.0, u), (kappa, ku), (dt, m.apply_G(pressure)), (-1.0, rhs))
        slots = (meter.maxabs(defect), meter.maxabs(rhs), phys, cont, meter.maxabs(u), meter.maxabs(pressure))

        def gate(d_v: float, rhs_v: float, phys_v: float, cont_v: float, u_v: float, p_v: float) -> None:
            residual = d_v / (1.0 + rhs_v)
            record.residual, record.divergence_inf, record.continuity_residual = residual, phys_v, cont_v
            if not _finite(residual, phys_v, cont_v, u_v, p_v) or residual > GATE or cont_v > GATE * (1.0 + u_v):
                raise RuntimeError(f"stage {i}: momentum residual {residual:.3e}, continuity {cont_v:.3e} "
                                   f"above {GATE:g}")
        meter.later(slots, gate)

    # ------------------------------------------------------------- mr-SAV
    def _sav_stage(self, i: int, rhs: Any, conv: Any, lifting: float, kappa: float) -> None:
        m, dt, meter, tab, gamma = self.m, self.dt, self.meter, self.tab, self.e.gamma
        solved = m.solve_columns([rhs, conv], mass=1.0, viscosity=kappa)
        v_vel, w_vel = solved[0].velocity, solved[1].velocity
        cv, cw, cc, ww = meter.read(meter.dot(conv, v_vel), meter.dot(conv, w_vel), meter.dot(conv, conv),
                                    meter.dot(w_vel, w_vel))
        alpha, beta = cv + lifting, cw
        # <B, W> = <B, L^-1 B> >= 0 in exact arithmetic; rounding may leave it slightly
        # negative, anything beyond that means the solve is wrong.
        if beta < 0.0 and beta < -1e-12 * (1.0 + float(np.sqrt(max(cc * ww, 0.0)))):
            raise RuntimeError("Negative Stokes response energy")
        beta = max(beta, 0.0)
        d = 1.0 + gamma * dt * tab.implicit[i, i]
        s = self.scalars[i] - gamma * dt * sum(tab.implicit[i, j] * self.scalars[j + 1] for j in range(i))
        upd = sav_update(self.scalars[i], dt, d, s, alpha, beta, order=tab.sav_order, count=tab.corrections)
        c = upd.coupling
        u = m.combine((1.0, v_vel), (-dt * c, w_vel))
        pressure = m.combine((1.0 / dt, solved[0].pressure), (-c, solved[1].pressure))
        ku = m.apply_K(u)
        defect = m.combine((1.0, u), (kappa, ku), (dt, m.apply_G(pressure)), (-1.0, rhs), (dt * c, conv))
        self.velocity.append(u)
        self.kvel.append(ku)
        self.scalars.append(upd.r)
        record = m.stage(pressure, float("nan"), float("nan"), upd.r, continuity_residual=float("nan"))
        _record_update(record, upd)
        self.records.append(record)
        phys, cont = meter.divergences(u)
        slots = (meter.maxabs(defect), meter.maxabs(rhs), meter.maxabs(conv), meter.dot(conv, u), phys, cont,
                 meter.maxabs(u))

        def gate(d_v: float, rhs_v: float, conv_v: float, cu: float, phys_v: float, cont_v: float,
                 u_v: float) -> None:
            residual = d_v / (1.0 + rhs_v + dt * abs(c) * conv_v)
            scalar = d * upd.r - s + dt * upd.factor * (cu + lifting)
            scalar_residual = abs(scalar) / (1.0 + abs(s) + abs(upd.r))
            record.residual, record.scalar_residual = residual, scalar_residual
            record.divergence_inf, record.continuity_residual = phys_v, cont_v
            if (not _finite(residual, scalar_residual, cont_v) or residual > GATE or scalar_residual > GATE
                    or cont_v > GATE * (1.0 + u_v)):
                raise RuntimeError(f"SAV stage {i}: momentum residual {residual:.3e}, scalar residual "
                                   f"{scalar_residual:.3e}, continuity {cont_v:.3e} above {GATE:g}")
        meter.later(slots, gate)


def _record_update(record: Stage, upd: SavUpdate) -> None:
    record.correction_predictor = upd.predictor
    record.corrections = list(upd.values)
    record.frozen_factor = upd.factor
