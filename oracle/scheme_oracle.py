"""Statement-by-statement restatement of the reference's time schemes (test oracle).

TEST INFRASTRUCTURE ONLY: imported by `tests/`, never by the product package.  The
product runs every scheme through one table-driven engine (`macsav.schemes.engine`);
these functions follow the reference's own per-scheme code instead, in its statement
and combination order, so a test can run both on the same model and backend and
compare trajectories, stage records and scalars.

* `sdirk2_step`          solver/schemes/sdirk2.py:17-43
* `sdirk2_mrsav_step`    solver/schemes/sdirk2_mrsav.py:14-75
* `sdirk3_step`          solver/schemes/sdirk3.py:8-57
* `sdirk3_mrsav_step`    solver/schemes/sdirk3_mrsav.py:11-65
* `imex_euler_step`      solver/schemes/imex_euler.py:10-32
* `imex_euler_mrsav_step` solver/schemes/imex_euler_mrsav.py
* `scalar_stage`         solver/schemes/scalar_correction.py:29-52
  (its substitutions are pinned against an independent block solve in the tests)
* momentum residual      solver/momentum.py:48-52: max|defect| / (1 + max|rhs| [+ c max|conv|])
Returned trials are (t, velocity, r, stages, embedded), stages as dicts.
"""
from __future__ import annotations

import numpy as np

ETA = 1 - 1 / np.sqrt(2.0)
DELTA = -1 / np.sqrt(2.0)
D3 = np.array([[1 / 4, 0, 0, 0], [3 / 10, 1 / 4, 0, 0], [-31 / 220, 1 / 11, 1 / 4, 0],
               [31 / 198, 5 / 132, -4 / 9, 1 / 4]])
DHAT3 = np.array([[1 / 4, 0, 0, 0], [-33 / 50, 121 / 100, 0, 0], [239 / 400, -71 / 100, 5 / 16, 0],
                  [-1 / 48, -17 / 66, 235 / 528, -1 / 6]])
C3 = np.array([0.0, 1 / 4, 4 / 5, 1.0, 1.0])


def scalar_stage(z, dt, d, rhs, alpha, beta, order):
    series = {1: lambda x: 1.0, 2: lambda x: 1 + x, 3: lambda x: 1 + x + x * x}[order]
    values, factor = [], 1.0
    for _ in range(2 if order == 3 else 1):
        factor = series(z)
        response = dt * dt * factor * factor * beta
        z = (rhs - dt * factor * alpha + response) / (d + response)
        values.append(z)
    return values, factor


def _residual(m, defect, rhs, conv=None, factor=0.0):
    scale = 1 + m.max_abs(rhs)
    if conv is not None:
        scale += factor * m.max_abs(conv)
    return m.max_abs(defect) / scale


def _div(m, u):
    d = m.max_abs(m.apply_D(u))
    return d, d


def _sav_finish(m, rhs, conv, solved, lifting, dt, d, scalar_rhs, predictor, order, kappa):
    """The shared tail of every mr-SAV stage (pairings, scalar, recombination, checks)."""
    V, W = solved[0].velocity, solved[1].velocity
    alpha = m.inner(conv, V) + lifting
    beta = m.inner(conv, W)
    if beta < 0.0 and beta < -1e-12 * (1 + np.sqrt(m.inner(conv, conv) * m.inner(W, W))):
        raise RuntimeError("Negative Stokes response energy")
    beta = max(beta, 0.0)
    values, factor = scalar_stage(predictor, dt, d, scalar_rhs, alpha, beta, order)
    r = values[-1]
    coupling = factor * (1 - r)
    velocity = m.combine((1.0, V), (-dt * coupling, W))
    pressure = m.combine((1.0 / dt, solved[0].pressure), (-coupling, solved[1].pressure))
    scalar = d * r - scalar_rhs + dt * factor * (m.inner(conv, velocity) + lifting)
    scalar_residual = abs(scalar) / (1 + abs(scalar_rhs) + abs(r))
    kv = m.apply_K(velocity)
    defect = m.combine((1.0, velocity), (kappa, kv), (dt, m.apply_G(pressure)), (-1.0, rhs), (dt * coupling, conv))
    residual = _residual(m, defect, rhs, conv, dt * abs(coupling))
    phys, cont = _div(m, velocity)
    if (not np.isfinite([residual, scalar_residual, cont]).all() or residual > 1e-8 or scalar_residual > 1e-8
            or cont > 1e-8 * (1 + m.max_abs(velocity))):
        raise RuntimeError(f"SAV stage residuals {residual}, {scalar_residual}, divergence {cont}")
    stage = dict(pressure=pressure, residual=residual, divergence_inf=phys, r=r, scalar_residual=scalar_residual,
                 continuity_residual=cont, corrections=values, frozen_factor=factor, correction_predictor=predictor)
    return velocity, kv, r, stage


def sdirk2_step(m, state, dt):
    old = m.vector(state)
    n_old = m.nonlinear(old)
    f0 = m.force(state.t)
    visc = m.nu * ETA * dt
    first = m.solve(m.combine((1.0, old), (ETA * dt, f0), (-ETA * dt, n_old)), mass=1.0, viscosity=visc)
    f1 = m.force(state.t + ETA * dt)
    n1 = m.nonlinear(first.velocity)
    rhs = m.combine((1.0, first.velocity), (-m.nu * dt * (1 - 2 * ETA), m.apply_K(first.velocity)), (-dt, f0),
                    (dt * (1 - DELTA), f1), (dt, n_old), (-dt * (1 - DELTA), n1))
    second = m.solve(rhs, mass=1.0, viscosity=visc)
    stages = []
    for s in (first, second):
        phys, cont = _div(m, s.velocity)
        stages.append(dict(pressure=m.combine((1.0 / dt, s.pressure)), residual=s.residual, divergence_inf=phys,
                           continuity_residual=cont, r=0.0))
    embedded = m.combine((1.0 - 1.0 / ETA, old), (1.0 / ETA, first.velocity))
    return state.t + dt, second.velocity, 0.0, stages, embedded


def sdirk2_mrsav_step(m, state, dt, gamma=1.0):
    old = m.vector(state)
    n_old, lift_old = m.nonlinear_with_lifting(old)
    f0 = m.force(state.t)
    kappa = m.nu * ETA * dt
    d = 1 + gamma * ETA * dt

    def stage(rhs, scalar_rhs, conv, predictor, lifting):
        solved = m.solve_columns([rhs, conv], mass=1.0, viscosity=kappa)
        return _sav_finish(m, rhs, conv, solved, lifting, dt, d, scalar_rhs, predictor, 2, kappa)

    u1, k1, r1, s1 = stage(m.combine((1.0, old), (ETA * dt, f0)), state.r, m.combine((ETA, n_old)), state.r,
                           ETA * lift_old)
    n1, lift1 = m.nonlinear_with_lifting(u1)
    f1 = m.force(state.t + ETA * dt)
    rhs = m.combine((1.0, u1), (-m.nu * dt * (1 - 2 * ETA), k1), (-dt, f0), (dt * (1 - DELTA), f1))
    u2, _, r2, s2 = stage(rhs, (1 - gamma * dt * (1 - 2 * ETA)) * r1, m.combine((-1.0, n_old), (1 - DELTA, n1)),
                          r1, -lift_old + (1 - DELTA) * lift1)
    embedded = m.combine((1.0 - 1.0 / ETA, old), (1.0 / ETA, u1))
    return state.t + dt, u2, r2, [s1, s2], embedded


def _known(m, nonlinear, forces, i):
    conv = m.combine(*[(DHAT3[i, j], nonlinear[j][0]) for j in range(i + 1)])
    lifting = float(sum(DHAT3[i, j] * nonlinear[j][1] for j in range(i + 1)))
    force = m.combine(*[(DHAT3[i, j], forces[j]) for j in range(i + 1)])
    return conv, force, lifting


def _embedded3(m, vs):
    return m.combine((2 / 3, vs[0]), (-128 / 99, vs[1]), (50 / 33, vs[2]), (1 / 9, vs[3]))


def sdirk3_step(m, state, dt):
    vs = [m.vector(state)]
    linear, stages = [], []
    nonlinear = [m.nonlinear_with_lifting(vs[0])]
    forces = [m.force(state.t + C3[j] * dt) for j in range(4)]
    for i in range(4):
        conv, force, _ = _known(m, nonlinear, forces, i)
        rhs = m.combine((1.0, vs[-1]), (dt, force), (-dt, conv), *[(-m.nu * dt * D3[i, j], linear[j]) for j in range(i)])
        s = m.solve(rhs, mass=1.0, viscosity=m.nu * dt * D3[i, i])
        vs.append(s.velocity)
        if i < 3:
            linear.append(m.apply_K(s.velocity))
            nonlinear.append(m.nonlinear_with_lifting(s.velocity))
        phys, cont = _div(m, s.velocity)
        stages.append(dict(pressure=m.combine((1 / dt, s.pressure)), residual=s.residual, divergence_inf=phys,
                           continuity_residual=cont, r=0.0))
    return state.t + dt, vs[-1], 0.0, stages, _embedded3(m, vs)


def sdirk3_mrsav_step(m, state, dt, gamma=1.0):
    vs = [m.vector(state)]
    linear, scalars, stages = [], [state.r], []
    nonlinear = [m.nonlinear_with_lifting(vs[0])]
    forces = [m.force(state.t + C3[j] * dt) for j in range(4)]
    for i in range(4):
        conv, force, lifting = _known(m, nonlinear, forces, i)
        rhs = m.combine((1.0, vs[-1]), (dt, force), *[(-m.nu * dt * D3[i, j], linear[j]) for j in range(i)])
        scalar_rhs = scalars[-1] - gamma * dt * sum(D3[i, j] * scalars[j + 1] for j in range(i))
        kappa = m.nu * dt * D3[i, i]
        solved = m.solve_columns([rhs, conv], mass=1.0, viscosity=kappa)
        u, kv, r, stage = _sav_finish(m, rhs, conv, solved, lifting, dt, 1 + gamma * dt * D3[i, i], scalar_rhs,
                                      scalars[-1], 3, kappa)
        vs.append(u)
        linear.append(kv)
        if i < 3:
            nonlinear.append(m.nonlinear_with_lifting(u))
        scalars.append(r)
        stages.append(stage)
    return state.t + dt, vs[-1], scalars[-1], stages, _embedded3(m, vs)


def imex_euler_step(m, state, dt):
    old = m.vector(state)
    conv = m.nonlinear(old)
    rhs = m.combine((1.0, old), (dt, m.force(state.t)), (-dt, conv))
    s = m.solve(rhs, mass=1.0, viscosity=m.nu * dt)
    u = s.velocity
    pressure = m.combine((1 / dt, s.pressure))
    defect = m.combine((1.0, u), (m.nu * dt, m.apply_K(u)), (dt, m.apply_G(pressure)), (-1.0, rhs))
    residual = _residual(m, defect, rhs)
    phys, cont = _div(m, u)
    if residual > 1e-8 or cont > 1e-8 * (1 + m.max_abs(u)):
        raise RuntimeError(f"IMEX Euler residual {residual}, divergence {cont}")
    return state.t + dt, u, 0.0, [dict(pressure=pressure, residual=residual, divergence_inf=phys,
                                       continuity_residual=cont, r=0.0)], None


def imex_euler_mrsav_step(m, state, dt, gamma=1.0):
    old = m.vector(state)
    conv, lifting = m.nonlinear_with_lifting(old)
    rhs = m.combine((1.0, old), (dt, m.force(state.t)))
    kappa = m.nu * dt
    solved = m.solve_columns([rhs, conv], mass=1.0, viscosity=kappa)
    u, _, r, stage = _sav_finish(m, rhs, conv, solved, lifting, dt, 1 + gamma * dt, state.r, state.r, 1, kappa)
    return state.t + dt, u, r, [stage], None


STEPS = {"sdirk2": sdirk2_step, "sdirk2_mrsav": sdirk2_mrsav_step, "sdirk3": sdirk3_step,
         "sdirk3_mrsav": sdirk3_mrsav_step, "imex_euler": imex_euler_step, "imex_euler_mrsav": imex_euler_mrsav_step}
