"""Host-side model, schemes and driver (the callers of the stage solve), on the CPU.

The schemes are checked against independent references: closed-form linear solves for
the scalar corrections, exact matrix-exponential solutions of a linear test ODE for the
orders, and the oracle Stokes backend injected into the MAC model.
"""
from dataclasses import dataclass

import numpy as np
import pytest
import scipy.linalg

from oracle import mac_oracle as mo


# ----------------------------------------------------------- a linear ODE "model"
@dataclass(frozen=True)
class ODEState:
    t: float
    x: np.ndarray
    r: float = 0.0


class ODEModel:
    """dx/dt = -nu K x - N x + f(t) in R^2 with skew N (energy-neutral, like convection);
    the 'Stokes solve' is (mass I + viscosity K)^-1 with a zero pressure."""

    def __init__(self, nu=0.7, omega=1.3):
        self.nu = nu
        self.K = np.array([[2.0, 0.5], [0.5, 1.0]])
        self.N = omega * np.array([[0.0, -1.0], [1.0, 0.0]])
        self.force = lambda t: np.array([np.cos(t), 0.5 * np.sin(2 * t)])
        self.calls = []

    def vector(self, s): return s.x
    def state(self, t, x, r=0.0): return ODEState(t, np.asarray(x, float), r)
    def nonlinear(self, x): return self.N @ x
    def nonlinear_with_lifting(self, x): return self.N @ x, 0.0
    def combine(self, *terms):
        out = np.zeros_like(np.asarray(terms[0][1], float))
        for c, f in terms:
            out = out + c * f
        return out
    def apply_K(self, x): return self.K @ x
    def apply_G(self, p): return np.zeros(2)
    def apply_D(self, x): return np.zeros(1)
    def physical_divergence_inf(self, x): return 0.0
    def inner(self, a, b): return float(np.dot(a, b))
    def max_abs(self, f): return float(np.max(np.abs(f)))

    def solve(self, rhs, *, mass, viscosity):
        from macsav.mac.stokes import StokesResult
        x = np.linalg.solve(mass * np.eye(2) + viscosity * self.K, rhs)
        return StokesResult(x, np.zeros(1), 0.0, 0.0)

    def solve_columns(self, cols, *, mass, viscosity):
        self.calls.append(len(cols))
        return [self.solve(c, mass=mass, viscosity=viscosity) for c in cols]

    def stage(self, pressure, residual, divergence_inf, r=0.0, candidates=(), root_residuals=(),
              scalar_residual=0.0, *, continuity_residual=float("nan")):
        from macsav.core import Stage
        return Stage(pressure, residual, divergence_inf, r, list(candidates), list(root_residuals),
                     scalar_residual, continuity_residual)

    def diagnostics(self, s): return {"kinetic": 0.5 * float(s.x @ s.x), "r": s.r}


def _ode_error(name, dt, T=1.0):
    from macsav.schemes import make_scheme
    m = ODEModel()
    m.force = lambda t: np.zeros(2)
    s = make_scheme(name)
    st = m.state(0.0, np.array([1.0, -0.5]))
    for _ in range(int(round(T / dt))):
        st = s.step(m, st, dt).state
    exact = scipy.linalg.expm(-(m.nu * m.K + m.N) * T) @ np.array([1.0, -0.5])
    return np.linalg.norm(st.x - exact)


@pytest.mark.parametrize("name,order", [("imex_euler", 1), ("sdirk2", 2), ("sdirk3", 3)])
def test_plain_schemes_reach_their_order_on_a_linear_ode(name, order):
    e1, e2 = _ode_error(name, 0.02), _ode_error(name, 0.01)
    assert np.log2(e1 / e2) > order - 0.15


@pytest.mark.parametrize("name,order", [("sdirk2", 2), ("sdirk2_mrsav", 2), ("sdirk3", 3),
                                        ("sdirk3_mrsav", 3)])
def test_schemes_reach_their_order_on_a_manufactured_solution(name, order):
    # Exact x(t) = (cos t, sin t); forcing f = x' + nu K x + N x; SAV scalar starts at 0.
    from macsav.schemes import make_scheme

    def error(dt):
        m = ODEModel()
        m.force = lambda t: (np.array([-np.sin(t), np.cos(t)]) + m.nu * m.K @ np.array([np.cos(t), np.sin(t)])
                             + m.N @ np.array([np.cos(t), np.sin(t)]))
        s = make_scheme(name)
        st = m.state(0.0, np.array([1.0, 0.0]))
        for _ in range(int(round(0.5 / dt))):
            st = s.step(m, st, dt).state
        return np.linalg.norm(st.x - np.array([np.cos(st.t), np.sin(st.t)]))

    e1, e2 = error(0.02), error(0.01)
    assert np.log2(e1 / e2) > order - 0.2


@pytest.mark.parametrize("order,count", [(1, 1), (2, 1), (3, 2)])
def test_scalar_updates_match_an_independent_block_solve(order, count):
    from macsav.schemes import sav_update
    from oracle import scheme_oracle as so
    L = np.array([[1.8, 0.2], [0.2, 1.4]])
    B, rhs_u = np.array([0.7, -0.4]), np.array([0.6, -0.2])
    dt, d, rhs_r, z = 0.2, 1.7, 0.23, -0.25
    U, V = np.linalg.solve(L, rhs_u), np.linalg.solve(L, B)
    res = sav_update(z, dt, d, rhs_r, B @ U, B @ V, order=order, count=count)
    oracle_values, _ = so.scalar_stage(z, dt, d, rhs_r, B @ U, B @ V, order)
    want = []
    for _ in range(count):
        q = sum(z**j for j in range(order))
        block = np.block([[L, -dt * q * B[:, None]], [dt * q * B[None, :], np.array([[d]])]])
        sol = np.linalg.solve(block, np.r_[rhs_u - dt * q * B, rhs_r])
        z = float(sol[-1])
        want.append(z)
    np.testing.assert_allclose(res.values, want, atol=2e-16, rtol=0)
    np.testing.assert_allclose(oracle_values, want, atol=2e-16, rtol=0)
    np.testing.assert_allclose(U - dt * res.coupling * V, sol[:-1], atol=2e-16, rtol=0)


@pytest.mark.parametrize("name", ["imex_euler_mrsav", "sdirk2_mrsav", "sdirk3_mrsav"])
def test_mrsav_stages_take_two_columns_and_fixed_corrections(name, monkeypatch):
    from macsav.schemes import make_scheme, scheme_info

    def forbidden(*a, **k):
        raise AssertionError("no root finding")
    monkeypatch.setattr(np, "roots", forbidden)
    m = ODEModel()
    trial = make_scheme(name, gamma=2.0).step(m, m.state(0.2, np.array([0.4, -0.2]), 0.3), 0.03)
    info = scheme_info(name)
    assert m.calls == [2] * info.n_stages
    for stage in trial.stages:
        assert len(stage.corrections) == info.n_corrections
        assert stage.r == stage.corrections[-1]


def test_registry_and_step_sizes():
    from macsav.integrate import step_sizes
    from macsav.schemes import DEFAULT_SCHEMES, SCHEMES, make_scheme, scheme_info
    assert DEFAULT_SCHEMES == ("sdirk2", "sdirk2_mrsav", "sdirk3", "sdirk3_mrsav")
    assert set(DEFAULT_SCHEMES) <= set(SCHEMES)
    assert scheme_info("sdirk3_mrsav").n_stages == 4
    with pytest.raises(ValueError):
        make_scheme("rk4")
    np.testing.assert_allclose(step_sizes(1.0, dt=0.3), [0.3, 0.3, 0.3, 0.1])
    assert step_sizes(1.0, dt=0.1).size == 10
    with pytest.raises(ValueError):
        step_sizes(1.0, steps=[0.5, 0.4])


# ------------------------------------------------------ the MAC model on the CPU
def _mac_model(n=12, backend=None, nu=0.01):
    from macsav import MACGrid
    from macsav.mac_ns import MACNavierStokes
    g = MACGrid(n, n - 2, 1.3, 0.8)
    oracle = backend or mo.TensorStokes(mo.Grid(g.nx, g.ny, g.lx, g.ly), tolerance=1e-10)
    return MACNavierStokes(g, nu, backend=oracle)


def test_product_operators_equal_the_oracle_operators():
    from macsav import MACGrid, MACOperators
    g = MACGrid(9, 7, 1.3, 0.8)
    a, b = MACOperators(g), mo.Operators(mo.Grid(9, 7, 1.3, 0.8))
    for name in ("K", "D", "G"):
        assert abs(getattr(a, name) - getattr(b, name)).max() < 1e-9


def test_mac_model_steps_with_an_injected_backend_and_keeps_divergence_free():
    from macsav.fields import stream_function_velocity
    from macsav.integrate import integrate
    from macsav.schemes import make_scheme
    m = _mac_model()
    x0 = stream_function_velocity(m.grid, seed=1, modes=4, amplitude=0.5)
    res = integrate(m, make_scheme("sdirk3_mrsav"), m.state(0.0, x0), [1e-3] * 3)
    assert res.status == "complete" and len(res.times) == 4
    assert res.diagnostics[-1]["divergence_inf"] < 1e-9
    assert all(len(stage_list) == 4 for stage_list in res.stages)
    # Without forcing, kinetic energy decays.
    assert res.diagnostics[-1]["kinetic"] < res.diagnostics[0]["kinetic"]


def test_direct_and_tensor_backends_give_the_same_trajectory():
    from macsav.fields import stream_function_velocity
    from macsav.integrate import integrate
    from macsav.schemes import make_scheme
    a = _mac_model()
    b = _mac_model(backend=mo.DirectStokes(mo.Grid(12, 10, 1.3, 0.8), tolerance=1e-10))
    x0 = stream_function_velocity(a.grid, seed=2, modes=4, amplitude=0.5)
    ra = integrate(a, make_scheme("sdirk2_mrsav"), a.state(0.0, x0), [2e-3] * 3)
    rb = integrate(b, make_scheme("sdirk2_mrsav"), b.state(0.0, x0), [2e-3] * 3)
    np.testing.assert_allclose(a.vector(ra.final), b.vector(rb.final), atol=1e-11)
    assert abs(ra.final.r - rb.final.r) < 1e-12


def test_failing_step_keeps_the_accepted_prefix():
    from macsav.integrate import integrate
    from macsav.schemes import make_scheme

    class FailsThird:
        name = "fails"

        def __init__(self):
            self.inner, self.count = make_scheme("sdirk2"), 0

        def step(self, model, state, dt):
            self.count += 1
            if self.count == 3:
                raise RuntimeError("boom")
            return self.inner.step(model, state, dt)

    m = _mac_model()
    from macsav.fields import stream_function_velocity
    x0 = stream_function_velocity(m.grid, seed=3, modes=3, amplitude=0.2)
    res = integrate(m, FailsThird(), m.state(0.0, x0), [1e-3] * 5, snapshots=[0.0, 2e-3, 4e-3])
    assert res.status == "failed" and "boom" in res.error
    assert len(res.times) == 3 and res.snapshot_times == [0.0, 2e-3]


# ------------------------------------- the engine against the per-scheme restatement
def _engine_vs_oracle(model_factory, name, steps, dt, x0, r0=0.3):
    from macsav.schemes import make_scheme
    from oracle import scheme_oracle as so
    m = model_factory()
    scheme = make_scheme(name)
    st = m.state(0.0, x0, r0)
    ost = m.state(0.0, x0, r0)
    for _ in range(steps):
        trial = scheme.step(m, st, dt)
        t, u, r, stages, emb = so.STEPS[name](m, ost, dt)
        ost = m.state(t, u, r)
        assert trial.state.t == t
        scale = max(np.abs(u).max(), 1.0)
        np.testing.assert_allclose(m.vector(trial.state), u, rtol=0, atol=1e-13 * scale)
        assert abs(trial.state.r - r) <= 1e-13 * max(abs(r), 1.0)
        if emb is None:
            assert trial.embedded_velocity is None
        else:
            np.testing.assert_allclose(trial.embedded_velocity, emb, rtol=0, atol=1e-12 * scale)
        assert len(trial.stages) == len(stages)
        for a, b in zip(trial.stages, stages):
            np.testing.assert_allclose(np.ravel(a.pressure), np.ravel(b["pressure"]), rtol=0,
                                       atol=1e-11 * max(np.abs(b["pressure"]).max(), 1.0))
            assert a.r == pytest.approx(b["r"], abs=1e-13)
            for key in ("residual", "divergence_inf", "continuity_residual"):
                # record values agree to the rounding level of the quantities they measure
                assert abs(getattr(a, key) - b[key]) <= 1e-12 * (1 + abs(b[key])) + 1e-14
            if "corrections" in b:
                np.testing.assert_allclose(a.corrections, b["corrections"], rtol=0, atol=1e-13)
                assert a.frozen_factor == pytest.approx(b["frozen_factor"], abs=1e-13)
                assert a.scalar_residual <= 1e-8
        st = trial.state


@pytest.mark.parametrize("name", ["sdirk2", "sdirk2_mrsav", "sdirk3", "sdirk3_mrsav", "imex_euler",
                                  "imex_euler_mrsav"])
def test_engine_matches_the_per_scheme_restatement_on_the_mac_model(name):
    from macsav.fields import stream_function_velocity
    factory = lambda: _mac_model(n=12, nu=0.05)  # noqa: E731
    m = factory()
    x0 = stream_function_velocity(m.grid, seed=5, modes=4, amplitude=0.5)
    _engine_vs_oracle(factory, name, steps=3, dt=2e-3, x0=x0)


@pytest.mark.parametrize("name", ["sdirk2", "sdirk2_mrsav", "sdirk3", "sdirk3_mrsav", "imex_euler",
                                  "imex_euler_mrsav"])
def test_engine_matches_the_per_scheme_restatement_on_a_forced_ode(name):
    def factory():
        m = ODEModel()
        m.force = lambda t: np.array([np.cos(3 * t), np.sin(t) - 0.2])
        return m
    _engine_vs_oracle(factory, name, steps=4, dt=0.05, x0=np.array([0.4, -0.7]), r0=0.8)
