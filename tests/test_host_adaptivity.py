"""Adaptive SDIRK3(2) / SDIRK2(1) step control on the CPU (the reference's
tests/core/test_adaptivity.py checks, restated): embedded-difference scaling, accept /
reject bookkeeping, failure without advancing, terminal alignment below min_step, and
the controllers' factor formulas against their closed forms."""
import numpy as np
import pytest
import scipy.linalg

from oracle import mac_oracle as mo
from tests.test_host_schemes import ODEModel


def _exact(model, t, x0):
    return scipy.linalg.expm(-(model.nu * model.K + model.N) * t) @ x0


def _ode():
    m = ODEModel()
    m.force = lambda t: np.zeros(2)
    return m


@pytest.mark.parametrize("name", ["sdirk3", "sdirk3_mrsav"])
def test_embedded_difference_has_cubic_local_scaling(name):
    from macsav.schemes import make_scheme
    m = _ode()
    s = make_scheme(name)
    x0 = np.array([1.0, -0.5])
    errors = []
    for h in (0.2, 0.1, 0.05):
        trial = s.step(m, m.state(0.0, x0), h)
        assert trial.embedded_velocity is not None
        errors.append(np.linalg.norm(trial.state.x - trial.embedded_velocity))
    assert np.log2(errors[0] / errors[1]) > 2.6 and np.log2(errors[1] / errors[2]) > 2.7


@pytest.mark.parametrize("name", ["sdirk3", "sdirk3_mrsav"])
def test_controller_rejects_and_commits_only_the_high_order_state(name):
    from macsav.adaptivity import IController, integrate_adaptive
    from macsav.schemes import make_scheme
    m = _ode()
    x0 = np.array([1.0, -0.5])
    control = IController(atol=1e-7, rtol=1e-5, max_step=0.4)
    res = integrate_adaptive(m, make_scheme(name), m.state(0.0, x0), 0.5, 0.4, controller=control,
                             snapshots=[0.0, 0.25, 0.5])
    assert res.status == "complete", res.error
    assert any(not a["accepted"] for a in res.attempts)
    assert len(res.times) - 1 == sum(bool(a["accepted"]) for a in res.attempts) == res.accepted_steps
    assert all(a["error"] <= 1 for a in res.attempts if a["accepted"])
    assert res.final.t == pytest.approx(0.5)
    assert np.linalg.norm(res.final.x - _exact(m, 0.5, x0)) < 1e-5
    assert len(res.snapshot_times) == 3 and res.snapshot_times[0] == 0.0
    assert res.snapshot_times[-1] == pytest.approx(0.5)
    assert all(len(stages) == 4 for stages in res.stages)


def test_a_failed_run_never_advances_the_accepted_state():
    from macsav.adaptivity import IController, integrate_adaptive
    from macsav.schemes import make_scheme
    m = _ode()
    initial = m.state(0.0, np.array([1.0, -0.5]))
    control = IController(atol=1e-15, rtol=0.0, min_step=0.1, max_rejections=1)
    res = integrate_adaptive(m, make_scheme("sdirk3"), initial, 0.5, 0.4, controller=control)
    assert res.status == "failed"
    assert res.final is initial and res.times == [0.0]
    assert all(not a["accepted"] for a in res.attempts)


@pytest.mark.parametrize("pi", [False, True])
@pytest.mark.parametrize("name", ["sdirk2", "sdirk2_mrsav", "sdirk3", "sdirk3_mrsav"])
def test_terminal_alignment_step_may_be_below_min_step(pi, name):
    from macsav import MACGrid
    from macsav.adaptivity import IController, PIController, integrate_adaptive
    from macsav.mac_ns import MACNavierStokes
    from macsav.schemes import make_scheme
    g = MACGrid(3, 3)
    model = MACNavierStokes(g, 0.1, backend=mo.TensorStokes(mo.Grid(3, 3), tolerance=1e-9))
    initial = model.state(1.0, np.zeros(g.size))
    end = initial.t + 0.0202
    seen = []
    cls = PIController if pi else IController
    control = cls(atol=1.0, rtol=0.0, min_step=0.001, max_step=0.01,
                  estimator_order=2 if name.startswith("sdirk2") else 3)
    res = integrate_adaptive(model, make_scheme(name), initial, 0.0202, 0.01, controller=control,
                             snapshots=[end], on_accept=lambda s: seen.append(s.t))
    assert res.status == "complete", res.error
    assert len(res.attempts) == 3
    last = res.attempts[-1]
    assert last["accepted"] and last["error"] <= 1.0
    assert 0 < last["step"] < control.min_step
    assert last["step"] == end - res.times[-2]
    assert res.final.t == res.times[-1] == res.snapshot_times[-1] == seen[-1] == end
    np.testing.assert_array_equal(model.vector(res.final), np.zeros(g.size))


def test_controller_factors_follow_their_formulas_and_validate():
    from macsav.adaptivity import IController, PIController, ScalarErrorControl
    i = IController()
    assert i.accepted_factor(0.0, None) == i.max_factor
    assert i.accepted_factor(0.5, None) == pytest.approx(0.9 * 0.5 ** (-1 / 3))
    assert i.accepted_factor(1e9, None) == i.min_factor
    p = PIController()
    assert p.accepted_factor(0.5, None) == i.accepted_factor(0.5, None)
    assert p.accepted_factor(0.5, 0.25) == pytest.approx(0.9 * 0.5 ** (-0.7 / 3) * 0.25 ** (0.4 / 3))
    assert p.accepted_factor(0.0, 0.0) == pytest.approx(min(0.9 * 1e-14 ** (-0.3 / 3), p.max_factor))
    for bad in (dict(atol=0.0, rtol=0.0), dict(safety=1.5), dict(min_step=0.0), dict(estimator_order=1)):
        with pytest.raises(ValueError):
            IController(**bad)
    with pytest.raises(ValueError):
        PIController(proportional=0.3, integral=0.4)
    sc = ScalarErrorControl(atol=1e-6, gamma=2.0)
    assert sc.defect(1.0, np.exp(-2.0 * 0.1), 0.1) == pytest.approx(0.0, abs=1e-15)


def _mms_model(backend_kind):
    from macsav import MACGrid
    from macsav.mac_ns import MACNavierStokes
    from oracle import mms
    grid = MACGrid(12, 10, 1.3, 0.8)
    og = mo.Grid(12, 10, 1.3, 0.8)
    backend = (mo.DirectStokes(og, tolerance=1e-9) if backend_kind == "direct"
               else mo.TensorStokes(og, tolerance=1e-9, krylov_tolerance=1e-12))
    model = MACNavierStokes(grid, 0.1, backend=backend)
    model.force = lambda t: mms.forcing(grid, 0.1, 0.2, t)
    return model, model.state(0.0, mms.initial_velocity(grid, 0.2))


def test_tensor_and_direct_backends_take_the_same_adaptive_decisions():
    """The reference's tests/core/test_tensor_stokes_adaptive.py:12-37 on the oracle
    backends: SDIRK3 under I control rejects the first trial, then accepts two; both
    backends take the same decisions and step sizes and end at the same state."""
    from macsav.adaptivity import IController, integrate_adaptive
    from macsav.schemes import SDIRK3
    runs = []
    for kind in ("direct", "tensor"):
        model, initial = _mms_model(kind)
        res = integrate_adaptive(model, SDIRK3(), initial, 0.002, 0.002,
                                 controller=IController(atol=1e-12, rtol=1e-8, min_step=1e-7, max_step=0.002))
        assert res.status == "complete", res.error
        assert res.final.t == 0.002
        assert model.max_abs(model.apply_D(model.vector(res.final))) < 1e-11
        assert max(s.continuity_residual for s in res.final_stages) < 1e-11
        runs.append((model, res))
    (dm, direct), (tm, tensor) = runs
    assert [a["accepted"] for a in direct.attempts] == [False, True, True]
    assert [a["accepted"] for a in tensor.attempts] == [False, True, True]
    np.testing.assert_allclose([a["step"] for a in tensor.attempts], [a["step"] for a in direct.attempts],
                               rtol=2e-8)
    np.testing.assert_allclose(tm.vector(tensor.final), dm.vector(direct.final), rtol=2e-11, atol=2e-11)


def test_strict_snapshots_become_nodes_and_slivers_are_split():
    from macsav.adaptivity import IController, integrate_adaptive
    from macsav.schemes import make_scheme
    m = _ode()
    x0 = np.array([1.0, -0.5])
    outputs = [0.0, 0.13, 0.2001, 0.5]
    control = IController(atol=1e-6, rtol=1e-4, min_step=1e-3, max_step=0.1)
    res = integrate_adaptive(m, make_scheme("sdirk3"), m.state(0.0, x0), 0.5, 0.05, controller=control,
                             snapshots=outputs, strict_snapshots=True)
    assert res.status == "complete", res.error
    for t in outputs:
        assert min(abs(np.array(res.times) - t)) <= 1e-15
    assert res.snapshot_times == pytest.approx(outputs, abs=1e-15)
    parts = [a for a in res.attempts if a["boundary_partition"]]
    assert parts, "a sliver in front of an output is replaced by equal partition steps"
    with pytest.raises(ValueError):
        integrate_adaptive(m, make_scheme("sdirk3"), m.state(0.0, x0), 0.5, 0.05, snapshots=[0.3, 0.2],
                           strict_snapshots=True)


def test_continuation_resumes_with_the_controller_history():
    from macsav.adaptivity import PIController, integrate_adaptive
    from macsav.schemes import make_scheme
    m = _ode()
    x0 = np.array([1.0, -0.5])
    control = PIController(atol=1e-7, rtol=1e-5, max_step=0.2)
    whole = integrate_adaptive(m, make_scheme("sdirk3"), m.state(0.0, x0), 0.6, 0.1, controller=control)
    first = integrate_adaptive(m, make_scheme("sdirk3"), m.state(0.0, x0), 0.3, 0.1, controller=control)
    assert first.continuation.accepted_steps == first.accepted_steps
    rest = integrate_adaptive(m, make_scheme("sdirk3"), first.final, 0.3, first.continuation.step,
                              controller=control, continuation=first.continuation)
    assert rest.status == "complete" and rest.final.t == pytest.approx(0.6)
    assert rest.continuation.attempted_steps == first.attempted_steps + rest.attempted_steps
    assert np.linalg.norm(rest.final.x - whole.final.x) < 1e-5


def test_observer_failures_end_the_run_with_status_failed():
    from macsav.adaptivity import integrate_adaptive
    from macsav.integrate import integrate
    from macsav.schemes import make_scheme
    m = _ode()
    initial = m.state(0.0, np.array([1.0, -0.5]))

    def boom(*_):
        raise KeyError("observer")
    for kw in ("on_attempt", "on_advance", "on_accept"):
        res = integrate_adaptive(m, make_scheme("sdirk3"), initial, 0.5, 0.1, **{kw: boom})
        assert res.status == "failed" and "observer failed" in res.error and "KeyError" in res.error
    res = integrate(m, make_scheme("sdirk3"), initial, [0.1] * 3, on_accept=boom)
    assert res.status == "failed" and len(res.times) == 2 and "observer failed" in res.error
