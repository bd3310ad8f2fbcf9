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
