"""The device-resident MAC model against the host model (the reference's operators).

Bars: convection, combination, K and G bitwise equal to the host model (same operation
order, no fused multiply-adds); D and the inner product to FP64 rounding; whole
sdirk2_mrsav / sdirk3_mrsav runs with the device model equal to the host-model runs to
1e-11 relative with identical Schur-CG iteration counts.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _models(nx=24, ny=17, lx=1.3, ly=0.8, nu=0.01):
    from macsav import GpuTensorStokes, MACGrid
    from macsav.mac_ns import MACNavierStokes
    from macsav.mac_ns_device import DeviceMACNavierStokes
    g = MACGrid(nx, ny, lx, ly)
    host = MACNavierStokes(g, nu, backend=GpuTensorStokes(g, tolerance=1e-10))
    dev = DeviceMACNavierStokes(g, nu, backend=GpuTensorStokes(g, tolerance=1e-10))
    return g, host, dev


def test_device_field_operations_match_the_host_model():
    g, host, dev = _models()
    rng = np.random.default_rng(3)
    a, b, c = rng.normal(size=(3, g.size))
    p = rng.normal(size=g.np)
    np.testing.assert_array_equal(dev.nonlinear(a).cpu().numpy(), host.nonlinear(a))
    np.testing.assert_array_equal(dev.combine((0.3, a), (-1.7, b), (2.5, c)).cpu().numpy(),
                                  host.combine((0.3, a), (-1.7, b), (2.5, c)))
    many = [(0.1 * k, x) for k, x in enumerate([a, b, c] * 4)]
    np.testing.assert_array_equal(dev.combine(*many).cpu().numpy(), host.combine(*many))
    np.testing.assert_array_equal(dev.apply_K(a).cpu().numpy(), host.apply_K(a))
    np.testing.assert_array_equal(dev.apply_G(p).cpu().numpy(), host.apply_G(p))
    np.testing.assert_allclose(dev.apply_D(a).cpu().numpy(), host.apply_D(a), rtol=0, atol=1e-12 * np.abs(a).max() / g.hx)
    assert abs(dev.inner(a, b) - host.inner(a, b)) <= 1e-13 * g.area * np.abs(a).dot(np.abs(b))
    assert dev.max_abs(a) == host.max_abs(a)
    bad = a.copy()
    bad[5] = np.nan
    assert dev.max_abs(bad) == np.inf and not dev.all_finite(bad)
    dh, dd = host.diagnostics(host.state(0.0, a, 0.4)), dev.diagnostics(dev.state(0.0, a, 0.4))
    for key, value in dh.items():
        assert abs(dd[key] - value) <= 1e-12 * max(abs(value), 1.0), key


@pytest.mark.parametrize("scheme", ["sdirk2_mrsav", "sdirk3_mrsav"])
def test_device_model_runs_equal_host_model_runs(scheme):
    from macsav.fields import stream_function_velocity
    from macsav.integrate import integrate
    from macsav.schemes import make_scheme
    g, host, dev = _models(32, 24, 1.0, 1.0)
    x0 = stream_function_velocity(g, seed=4, modes=6, amplitude=0.5)
    rh = integrate(host, make_scheme(scheme), host.state(0.0, x0), [2e-3] * 10)
    its_host = host.backend.total_iterations
    rd = integrate(dev, make_scheme(scheme), dev.state(0.0, x0), [2e-3] * 10)
    assert rh.status == rd.status == "complete", (rh.error, rd.error)
    uh, ud = host.vector(rh.final), dev.vector(rd.final).cpu().numpy()
    assert np.linalg.norm(ud - uh) <= 1e-11 * np.linalg.norm(uh)
    assert abs(rd.final.r - rh.final.r) <= 1e-12
    assert dev.backend.total_iterations == its_host
    assert rd.diagnostics[-1]["divergence_inf"] < 1e-9
