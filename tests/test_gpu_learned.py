"""Learned warm start: device training kernels against the NumPy oracle, and what the
trained network gains over the solve without it (results unchanged to tolerance)."""
import numpy as np
import pytest

from oracle import cnn_oracle as co
from oracle import mac_oracle as mo

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _net(width=8, seed=1):
    from macsav.learned import WarmStartNet
    net = WarmStartNet(width=width, seed=seed)
    rng = np.random.default_rng(seed)
    flat = net.get_params()
    flat += (0.02 * rng.normal(size=flat.size)).astype(np.float32)   # nonzero biases too
    net.set_params(flat)
    return net


def test_device_forward_and_gradients_match_the_oracle():
    net = _net(width=8)
    rng = np.random.default_rng(4)
    x = rng.normal(size=(2, 2, 32, 24)).astype(np.float32)
    t = rng.normal(size=(2, 1, 32, 24)).astype(np.float32)
    params = co.unflatten(net.get_params(), width=8)
    y = net.forward(torch.as_tensor(x, device="cuda")).cpu().numpy()
    for n in range(2):
        np.testing.assert_allclose(y[n], co.forward(params, x[n].astype(np.float64)), rtol=2e-4, atol=2e-5)
    loss, grad = net.loss_grad(torch.as_tensor(x, device="cuda"), torch.as_tensor(t, device="cuda"))
    # Oracle over the batch: mean over both samples' outputs.
    total, gsum = 0.0, None
    for n in range(2):
        l, g = co.loss_and_grads(params, x[n].astype(np.float64), t[n].astype(np.float64))
        flat = co.flatten([(gw, gb) for gw, gb in g]).astype(np.float64)
        total += l / 2
        gsum = flat / 2 if gsum is None else gsum + flat / 2
    assert abs(loss - total) <= 1e-4 * abs(total)
    got = grad.cpu().numpy().astype(np.float64)
    assert np.linalg.norm(got - gsum) <= 2e-4 * np.linalg.norm(gsum)


def test_training_steps_reduce_the_loss():
    net = _net(width=16, seed=2)
    rng = np.random.default_rng(5)
    x = torch.as_tensor(rng.normal(size=(4, 2, 32, 32)).astype(np.float32), device="cuda")
    t = 0.5 * x[:, :1] - 0.25 * x[:, 1:]
    losses = [net.train_step(x, t, 3e-3) for _ in range(60)]
    assert losses[-1] < 0.5 * losses[0]


def test_prepare_matches_pooling_and_rms():
    from macsav import MACGrid
    from macsav.learned import prepare
    g = MACGrid(64, 32)
    rng = np.random.default_rng(6)
    b, pp, p = (torch.as_tensor(rng.normal(size=(g.np, 2)), device="cuda") for _ in range(3))
    x, t, s = prepare(g, b, pp, p, alpha=0.3)
    pool = lambda a: a.reshape(g.ny // 4, 4, g.nx // 4, 4).mean(axis=(1, 3))  # noqa: E731
    for c in range(2):
        pb = pool(b[:, c].cpu().numpy().reshape(g.ny, g.nx))
        rms = np.sqrt(np.mean(pb**2))
        assert abs(s[c].item() - rms) <= 1e-6 * rms
        np.testing.assert_allclose(x[c, 0].cpu().numpy(), pb / rms, rtol=1e-5, atol=1e-6)
        pbase = pool(pp[:, c].cpu().numpy().reshape(g.ny, g.nx))
        np.testing.assert_allclose(x[c, 1].cpu().numpy(), pbase / np.sqrt(np.mean(pbase**2)), rtol=1e-5, atol=1e-6)
        np.testing.assert_allclose(t[c, 0].cpu().numpy(), pool(p[:, c].cpu().numpy().reshape(g.ny, g.nx)) / (0.3 * rms),
                                   rtol=1e-5, atol=1e-6)


def _run(backend, seed, steps, dt=2e-4, n=128, nu=1e-3, modes=8):
    from macsav import MACGrid
    from macsav.fields import stream_function_velocity
    from macsav.mac_ns import MACNavierStokes
    from macsav.schemes import make_scheme
    grid = MACGrid(n, n)
    model = MACNavierStokes(grid, nu, backend=backend)
    u0 = stream_function_velocity(grid, seed=seed, modes=modes)
    return model, make_scheme("sdirk3_mrsav"), model.state(0.0, u0)


def _train(grid, seeds, steps, period=None, depth=3, epochs=60):
    from macsav import GpuTensorStokes
    from macsav.learned import Pairs, WarmStartNet, collect_pairs, fit_alpha, train
    plain = GpuTensorStokes(grid, tolerance=1e-9)
    pairs = Pairs([], [], [])
    for seed in seeds:
        model, scheme, state = _run(plain, seed, 0)
        got = collect_pairs(model, scheme, state, steps=steps, dt=2e-4, period=period, depth=depth)
        pairs.b += got.b
        pairs.base += got.base
        pairs.p += got.p
    net = WarmStartNet(width=32, seed=0, alpha=fit_alpha(grid, pairs))
    hist = train(net, grid, pairs, epochs=epochs, lr=1e-3, batch=8)
    assert hist[-1] < hist[0]
    return net, hist


def _held_out(grid, start, precision="fp64", period=None, network_depth=None, steps=6):
    from macsav import GpuTensorStokes
    s = GpuTensorStokes(grid, tolerance=1e-9, precision=precision)
    s.set_warm_start(start, period=period, network_depth=network_depth)
    model, scheme, state = _run(s, seed=11, steps=0)
    for _ in range(steps):
        state = scheme.step(model, state, 2e-4).state
    return s, model.vector(state)


def test_network_start_saves_iterations_over_the_zero_start_and_keeps_results():
    """Held-out seed, 6 steps of sdirk3_mrsav at 128^2 (24 two-RHS stage solves), FP64:
    Schur-CG iterations from the zero start (the reference), the previous-stage pressure
    (Galerkin-scaled) and the network's two-term start over {previous, correction}
    (trained on 3 other seeds), whose coefficients minimise the starting residual over a
    space that contains every multiple of the base.  The network start saves iterations
    over the reference's zero start; against its previous-stage base it is measured and
    printed (it does not gain there: the iterations the preconditioned CG spends go to
    the wall modes of the Schur complement, which the pooled correction does not reach)."""
    from macsav import MACGrid
    grid = MACGrid(128, 128)
    net, hist = _train(grid, range(3), steps=6)
    zero, zero_u = _held_out(grid, None)
    prev, prev_u = _held_out(grid, "previous")
    netw, net_u = _held_out(grid, net)
    print(f"held-out seed, 24 two-RHS stage solves: FP64 Schur-CG iterations zero / previous / network: "
          f"{zero.total_iterations} / {prev.total_iterations} / {netw.total_iterations}; network starts "
          f"{netw.network_starts}; training loss {hist[0]:.3g} -> {hist[-1]:.3g}", flush=True)
    assert netw.network_starts > 0
    assert netw.total_iterations < zero.total_iterations and prev.total_iterations < zero.total_iterations
    for u in (prev_u, net_u):
        assert np.linalg.norm(u - zero_u) <= 1e-8 * np.linalg.norm(zero_u)


def test_history_and_cold_start_network_keep_results():
    """The same-stage history (depth 6) saves work in both forms; with a history period
    the network corrects the extrapolation only while fewer than `network_depth` steps
    are recorded, and every start leaves the results unchanged to the solve tolerance."""
    from macsav import MACGrid
    grid = MACGrid(128, 128)
    net, _ = _train(grid, range(3), steps=3, period=4, depth=2, epochs=40)
    zero, zero_u = _held_out(grid, None)
    hist, hist_u = _held_out(grid, "history", period=4)
    mz, mixed_u = _held_out(grid, None, "mixed")
    mh, _ = _held_out(grid, "history", "mixed", period=4)
    mn, net_mixed_u = _held_out(grid, net, "mixed", period=4, network_depth=3)
    fn, net_u = _held_out(grid, net, period=4, network_depth=3)
    print(f"FP64 iterations zero / history / network+history: {zero.total_iterations} / {hist.total_iterations} / "
          f"{fn.total_iterations}; mixed sweeps zero / history / network+history: {mz.sweeps} / {mh.sweeps} / "
          f"{mn.sweeps}; network starts FP64 {fn.network_starts}, mixed {mn.network_starts}", flush=True)
    assert 0 < fn.network_starts <= 2 * 4 and 0 < mn.network_starts <= 2 * 4
    assert hist.total_iterations < zero.total_iterations and mh.sweeps < mz.sweeps
    for u in (hist_u, net_u, mixed_u, net_mixed_u):
        assert np.linalg.norm(u - zero_u) <= 1e-8 * np.linalg.norm(zero_u)


def test_device_adam_steps_match_the_oracle():
    """Two training steps on the device against the oracle's Adam (float64) applied to the
    device's own gradients: the FP32 master weights agree to FP32 rounding."""
    net = _net(width=8, seed=3)
    rng = np.random.default_rng(8)
    x = torch.as_tensor(rng.normal(size=(2, 2, 16, 16)).astype(np.float32), device="cuda")
    t = torch.as_tensor(rng.normal(size=(2, 1, 16, 16)).astype(np.float32), device="cuda")
    w = net.get_params().astype(np.float64)
    m = np.zeros_like(w)
    v = np.zeros_like(w)
    for step in (1, 2):
        _, grad = net.loss_grad(x, t)
        g = grad.cpu().numpy().astype(np.float64)
        net.train_step(x, t, lr=1e-2)
        w, m, v = co.adam_step(w, g, m, v, step, lr=1e-2)
        got = net.get_params().astype(np.float64)
        np.testing.assert_allclose(got, w, rtol=2e-6, atol=2e-7)
