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


def test_trained_warm_start_saves_iterations_and_keeps_results():
    """Held-out seed, 6 steps of sdirk3_mrsav at 128^2 (24 two-RHS stage solves): Schur-CG
    iterations and mixed-solve sweeps from the zero start (the reference), the previous
    solve's pressure, the same-stage history extrapolation, and the network's correction
    of that extrapolation (trained on 3 other seeds)."""
    from macsav import GpuTensorStokes, MACGrid
    from macsav.learned import WarmStartNet, collect_pairs, fit_alpha, train
    grid = MACGrid(128, 128)
    plain = GpuTensorStokes(grid, tolerance=1e-9)
    pairs = None
    for seed in range(3):
        model, scheme, state = _run(plain, seed, 0)
        got = collect_pairs(model, scheme, state, steps=6, dt=2e-4, period=4)
        if pairs is None:
            pairs = got
        else:
            pairs.b += got.b; pairs.base += got.base; pairs.p += got.p
    alpha = fit_alpha(grid, pairs)
    net = WarmStartNet(width=32, seed=0, alpha=alpha)
    hist = train(net, grid, pairs, epochs=60, lr=1e-3, batch=8)
    assert hist[-1] < hist[0]

    def run(start, precision="fp64", period=None):
        s = GpuTensorStokes(grid, tolerance=1e-9, precision=precision)
        s.set_warm_start(start, period=period)
        model, scheme, state = _run(s, seed=11, steps=0)
        for _ in range(6):
            state = scheme.step(model, state, 2e-4).state
        return s.total_iterations, s.sweeps, model.vector(state), s.warm_starts

    zero_its, _, zero_u, _ = run(None)
    prev_its, _, prev_u, _ = run("previous")
    hist_its, _, hist_u, _ = run("history", period=4)
    net_its, _, net_u, used = run(net, period=4)
    _, zero_sw, mixed_u, _ = run(None, "mixed")
    _, hist_sw, _, _ = run("history", "mixed", period=4)
    _, net_sw, net_mixed_u, _ = run(net, "mixed", period=4)
    print(f"held-out seed, 24 two-RHS stage solves: FP64 Schur-CG iterations zero / previous / history / "
          f"network: {zero_its} / {prev_its} / {hist_its} / {net_its}; mixed sweeps zero / history / network: "
          f"{zero_sw} / {hist_sw} / {net_sw}; training loss {hist[0]:.3g} -> {hist[-1]:.3g}")
    assert used > 0
    # Warm starts change the work, never the results (to the solve tolerance).
    for u in (prev_u, hist_u, net_u, mixed_u, net_mixed_u):
        assert np.linalg.norm(u - zero_u) <= 1e-8 * np.linalg.norm(zero_u)
    # The same-stage history saves work in both forms.
    assert hist_its < zero_its and hist_sw < zero_sw
