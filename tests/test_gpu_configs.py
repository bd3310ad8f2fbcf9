"""The workloads' own shapes on the GPU, against the oracle or exact properties.

* Batched uneven cavities (nx, ny uniform on {64, ..., 512}, nu log-uniform on
  [1e-4, 1e-2], lid velocity 1 on the top wall, sdirk2_mrsav with dt = 5e-4): the
  batched solve of 16 such cavities against the oracle, cavity by cavity, in the FP64
  and in the mixed form.
* 2048^2 (nu = 2e-4, lid plus a stream-function perturbation with modes <= 64): the
  stage solve by exact properties (linearity, a gradient load gives zero velocity and
  returns the pressure) and the absolute 1e-10 bar for the mixed and refined FP64 forms.
* Adaptive SDIRK3 step control with the GPU backend against the oracle's sparse-direct
  backend (the reference's tests/core/test_tensor_stokes_adaptive.py:12-37).
* The same-stage history forgets a rejected trial's solves.
"""
import numpy as np
import pytest

from oracle import mac_oracle as mo

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

ETA = 1 - 1 / np.sqrt(2.0)


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def _lid_stage(g, nu, dt, seed, modes, amplitude=1.0, lid=1.0):
    """First sdirk2_mrsav stage pair of a lid cavity: [U + eta dt f_lid, eta N(U)]."""
    from macsav.fields import stream_function_velocity
    from macsav.mac.kernels import lid_load, packed_convection
    u = stream_function_velocity(g, seed=seed, modes=modes, amplitude=amplitude)
    return np.column_stack([u + ETA * dt * nu * lid_load(g, lid), ETA * packed_convection(g, u)])


def test_sixteen_lid_cavities_of_the_batched_workload_match_the_oracle():
    from macsav import GpuBatchedStokes, MACGrid
    rng = np.random.default_rng(2026)
    sizes = rng.choice(np.arange(64, 513, 32), size=(16, 2))
    nu = 10.0 ** rng.uniform(-4, -2, size=16)
    dt = 5e-4
    grids = [MACGrid(int(a), int(b)) for a, b in sizes]
    rhs = [_lid_stage(g, nu[i], dt, seed=i, modes=8) for i, g in enumerate(grids)]
    visc = nu * ETA * dt
    batch = GpuBatchedStokes(grids, ncols=2, tolerance=1e-9)
    print(f"16 cavities {[tuple(s) for s in sizes]}, {batch.total_faces} faces", flush=True)
    out = batch.solve(rhs, mass=1.0, viscosity=visc)
    for i, g in enumerate(grids):
        orc = mo.TensorStokes(mo.Grid(g.nx, g.ny, g.lx, g.ly), tolerance=1e-9)
        want = orc.solve(rhs[i], mass=1.0, viscosity=visc[i])
        assert _rel(out[i].velocity, want.velocity) <= 1e-10, i
        assert _rel(out[i].pressure, want.pressure) <= 1e-10, i
        assert batch.last_iterations[i] == list(orc.last_iterations), i
        assert out[i].residual <= 1e-9 and out[i].divergence_inf <= 1e-9 * (1 + np.abs(want.velocity).max())


def test_sixteen_lid_cavities_in_the_mixed_batch_meet_the_bar_against_the_oracle():
    """The workload's arithmetic: FP64 CG with split-bf16 transforms under FP64 defect
    correction, per cavity within rel L2 1e-9 of the oracle with residual and max|D u|
    <= 1e-10 absolute."""
    from macsav import GpuBatchedStokes, MACGrid
    rng = np.random.default_rng(2027)
    sizes = rng.choice(np.arange(64, 513, 32), size=(16, 2))
    nu = 10.0 ** rng.uniform(-4, -2, size=16)
    dt = 5e-4
    grids = [MACGrid(int(a), int(b)) for a, b in sizes]
    rhs = [_lid_stage(g, nu[i], dt, seed=100 + i, modes=8) for i, g in enumerate(grids)]
    visc = nu * ETA * dt
    batch = GpuBatchedStokes(grids, ncols=2, tolerance=1e-9, precision="mixed")
    print(f"mixed batch of 16 cavities {[tuple(s) for s in sizes]}", flush=True)
    out = batch.solve(rhs, mass=1.0, viscosity=visc)
    print(f"sweeps {batch.last_sweeps}", flush=True)
    for i, g in enumerate(grids):
        orc = mo.TensorStokes(mo.Grid(g.nx, g.ny, g.lx, g.ly), tolerance=1e-9, krylov_tolerance=1e-14)
        want = orc.solve(rhs[i], mass=1.0, viscosity=visc[i])
        assert out[i].residual <= 1e-10 and out[i].divergence_inf <= 1e-10, i
        assert _rel(out[i].velocity, want.velocity) <= 1e-9, i
        assert _rel(out[i].pressure, want.pressure) <= 1e-9, i


@pytest.fixture(scope="module")
def stage2048():
    g = mo.Grid(2048, 2048)
    from macsav import MACGrid
    rhs = _lid_stage(MACGrid(2048, 2048), 2e-4, 1e-4, seed=0, modes=64, amplitude=0.05)
    return g, rhs, 2e-4 * 1e-4 / 4


@pytest.mark.parametrize("form", [dict(precision="mixed"), dict(krylov_tolerance=1e-12, fp64_refine=3)])
def test_2048_stage_solve_meets_the_absolute_bar_and_is_linear(stage2048, form):
    from tests.test_gpu_stokes import _backend
    g, rhs, kappa = stage2048
    s = _backend(g, tolerance=1e-10, **form)
    dev = torch.as_tensor(rhs, device="cuda:0")
    both = s.solve(dev, mass=1.0, viscosity=kappa)
    print(f"2048^2 {form}: residual {both.residual:.2e}, max|D u| {both.divergence_inf:.2e}, "
          f"sweeps {s.last_sweeps}, iterations {s.last_iterations}", flush=True)
    assert both.residual <= 1e-10 and both.divergence_inf <= 1e-10
    summed = s.solve(dev[:, :1] + dev[:, 1:], mass=1.0, viscosity=kappa)
    lin = (both.velocity[:, 0] + both.velocity[:, 1]).cpu().numpy()
    assert _rel(summed.velocity[:, 0].cpu().numpy(), lin) <= 1e-9


def test_2048_gradient_load_gives_zero_velocity_and_returns_the_pressure(stage2048):
    from tests.test_gpu_stokes import _backend
    g, _, kappa = stage2048
    x, y = np.meshgrid((np.arange(g.nx) + 0.5) * g.hx, (np.arange(g.ny) + 0.5) * g.hy)
    p = np.cos(3 * np.pi * x) * np.cos(2 * np.pi * y)
    p -= p.mean()
    grad = -mo.packed_divergence_transpose(p.reshape(-1), g)   # G p = -D^T p
    s = _backend(g, tolerance=1e-8, precision="mixed")
    res = s.solve(torch.as_tensor(grad, device="cuda:0"), mass=1.0, viscosity=kappa)
    assert float(res.velocity.abs().max()) <= 1e-10 * float(np.abs(grad).max())
    assert _rel(res.pressure.cpu().numpy().reshape(-1), p.reshape(-1)) <= 1e-9


def test_adaptive_decisions_with_the_gpu_backend_match_sparse_direct():
    from macsav import GpuTensorStokes
    from macsav.adaptivity import IController, integrate_adaptive
    from macsav.schemes import SDIRK3
    from tests.test_host_adaptivity import _mms_model
    runs = []
    for kind in ("direct", "gpu"):
        model, initial = _mms_model("direct")
        if kind == "gpu":
            model.backend = GpuTensorStokes(model.grid, tolerance=1e-9, krylov_tolerance=1e-12)
        res = integrate_adaptive(model, SDIRK3(), initial, 0.002, 0.002,
                                 controller=IController(atol=1e-12, rtol=1e-8, min_step=1e-7, max_step=0.002))
        assert res.status == "complete", res.error
        assert res.final.t == 0.002
        assert model.max_abs(model.apply_D(model.vector(res.final))) < 1e-11
        assert max(s.continuity_residual for s in res.final_stages) < 1e-11
        runs.append((model, res))
    (dm, direct), (gm, gpu) = runs
    assert [a["accepted"] for a in gpu.attempts] == [a["accepted"] for a in direct.attempts] == [False, True, True]
    np.testing.assert_allclose([a["step"] for a in gpu.attempts], [a["step"] for a in direct.attempts], rtol=2e-8)
    np.testing.assert_allclose(gm.vector(gpu.final), dm.vector(direct.final), rtol=2e-11, atol=2e-11)


def test_history_forgets_a_rejected_trial():
    """Steps A (accepted), B (rejected) and C on one solver give C the same guesses, hence
    the same iteration counts, as A then C on another: the history discards B."""
    from macsav import GpuTensorStokes, MACGrid
    from tests.test_gpu_stokes import _stage_rhs
    g = mo.Grid(96, 96)
    rng = np.random.default_rng(4)
    base = _stage_rhs(g, seed=5)

    def stage(k, scale):
        return torch.as_tensor(base * (1.0 + scale * k) + 1e-3 * rng.normal(size=base.shape), device="cuda:0")
    steps_a = [[stage(k, 0.01) for k in range(4)] for _ in range(3)]
    step_b = [stage(k, 0.5) for k in range(4)]
    step_c = [stage(k, 0.02) for k in range(4)]

    def run(with_b):
        s = GpuTensorStokes(MACGrid(96, 96), tolerance=1e-9)
        s.set_warm_start("history", period=4, history_steps=3)
        for step in steps_a:
            for r in step:
                s.solve(r, mass=1.0, viscosity=1e-4)
            s.commit_history()
        if with_b:
            for r in step_b:
                s.solve(r, mass=1.0, viscosity=1e-4)
            s.discard_history()
        its = []
        for r in step_c:
            s.solve(r, mass=1.0, viscosity=1e-4)
            its.append(s.last_iterations)
        return its
    assert run(True) == run(False)
