"""GPU parity of the gfx950 stage solve against the CPU oracle (FP64 path).

Bars: D and D^T bitwise; transforms and the solve within the reference's own test
tolerances (tests/core/test_stokes.py: 3e-11 absolute on unit-scale data) and
relative L2 <= 1e-10 on realistic stage right-hand sides; Schur-CG iteration counts
equal to the oracle's (scipy cg loop).
"""
import numpy as np
import pytest

from oracle import mac_oracle as mo

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda:0")


def _backend(g, **kw):
    from macsav import GpuTensorStokes, MACGrid
    return GpuTensorStokes(MACGrid(g.nx, g.ny, g.lx, g.ly), **kw)


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))


@pytest.mark.parametrize("shape", [(7, 5, 2.0, 1.0), (11, 7, 2.0, 0.7), (130, 97, 1.3, 0.8), (1100, 1000, 1.0, 1.0)])
def test_divergence_and_transpose_are_bitwise_the_reference_kernels(shape):
    g = mo.Grid(*shape)
    rng = np.random.default_rng(79)
    vel, pre = rng.normal(size=(g.size, 2)), rng.normal(size=(g.np, 2))
    s = _backend(g)
    np.testing.assert_array_equal(s.apply_divergence(_dev(vel)).cpu().numpy(), mo.packed_divergence(vel, g))
    np.testing.assert_array_equal(s.apply_divergence_transpose(_dev(pre)).cpu().numpy(),
                                  mo.packed_divergence_transpose(pre, g))


@pytest.mark.parametrize("shape", [(7, 5, 2.0, 1.0), (64, 48, 1.0, 1.0), (257, 130, 1.3, 0.8)])
@pytest.mark.parametrize("mass,visc", [(1.0, 0.07), (1.0, 5e-8), (0.0, 0.05)])
def test_velocity_inverse_matches_oracle(shape, mass, visc):
    g = mo.Grid(*shape)
    x = np.random.default_rng(5).normal(size=(g.size, 2))
    want = mo.TensorStokes(g).velocity_inverse(x, mass, visc)
    got = _backend(g).velocity_inverse(_dev(x), mass, visc).cpu().numpy()
    assert _rel(got, want) < 1e-13


@pytest.mark.parametrize("shape", [(7, 5, 2.0, 1.0), (64, 48, 1.0, 1.0), (257, 130, 1.3, 0.8)])
def test_pressure_poisson_inverse_matches_oracle(shape):
    g = mo.Grid(*shape)
    x = np.random.default_rng(6).normal(size=(g.np, 2))
    want = mo.TensorStokes(g).pressure_poisson_inverse(x)
    got = _backend(g).pressure_poisson_inverse(_dev(x)).cpu().numpy()
    assert _rel(got, want) < 1e-12


def test_solve_matches_direct_and_oracle_for_multi_rhs_and_limit_shifts():
    g = mo.Grid(7, 5, 2.0, 1.0)
    rhs = np.random.default_rng(19).normal(size=(g.size, 2))
    gpu = _backend(g, tolerance=1e-10)
    cpu = mo.TensorStokes(g, tolerance=1e-10)
    for mass, visc in ((1.0, 0.07), (1.0, 0.0), (0.0, 0.07)):
        want = mo.DirectStokes(g, tolerance=1e-10).solve(rhs, mass=mass, viscosity=visc)
        ref = cpu.solve(rhs, mass=mass, viscosity=visc)
        got = gpu.solve(rhs, mass=mass, viscosity=visc)
        np.testing.assert_allclose(got.velocity, want.velocity, atol=3e-11)
        np.testing.assert_allclose(got.pressure, want.pressure, atol=3e-11)
        np.testing.assert_allclose(got.velocity, ref.velocity, atol=3e-11)
        assert gpu.last_iterations == cpu.last_iterations
        assert got.residual < 1e-10 and got.divergence_inf < 1e-10
        assert abs(got.pressure.mean(axis=(0, 1))).max() < 1e-12


def test_gradient_rhs_gives_zero_velocity_and_setup_is_reused():
    g = mo.Grid(8, 6)
    ops = mo.Operators(g)
    p = np.random.default_rng(23).normal(size=g.np)
    p -= p.mean()
    s = _backend(g, tolerance=1e-10)
    setup = s.setup_seconds
    for visc in (0.02, 0.05, 0.11):
        res = s.solve(ops.G @ p, mass=1.0, viscosity=visc)
        np.testing.assert_allclose(res.velocity, 0.0, atol=2e-11)
        np.testing.assert_allclose(res.pressure.ravel(), p, atol=2e-10)
    assert s.setup_seconds == setup and s.solves == 3 and s.total_iterations > 0


def test_columns_solve_as_if_alone_and_rhs_is_not_mutated():
    g = mo.Grid(11, 7, 2.0, 1.0)
    rhs = np.random.default_rng(37).normal(size=(g.size, 2))
    original = rhs.copy()
    s = _backend(g, tolerance=1e-10)
    both = s.solve(rhs, mass=0.7, viscosity=0.09)
    its = s.last_iterations
    np.testing.assert_array_equal(rhs, original)
    for c in range(2):
        one = s.solve(rhs[:, c], mass=0.7, viscosity=0.09)
        np.testing.assert_allclose(one.velocity, both.velocity[:, c], atol=1e-13)
        np.testing.assert_allclose(one.pressure, both.pressure[:, :, c], atol=1e-13)
        assert s.last_iterations == (its[c],)


def _stage_rhs(grid_mo, seed, modes=8):
    # A realistic mr-SAV stage pair: [U + dt f, eta*N(U)] for a seeded stream-function U.
    from macsav.fields import stream_function_velocity
    from macsav import MACGrid
    g = grid_mo
    U = stream_function_velocity(MACGrid(g.nx, g.ny, g.lx, g.ly), seed, modes)
    u, v = g.unpack(U)
    cu, cv = mo.convection(u, v, g.hx, g.hy)
    return np.column_stack([U, (1 - 1 / np.sqrt(2)) * g.pack(cu, cv)])


@pytest.mark.parametrize("shape,visc", [((64, 64, 1.0, 1.0), 2.929e-6), ((128, 96, 1.0, 1.0), 1e-4),
                                        ((512, 512, 1.0, 1.0), 5e-8)])
def test_stage_solve_matches_oracle_on_stage_right_hand_sides(shape, visc):
    g = mo.Grid(*shape)
    rhs = _stage_rhs(g, seed=3)
    cpu = mo.TensorStokes(g, tolerance=1e-10)
    gpu = _backend(g, tolerance=1e-10)
    want = cpu.solve(rhs, mass=1.0, viscosity=visc)
    got = gpu.solve(rhs, mass=1.0, viscosity=visc)
    assert _rel(got.velocity, want.velocity) < 1e-10
    assert _rel(got.pressure, want.pressure) < 1e-10
    # Same Krylov tolerance as the reference: the divergence is the CG residual's, so it
    # is compared with the oracle's own, not with a fixed bar.
    assert got.residual < 1e-10 and got.divergence_inf < 4 * want.divergence_inf + 1e-12
    assert abs(np.subtract(gpu.last_iterations, cpu.last_iterations)).max() <= 1


def test_tight_krylov_tolerance_passes_the_1e10_gate_at_1024():
    # Schur-CG rtol 1e-14 brings max|D u| down to the FP64 transform rounding floor
    # (~1e-10 here: D multiplies the ~1e-15-relative GEMM rounding of L^-1 D^T p by 1/h),
    # inside the reference's gate tolerance*(1 + max|u|) with tolerance 1e-10.
    g = mo.Grid(1024, 1024)
    rhs = _stage_rhs(g, seed=2, modes=32)
    s = _backend(g, tolerance=1e-10, krylov_tolerance=1e-14)
    res = s.solve(_dev(rhs), mass=1.0, viscosity=5e-8)
    vmax = float(res.velocity.abs().max())
    assert res.residual < 1e-10 and res.divergence_inf <= 1e-10 * (1 + vmax)


def test_1024_stage_solve_properties_on_device_tensors():
    from macsav import MACGrid
    g = mo.Grid(1024, 1024)
    rhs = _stage_rhs(g, seed=11, modes=32)
    s = _backend(g, tolerance=1e-10, krylov_tolerance=1e-14)
    res = s.solve(_dev(rhs), mass=1.0, viscosity=5e-8)
    assert res.residual < 1e-10 and res.divergence_inf <= 1e-10 * (1 + float(res.velocity.abs().max()))
    assert res.velocity.shape == (g.size, 2) and res.pressure.shape == (g.ny, g.nx, 2)
    # Linearity: the solve of a sum is the sum of the solves (up to the CG tolerance).
    both = s.solve(_dev(rhs[:, :1] + rhs[:, 1:]), mass=1.0, viscosity=5e-8)
    summed = (res.velocity[:, 0] + res.velocity[:, 1]).cpu().numpy()
    assert _rel(both.velocity[:, 0].cpu().numpy(), summed) < 1e-9
    # Gradient load: zero velocity, pressure returned.  The exact velocity is 0, so the
    # computed one is pure rounding: ~eps*sqrt(n)*|rhs| from the transforms, which D
    # multiplies by ~2/h.  With |rhs| ~ 10 here that floor is ~1e-9 in max|D u|, so this
    # sub-case gates divergence at 1e-8 and checks velocity and pressure directly.
    yy, xx = np.meshgrid((np.arange(g.ny) + 0.5) / g.ny, (np.arange(g.nx) + 0.5) / g.nx, indexing="ij")
    p = (np.cos(np.pi * xx) * np.cos(2 * np.pi * yy) + 0.3 * np.cos(3 * np.pi * xx) * np.cos(np.pi * yy)).ravel()
    p -= p.mean()
    grad = mo.packed_divergence_transpose(-p, g)
    res = _backend(g, tolerance=1e-8, krylov_tolerance=1e-14).solve(_dev(grad), mass=1.0, viscosity=5e-8)
    assert float(res.velocity.abs().max()) < 1e-9
    assert _rel(res.pressure.cpu().numpy().ravel(), p) < 1e-10
    del MACGrid


def test_warm_start_with_the_solution_needs_no_iterations():
    g = mo.Grid(32, 24)
    rhs = _stage_rhs(g, seed=5)
    s = _backend(g, tolerance=1e-10)
    first = s.solve(rhs, mass=1.0, viscosity=1e-3)
    again = s.solve(rhs, mass=1.0, viscosity=1e-3, pressure_guess=first.pressure.reshape(g.np, 2))
    assert max(s.last_iterations) <= 1
    assert _rel(again.velocity, first.velocity) < 1e-10


def test_failures_raise_like_the_reference():
    g = mo.Grid(16, 12)
    rhs = np.random.default_rng(2).normal(size=g.size)
    with pytest.raises(RuntimeError, match="Pressure Schur CG failed"):
        _backend(g, max_iterations=1).solve(rhs, mass=1.0, viscosity=0.1)
    s = _backend(g)
    with pytest.raises(ValueError):
        s.solve(rhs, mass=-1.0, viscosity=0.1)
    with pytest.raises(ValueError):
        s.solve(np.full(g.size, np.nan), mass=1.0, viscosity=0.1)
    with pytest.raises(ValueError):
        s.solve(rhs[:-1], mass=1.0, viscosity=0.1)
    # Gate failure: counters are updated, then RuntimeError (tensor_stokes.py:340-351).
    tight = _backend(g, tolerance=1e-300)
    with pytest.raises(RuntimeError, match="Stokes residual"):
        tight.solve(rhs, mass=1.0, viscosity=0.1)
    assert tight.solves == 1
    # A zero right-hand side takes no iterations and returns zeros.
    zero = s.solve(np.zeros((g.size, 2)), mass=1.0, viscosity=0.1)
    assert s.last_iterations == (0, 0) and not zero.velocity.any() and not zero.pressure.any()
