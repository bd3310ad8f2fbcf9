"""Mixed precision (bf16 MFMA transforms, FP32 accumulation, FP64 defect correction).

The bf16 transforms alone are accurate to bf16 rounding (~1e-2 relative); the mixed
solve converges to the FP64 solution: relative L2 <= 1e-9 against the oracle / the FP64
GPU solve, residual and max|D u| <= 1e-10 (absolute).
"""
import numpy as np
import pytest

from oracle import mac_oracle as mo
from tests.test_gpu_stokes import _backend, _dev, _rel, _stage_rhs

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.mark.parametrize("shape", [(7, 5, 2.0, 1.0), (64, 48, 1.0, 1.0), (257, 130, 1.3, 0.8)])
def test_bf16_velocity_inverse_is_accurate_to_bf16_rounding(shape):
    g = mo.Grid(*shape)
    x = np.random.default_rng(5).normal(size=(g.size, 2))
    want = mo.TensorStokes(g).velocity_inverse(x, 1.0, 1e-3)
    got = _backend(g, precision="mixed").velocity_inverse(_dev(x), 1.0, 1e-3).cpu().numpy()
    assert _rel(got, want) < 2e-2


@pytest.mark.parametrize("shape", [(7, 5, 2.0, 1.0), (64, 48, 1.0, 1.0), (257, 130, 1.3, 0.8)])
def test_bf16_pressure_inverse_is_accurate_to_bf16_rounding(shape):
    g = mo.Grid(*shape)
    x = np.random.default_rng(6).normal(size=(g.np, 2))
    want = mo.TensorStokes(g).pressure_poisson_inverse(x)
    got = _backend(g, precision="mixed").pressure_poisson_inverse(_dev(x)).cpu().numpy()
    assert _rel(got, want) < 2e-2


@pytest.mark.parametrize("shape", [(257, 130, 1.3, 0.8), (520, 512, 1.0, 1.0), (1024, 1024, 1.0, 1.0)])
def test_split_transforms_are_accurate_on_every_tile_path(shape):
    """The split (bf16x3) velocity and pressure inverses against the oracle at sizes whose
    launches take the 64x64, 128x64 and 8-wave 128x128 tiles: about 16 significant bits
    per operand, so well inside 2e-4 relative (plain bf16 is near 1e-2)."""
    g = mo.Grid(*shape)
    rng = np.random.default_rng(12)
    x = rng.normal(size=(g.size, 2))
    want = mo.TensorStokes(g).velocity_inverse(x, 1.0, 1e-4)
    # mixed_split 7: the split GEMMs even where the FFT passes would apply.
    got = _backend(g, precision="mixed", mixed_split=7).velocity_inverse(_dev(x), 1.0, 1e-4).cpu().numpy()
    err_v = _rel(got, want)
    y = rng.normal(size=(g.np, 2))
    y -= y.mean(axis=0)
    want_p = mo.TensorStokes(g).pressure_poisson_inverse(y)
    got_p = _backend(g, precision="mixed", mixed_split=7).pressure_poisson_inverse(_dev(y)).cpu().numpy()
    err_p = _rel(got_p, want_p)
    print(f"{shape}: split velocity inverse {err_v:.2e}, pressure inverse {err_p:.2e}")
    assert err_v < 2e-4 and err_p < 2e-4


@pytest.mark.parametrize("shape,visc", [((7, 5, 2.0, 1.0), 0.07), ((64, 64, 1.0, 1.0), 2.929e-6),
                                        ((256, 192, 1.0, 1.0), 1e-4)])
def test_mixed_solve_matches_oracle_after_refinement(shape, visc):
    g = mo.Grid(*shape)
    rhs = _stage_rhs(g, seed=3)
    want = mo.TensorStokes(g, tolerance=1e-10, krylov_tolerance=1e-14).solve(rhs, mass=1.0, viscosity=visc)
    s = _backend(g, tolerance=1e-10, precision="mixed")
    got = s.solve(rhs, mass=1.0, viscosity=visc)
    assert _rel(got.velocity, want.velocity) < 1e-9
    assert _rel(got.pressure, want.pressure) < 1e-9
    assert got.residual <= 1e-10 and got.divergence_inf <= 1e-10
    assert s.last_sweeps >= 2


def test_mixed_limit_shifts_match_direct():
    g = mo.Grid(7, 5, 2.0, 1.0)
    rhs = np.random.default_rng(19).normal(size=(g.size, 2))
    s = _backend(g, tolerance=1e-10, precision="mixed")
    for mass, visc in ((1.0, 0.07), (1.0, 0.0), (0.0, 0.07)):
        want = mo.DirectStokes(g, tolerance=1e-10).solve(rhs, mass=mass, viscosity=visc)
        got = s.solve(rhs, mass=mass, viscosity=visc)
        np.testing.assert_allclose(got.velocity, want.velocity, atol=3e-11)
        np.testing.assert_allclose(got.pressure, want.pressure, atol=3e-11)


def test_mixed_matches_fp64_gpu_at_1024():
    g = mo.Grid(1024, 1024)
    rhs = _dev(_stage_rhs(g, seed=9, modes=32))
    plain = _backend(g, tolerance=1e-10, krylov_tolerance=1e-14).solve(rhs, mass=1.0, viscosity=5e-8)
    s = _backend(g, tolerance=1e-10, precision="mixed")
    mixed = s.solve(rhs, mass=1.0, viscosity=5e-8)
    assert mixed.residual <= 1e-10 and mixed.divergence_inf <= 1e-10
    assert _rel(mixed.velocity.cpu().numpy(), plain.velocity.cpu().numpy()) < 1e-9
    assert _rel(mixed.pressure.cpu().numpy(), plain.pressure.cpu().numpy()) < 1e-9


def test_solves_run_on_the_callers_stream():
    """The wrapper hands the library torch's current stream: a solve issued inside another
    stream's context is ordered after that stream's work (the right-hand side is formed
    there and not synchronised) and equals the same solve on the default stream; the
    outputs are ready for the stream's later work without a host wait."""
    g = mo.Grid(64, 64, 1.0, 1.0)
    rhs = _stage_rhs(g, seed=5)
    s = _backend(g, tolerance=1e-10, precision="mixed")
    ref = s.solve(_dev(rhs), mass=1.0, viscosity=1e-4)
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        f = _dev(rhs) * 3.0 - _dev(rhs) * 2.0       # queued on `side`, not synchronised
        got = s.solve(f, mass=1.0, viscosity=1e-4)
        vel_norm = torch.linalg.vector_norm(got.velocity - ref.velocity)  # also on `side`
    side.synchronize()
    assert float(vel_norm) <= 1e-12 * float(torch.linalg.vector_norm(ref.velocity))
    again = s.solve(_dev(rhs), mass=1.0, viscosity=1e-4)     # back on the default stream
    assert float(torch.linalg.vector_norm(again.pressure - ref.pressure)) <= 1e-12 * float(
        torch.linalg.vector_norm(ref.pressure))
