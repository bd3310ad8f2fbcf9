"""FP32 FFT transform passes of the mixed form (mixed_split bit 3) against the oracle.

Where every transform length has a power-of-two FFT length (DCT-II / DST-II: n = 2^m,
DST-I: n + 1 = 2^m), the mixed handle's velocity and pressure inverses run as three FFT
passes (rows forward, column strips forward-divide-inverse, rows inverse) instead of the
split-bf16 GEMMs.  FP32 FFTs keep about 24 bits: the pressure inverse (DCT-II) agrees
with the oracle's FP64 transforms to a relative L2 of 2e-6.  The velocity inverse's DST-I
takes its odd sine sums from a prefix sum over the FFT's outputs, whose rounding grows as
sqrt(N) eps (2.7e-6 measured at N = 2048), so its bar is 1e-5 (the split GEMMs reach
~1e-5).  The mixed solve still converges to the oracle to 1e-9 with residual and max|D u|
<= 1e-10.  The event profile shows which path ran: FFT launches only, no transform GEMM.
"""
import numpy as np
import pytest

from oracle import mac_oracle as mo
from tests.test_gpu_stokes import _backend, _dev, _rel, _stage_rhs

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

SHAPES = [(16, 16, 1.0, 1.0), (64, 32, 1.3, 0.8), (256, 128, 1.0, 1.0), (1024, 1024, 1.0, 1.0),
          (2048, 512, 2.0, 0.5)]
VELOCITY_BAR = 1e-5
PRESSURE_BAR = 2e-6


def _profiled(g, **kw):
    s = _backend(g, precision="mixed", **kw)
    s.set_profiling(True)
    return s


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("mass,visc,ncols", [(1.0, 5e-8, 2), (1.0, 1e-3, 6), (0.0, 0.05, 1)])
def test_fft_velocity_inverse_matches_oracle(shape, mass, visc, ncols):
    g = mo.Grid(*shape)
    x = np.random.default_rng(21).normal(size=(g.size, ncols))
    want = mo.TensorStokes(g).velocity_inverse(x, mass, visc)
    s = _profiled(g)
    got = s.velocity_inverse(_dev(x), mass, visc).cpu().numpy().reshape(want.shape)
    prof = s.kernel_profile()
    err = _rel(got, want)
    print(f"{shape} ({mass}, {visc}) x{ncols}: FFT velocity inverse rel L2 {err:.2e}, "
          f"{prof['fft_launches']} FFT launches", flush=True)
    assert prof["fft_launches"] == 3 * ((ncols + 3) // 4) and prof["gemm_launches"] == 0
    assert err < VELOCITY_BAR
    # Column by column as well: every column of a multi-launch call is its own transform.
    for c in range(ncols):
        assert _rel(got[:, c], want[:, c]) < VELOCITY_BAR


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("ncols", [1, 2, 10])
def test_fft_pressure_inverse_matches_oracle(shape, ncols):
    g = mo.Grid(*shape)
    y = np.random.default_rng(22).normal(size=(g.np, ncols))
    y -= y.mean(axis=0)
    want = mo.TensorStokes(g).pressure_poisson_inverse(y)
    s = _profiled(g)
    got = s.pressure_poisson_inverse(_dev(y)).cpu().numpy().reshape(want.shape)
    prof = s.kernel_profile()
    err = _rel(got, want)
    print(f"{shape} x{ncols}: FFT pressure inverse rel L2 {err:.2e}", flush=True)
    assert prof["fft_launches"] == 3 * ((ncols + 7) // 8) and prof["gemm_launches"] == 0
    assert err < PRESSURE_BAR


@pytest.mark.parametrize("shape", [(7, 5, 2.0, 1.0), (48, 32, 1.0, 1.0), (33, 64, 1.0, 1.0)])
def test_lengths_without_an_fft_plan_take_the_split_gemm(shape):
    g = mo.Grid(*shape)
    x = np.random.default_rng(23).normal(size=(g.size, 2))
    want = mo.TensorStokes(g).velocity_inverse(x, 1.0, 1e-3)
    s = _profiled(g)
    got = s.velocity_inverse(_dev(x), 1.0, 1e-3).cpu().numpy()
    prof = s.kernel_profile()
    assert prof["fft_launches"] == 0 and prof["gemm_launches"] > 0
    assert _rel(got, want) < 2e-4


@pytest.mark.parametrize("shape,visc", [((64, 64, 1.0, 1.0), 2.929e-6), ((256, 256, 1.0, 1.0), 1e-4),
                                        ((1024, 1024, 1.0, 1.0), 5e-8)])
def test_mixed_solve_with_fft_passes_matches_oracle(shape, visc):
    g = mo.Grid(*shape)
    rhs = _stage_rhs(g, seed=3, modes=8 if shape[0] < 1024 else 32)
    want = mo.TensorStokes(g, tolerance=1e-10, krylov_tolerance=1e-14).solve(rhs, mass=1.0, viscosity=visc)
    s = _profiled(g, tolerance=1e-10)
    got = s.solve(rhs, mass=1.0, viscosity=visc)
    prof = s.kernel_profile()
    v, p = _rel(got.velocity, want.velocity), _rel(got.pressure, want.pressure)
    print(f"{shape}: mixed+FFT rel L2 velocity {v:.2e} pressure {p:.2e}, residual {got.residual:.2e}, "
          f"max|D u| {got.divergence_inf:.2e}, {s.last_sweeps} sweeps, {s.last_iterations} inner its", flush=True)
    assert prof["fft_launches"] > 0 and prof["gemm_launches"] == 0
    assert v < 1e-9 and p < 1e-9
    assert got.residual <= 1e-10 and got.divergence_inf <= 1e-10
