"""Whole-run parity: the 64^2 sdirk2_mrsav benchmark run (100 steps) on the GPU backend
against the same run on the CPU oracle backend.

Workload: MAC 64x64 on the unit square, nu = 0.01, sdirk2_mrsav (gamma = 1), dt = 1e-3,
100 steps (200 two-RHS stage solves), FP64, gate 1e-10, Schur-CG rtol 1e-12, initial
field from the seeded stream function with modes <= 8.  Bars: final velocity and SAV
scalar within 1e-10 relative / 1e-12 absolute of the oracle run, identical per-solve
Schur-CG iteration counts, and discretely divergence-free fields throughout.
"""
import numpy as np
import pytest

from oracle import mac_oracle as mo

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


class _Counting:
    """Wraps a backend and records every solve's iteration counts."""

    def __init__(self, inner):
        self.inner = inner
        self.iterations = []

    def solve(self, rhs, *, mass, viscosity):
        res = self.inner.solve(rhs, mass=mass, viscosity=viscosity)
        self.iterations.append(tuple(self.inner.last_iterations))
        return res


def _run(backend):
    from macsav import MACGrid
    from macsav.fields import stream_function_velocity
    from macsav.integrate import integrate
    from macsav.mac_ns import MACNavierStokes
    from macsav.schemes import make_scheme
    grid = MACGrid(64, 64)
    counted = _Counting(backend)
    model = MACNavierStokes(grid, 0.01, backend=counted)
    x0 = stream_function_velocity(grid, seed=0, modes=8)
    res = integrate(model, make_scheme("sdirk2_mrsav"), model.state(0.0, x0), [1e-3] * 100)
    return model, res, counted.iterations


def test_64_sdirk2_mrsav_run_matches_the_oracle_run():
    from macsav import GpuTensorStokes, MACGrid
    grid = MACGrid(64, 64)
    gpu_model, gpu, gpu_its = _run(GpuTensorStokes(grid, tolerance=1e-10, krylov_tolerance=1e-12))
    print(f"GPU run: status {gpu.status}, {len(gpu_its)} stage solves", flush=True)
    cpu_model, cpu, cpu_its = _run(mo.TensorStokes(mo.Grid(64, 64), tolerance=1e-10, krylov_tolerance=1e-12))
    assert gpu.status == "complete" and cpu.status == "complete"
    assert len(gpu_its) == len(cpu_its) == 200
    assert gpu_its == cpu_its
    u_gpu, u_cpu = gpu_model.vector(gpu.final), cpu_model.vector(cpu.final)
    assert np.linalg.norm(u_gpu - u_cpu) <= 1e-10 * np.linalg.norm(u_cpu)
    assert abs(gpu.final.r - cpu.final.r) <= 1e-12
    assert max(d["divergence_inf"] for d in gpu.diagnostics) < 1e-9
    for a, b in zip(gpu.diagnostics, cpu.diagnostics):
        assert abs(a["kinetic"] - b["kinetic"]) <= 1e-12 * max(abs(b["kinetic"]), 1.0)
