"""GPU parity of the batched cavity solve (mac_stokes_solve_batched).

Every cavity of an uneven batch must come out bitwise equal to the same cavity solved
alone by the single-cavity library path (same reductions, same GEMM arithmetic), with
the same Schur-CG iteration counts, and within the oracle bar (rel L2 <= 1e-10, equal
iteration counts) of the CPU restatement of the reference's TensorStokes.
"""
from types import SimpleNamespace

import numpy as np
import pytest

from oracle import mac_oracle as mo

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

# Uneven shapes and domains, including odd sizes and a skinny cavity.
SHAPES = [(16, 12, 1.0, 1.0), (24, 16, 1.3, 0.8), (32, 32, 1.0, 1.0), (40, 24, 2.0, 1.0),
          (12, 48, 0.5, 1.5), (64, 20, 1.0, 0.4), (33, 17, 1.1, 0.9)]


def _grids():
    from macsav import MACGrid
    return [MACGrid(nx, ny, lx, ly) for nx, ny, lx, ly in SHAPES]


def _rhs(g, seed):
    # Stage-like right-hand side: a smooth divergence-free field plus its convection.
    from macsav.fields import stream_function_velocity
    from macsav.mac.kernels import packed_convection
    u = stream_function_velocity(g, seed=seed, modes=4)
    return np.column_stack([u, 0.25 * packed_convection(g, u) + 0.1 * np.random.default_rng(seed).normal(size=g.size)])


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300))


def _shifts():
    visc = np.array([1e-3, 3e-2, 5e-8, 0.1, 2e-4, 1e-2, 7e-3])
    return np.ones(len(SHAPES)), visc


def test_batch_equals_each_cavity_alone_and_the_oracle():
    from macsav import GpuBatchedStokes, GpuTensorStokes
    grids = _grids()
    rhs = [_rhs(g, 10 + i) for i, g in enumerate(grids)]
    mass, visc = _shifts()
    batch = GpuBatchedStokes(grids, ncols=2, tolerance=1e-9)
    print(f"batch of {len(grids)} cavities, {batch.total_faces} faces, {batch.total_cells} cells", flush=True)
    out = batch.solve(rhs, mass=mass, viscosity=visc)
    for i, g in enumerate(grids):
        alone = GpuTensorStokes(g, tolerance=1e-9)
        ref = alone.solve(rhs[i], mass=mass[i], viscosity=visc[i])
        np.testing.assert_array_equal(out[i].velocity, ref.velocity)
        np.testing.assert_array_equal(out[i].pressure, ref.pressure)
        assert batch.last_iterations[i] == list(alone.last_iterations)
        assert out[i].residual == ref.residual and out[i].divergence_inf == ref.divergence_inf
        orc = mo.TensorStokes(mo.Grid(g.nx, g.ny, g.lx, g.ly), tolerance=1e-9)
        want = orc.solve(rhs[i], mass=mass[i], viscosity=visc[i])
        assert _rel(out[i].velocity, want.velocity) < 1e-10
        assert _rel(out[i].pressure, want.pressure) < 1e-10
        assert batch.last_iterations[i] == list(orc.last_iterations)


def test_launch_shapes_and_new_shifts_give_identical_results():
    from macsav import GpuBatchedStokes
    grids = _grids()
    batch = GpuBatchedStokes(grids, ncols=2, tolerance=1e-9)
    flat = torch.as_tensor(batch.pack([_rhs(g, 30 + i) for i, g in enumerate(grids)]), device="cuda:0")
    mass, visc = _shifts()
    first = batch.solve_packed(flat, mass=mass, viscosity=visc)
    base_v, base_p = first["velocity"].clone(), first["pressure"].clone()
    for shape in [(1, 1), (1, 3), (0, 1), (0, 2)]:
        batch.set_launch(*shape)
        again = batch.solve_packed(flat, mass=mass, viscosity=visc)
        assert torch.equal(again["velocity"], base_v) and torch.equal(again["pressure"], base_p), shape
        assert again["iterations"] == first["iterations"]
    # New per-cavity shifts reach the device tables (and the captured graph replays them).
    batch.set_launch(1, 1)
    other = batch.solve_packed(flat, mass=0.7 * mass, viscosity=2 * visc)
    from macsav import GpuTensorStokes
    g = grids[3]
    s = 3 * batch.ncols
    seg = [other["velocity"][batch.velocity_offsets[s + c]:batch.velocity_offsets[s + c] + g.size] for c in (0, 1)]
    alone = GpuTensorStokes(g, tolerance=1e-9).solve(
        torch.stack([flat[batch.velocity_offsets[s + c]:batch.velocity_offsets[s + c] + g.size] for c in (0, 1)], 1),
        mass=0.7 * mass[3], viscosity=2 * visc[3])
    assert torch.equal(torch.stack(seg, 1), alone.velocity)


def test_autotune_picks_a_listed_shape_and_keeps_results():
    from macsav import GpuBatchedStokes
    grids = _grids()
    batch = GpuBatchedStokes(grids, ncols=2, tolerance=1e-9)
    flat = torch.as_tensor(batch.pack([_rhs(g, 50 + i) for i, g in enumerate(grids)]), device="cuda:0")
    mass, visc = _shifts()
    ref = batch.solve_packed(flat, mass=mass, viscosity=visc)["velocity"].clone()
    timings = batch.autotune(flat, mass=mass, viscosity=visc, repeats=1)
    print("launch shape timings (s):", {k: round(v, 5) for k, v in timings.items()}, "chosen", batch.launch)
    assert batch.launch in timings
    assert torch.equal(batch.solve_packed(flat, mass=mass, viscosity=visc)["velocity"], ref)


def test_batch_failures_raise_like_the_reference():
    from macsav import GpuBatchedStokes
    grids = _grids()[:3]
    rhs = [_rhs(g, 70 + i) for i, g in enumerate(grids)]
    with pytest.raises(RuntimeError, match="Pressure Schur CG failed"):
        GpuBatchedStokes(grids, ncols=2, max_iterations=1).solve(rhs, mass=1.0, viscosity=0.1)
    b = GpuBatchedStokes(grids, ncols=2)
    with pytest.raises(ValueError):
        b.solve(rhs, mass=[1.0, -1.0, 1.0], viscosity=0.1)
    bad = [r.copy() for r in rhs]
    bad[1][3, 0] = np.nan
    with pytest.raises(ValueError):
        b.solve(bad, mass=1.0, viscosity=0.1)
    with pytest.raises(ValueError):
        b.solve(rhs[:2], mass=1.0, viscosity=0.1)
    with pytest.raises(ValueError):
        GpuBatchedStokes([SimpleNamespace(nx=2, ny=8, lx=1.0, ly=1.0)], ncols=2)
    # Still usable after the errors.
    ok = b.solve(rhs, mass=1.0, viscosity=0.1)
    assert all(r.residual < 1e-9 for r in ok)


# ---------------------------------------------------------------- mixed form
def test_mixed_batch_meets_the_bar_and_matches_the_oracle_cavity_by_cavity():
    """Mixed form (bf16x3 transforms, FP64 defect correction) on the uneven batch: every
    cavity within rel L2 1e-9 of the oracle's FP64 solve, residual and max|D u| <= 1e-10
    absolute, and as close to the oracle as the single-cavity mixed solve."""
    from macsav import GpuBatchedStokes, GpuTensorStokes
    grids = _grids()
    rhs = [_rhs(g, 50 + i) for i, g in enumerate(grids)]
    mass, visc = _shifts()
    batch = GpuBatchedStokes(grids, ncols=2, tolerance=1e-9, precision="mixed")
    print(f"mixed batch of {len(grids)} cavities", flush=True)
    out = batch.solve(rhs, mass=mass, viscosity=visc)
    print(f"sweeps {batch.last_sweeps}, inner iterations {batch.last_iterations}", flush=True)
    assert 1 <= batch.last_sweeps <= 30
    for i, g in enumerate(grids):
        orc = mo.TensorStokes(mo.Grid(g.nx, g.ny, g.lx, g.ly), tolerance=1e-9, krylov_tolerance=1e-14)
        want = orc.solve(rhs[i], mass=mass[i], viscosity=visc[i])
        assert out[i].residual <= 1e-10 and out[i].divergence_inf <= 1e-10, i
        assert _rel(out[i].velocity, want.velocity) <= 1e-9, i
        assert _rel(out[i].pressure, want.pressure) <= 1e-9, i
        single = GpuTensorStokes(g, tolerance=1e-9, precision="mixed").solve(rhs[i], mass=mass[i],
                                                                             viscosity=visc[i])
        assert _rel(out[i].velocity, single.velocity) <= 1e-9, i
        # the defect measured on the returned fields, in FP64 on the host
        d = mo.packed_divergence(np.asarray(out[i].velocity)[:, 0], mo.Grid(g.nx, g.ny, g.lx, g.ly))
        assert np.abs(d).max() <= 1e-10, i


def test_mixed_batch_launch_shapes_give_identical_results_and_finished_cavities_drop_out():
    """Graph replay or direct launches, any CG burst length: bitwise the same fields.  A
    cavity with a zero right-hand side is finished at the first defect: zero fields and no
    inner iterations, while the others sweep on."""
    from macsav import GpuBatchedStokes
    grids = _grids()
    batch = GpuBatchedStokes(grids, ncols=2, tolerance=1e-9, precision="mixed")
    rhs = [_rhs(g, 70 + i) for i, g in enumerate(grids)]
    rhs[3] = np.zeros_like(rhs[3])
    flat = torch.as_tensor(batch.pack(rhs), device="cuda:0")
    mass, visc = _shifts()
    first = batch.solve_packed(flat, mass=mass, viscosity=visc)
    base_v, base_p = first["velocity"].clone(), first["pressure"].clone()
    vo, po = batch.layout()
    assert first["iterations"][3] == [0, 0]
    assert all(sum(its) > 0 for k, its in enumerate(first["iterations"]) if k != 3)
    assert float(base_v[vo[6]:vo[8]].abs().max()) == 0.0 and float(base_p[po[6]:po[8]].abs().max()) == 0.0
    assert batch.launch == (2, 4)  # the mixed form's default: one stage graph launch per stage
    # Every launch shape, the stage graph with CG rounds shorter and longer than the inner
    # solves included, gives the same results bitwise.
    for shape in [(1, 1), (1, 3), (0, 1), (0, 2), (2, 1), (2, 3), (2, 16)]:
        batch.set_launch(*shape)
        again = batch.solve_packed(flat, mass=mass, viscosity=visc)
        assert torch.equal(again["velocity"], base_v) and torch.equal(again["pressure"], base_p), shape
        assert again["iterations"] == first["iterations"]


def test_mixed_batch_rejects_a_non_finite_right_hand_side():
    from macsav import GpuBatchedStokes
    grids = _grids()
    batch = GpuBatchedStokes(grids, ncols=2, tolerance=1e-9, precision="mixed")
    rhs = [_rhs(g, 90 + i) for i, g in enumerate(grids)]
    rhs[2][5, 1] = np.nan
    mass, visc = _shifts()
    with pytest.raises(ValueError):
        batch.solve(rhs, mass=mass, viscosity=visc)


def _step_sequence(grids, steps, seed):
    """Right-hand sides of `steps` consecutive steps with two stages each: a fixed field
    that drifts slowly from step to step, as a run's stage right-hand sides do."""
    rng = np.random.default_rng(seed)
    base = [_rhs(g, seed + i) for i, g in enumerate(grids)]
    drift = [rng.normal(size=b.shape) for b in base]
    return [[[b * (1.0 + 0.02 * k + 0.01 * stage) + 1e-3 * k * d for b, d in zip(base, drift)]
             for stage in range(2)] for k in range(steps)]


def test_mixed_batch_history_start_cuts_sweeps_and_keeps_results():
    """Same-stage history (period 2, three steps): every solve meets the bar and matches
    the oracle to 1e-9, and the warm-started solves need fewer outer sweeps than the same
    solves from zero."""
    from macsav import GpuBatchedStokes
    grids = _grids()
    mass, visc = _shifts()
    seq = _step_sequence(grids, 5, 200)
    cold = GpuBatchedStokes(grids, ncols=2, tolerance=1e-9, precision="mixed")
    warm = GpuBatchedStokes(grids, ncols=2, tolerance=1e-9, precision="mixed")
    warm.set_history(2, 3)
    sweeps_cold, sweeps_warm = [], []
    for k, step in enumerate(seq):
        for rhs in step:
            a = cold.solve(rhs, mass=mass, viscosity=visc)
            sweeps_cold.append(cold.last_sweeps)
            b = warm.solve(rhs, mass=mass, viscosity=visc)
            sweeps_warm.append(warm.last_sweeps)
            for i in range(len(grids)):
                assert b[i].residual <= 1e-10 and b[i].divergence_inf <= 1e-10, (k, i)
                assert _rel(b[i].velocity, a[i].velocity) <= 1e-9, (k, i)
        warm.commit_history()
    print(f"sweeps cold {sweeps_cold}, warm {sweeps_warm}", flush=True)
    assert sweeps_warm[:2] == sweeps_cold[:2]          # nothing recorded yet
    assert sum(sweeps_warm[4:]) < sum(sweeps_cold[4:])  # two steps or more recorded
    g = grids[2]
    orc = mo.TensorStokes(mo.Grid(g.nx, g.ny, g.lx, g.ly), tolerance=1e-9, krylov_tolerance=1e-14)
    want = orc.solve(seq[-1][-1][2], mass=mass[2], viscosity=visc[2])
    assert _rel(b[2].velocity, want.velocity) <= 1e-9 and _rel(b[2].pressure, want.pressure) <= 1e-9


def test_mixed_batch_history_forgets_a_rejected_trial():
    """Steps A (accepted), B (rejected) and C on one batch give C bitwise the fields of A
    then C on another: the history discards B."""
    from macsav import GpuBatchedStokes
    grids = _grids()
    mass, visc = _shifts()
    seq = _step_sequence(grids, 4, 300)
    rejected = _step_sequence(grids, 1, 301)[0]

    def run(with_b):
        batch = GpuBatchedStokes(grids, ncols=2, tolerance=1e-9, precision="mixed")
        batch.set_history(2, 3)
        for step in seq[:3]:
            for rhs in step:
                batch.solve(rhs, mass=mass, viscosity=visc)
            batch.commit_history()
        if with_b:
            for rhs in rejected:
                batch.solve(rhs, mass=mass, viscosity=visc)
            batch.discard_history()
        return [batch.solve(rhs, mass=mass, viscosity=visc) for rhs in seq[3]]
    first, second = run(True), run(False)
    for x, y in zip(first, second):
        for i in range(len(grids)):
            np.testing.assert_array_equal(x[i].velocity, y[i].velocity)
            np.testing.assert_array_equal(x[i].pressure, y[i].pressure)


def test_history_is_refused_for_the_fp64_batch():
    from macsav import GpuBatchedStokes
    batch = GpuBatchedStokes(_grids(), ncols=2, tolerance=1e-9)
    with pytest.raises(ValueError):
        batch.set_history(2, 3)
