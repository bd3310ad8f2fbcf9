#!/usr/bin/env python3
"""Stage-solve throughput of the MAC shifted-Stokes solve at 1024^2 on one MI355X.

Workload (the headline configuration): MAC 1024x1024 on the unit square, nu=1e-3,
sdirk3_mrsav with dt=2e-4, so every stage solve has mass=1 and viscosity
nu*dt*D[i,i] = 5e-8, and carries the two right-hand sides [stage rhs, convection]
sharing one operator.  The stage pairs are the first-stage pairs of sdirk3_mrsav,
[U_n, D_hat[0,0] N(U_n)], for four seeded stream-function fields with modes <= 32
(synthetic data of the workload's shape); they are resident in HBM before timing.

One "step" is one 2-RHS stage solve to the bar residual <= 1e-10 and max|D u| <= 1e-10
(absolute, the reference's definitions).  The plain FP64 form reaches it with Schur-CG
rtol 1e-14; the mixed form (bf16 MFMA transforms, FP64 defect correction) stops its
outer loop there.  `value` is the default (mixed) form's stage solves per second on one
GPU; the plain FP64 figure is reported beside it.  The dominant kernel (the transform
GEMM) is timed live with HIP events on the solver's own stream over a second pass of
the same K solves.  The CPU baseline times the oracle (NumPy/SciPy restatement of the
reference path) on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
for _p in (str(ROOT), str(ROOT / "mac-mr-sav-rk_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

PEAK_TFLOPS = {"fp64": 78.6, "mixed": 2500.0}  # MI355X dense: FP64 MFMA, BF16 MFMA
NU, DT, STAGE_DIAG = 1e-3, 2e-4, 0.25            # sdirk3 D[i,i] = 1/4
CALIBRATION_BYTES = 512 << 20                    # one plain copy of known size, before timing
TRAFFIC_FILE = ROOT / "profiles" / "k_gemm_traffic_1024.md"


def calibration_copy():
    """One device copy of CALIBRATION_BYTES (read and write each), outside the timed
    region: hardware byte counters of the transform GEMM are scaled by what the same
    counters read for this copy."""
    import torch
    src = torch.ones(CALIBRATION_BYTES // 8, dtype=torch.float64, device="cuda:0")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst


def measured_traffic(precision: str):
    """HBM bytes per transform-GEMM launch from the committed counter summary (a line
    'traffic_bytes_per_launch <precision> <bytes>'), or None."""
    try:
        for line in TRAFFIC_FILE.read_text().splitlines():
            parts = line.strip().strip("`").split()
            if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch" and parts[1] == precision:
                return float(parts[2])
    except OSError:
        return None
    return None


def stage_pairs(grid, count: int, modes: int):
    import numpy as np
    from macsav.fields import stream_function_velocity
    from macsav.mac.kernels import packed_convection
    pairs = []
    for seed in range(count):
        U = stream_function_velocity(grid, seed=seed, modes=modes)
        pairs.append(np.column_stack([U, STAGE_DIAG * packed_convection(grid, U)]))
    return pairs


def cpu_baseline(grid, pair, kappa: float, budget_s: float) -> dict:
    """Oracle (scipy.fft + scipy-cg loop) on host cores, single-threaded FFTs."""
    import scipy.fft
    from oracle import mac_oracle as mo
    g = mo.Grid(grid.nx, grid.ny, grid.lx, grid.ly)
    solver = mo.TensorStokes(g, tolerance=1e-10, krylov_tolerance=1e-14)
    done, t0 = 0, time.perf_counter()
    with scipy.fft.set_workers(1):
        while True:
            solver.solve(pair, mass=1.0, viscosity=kappa)
            done += 1
            if time.perf_counter() - t0 > budget_s or done >= 20:
                break
    secs = time.perf_counter() - t0
    return {"value": done / secs, "unit": "stage solves/s", "cores": 1, "kind": "port",
            "sample": f"{done} two-RHS stage solves at {grid.nx}x{grid.ny} (oracle TensorStokes, "
                      f"scipy.fft workers=1), {secs:.1f} s"}


ETA2 = 1.0 - 0.5 ** 0.5                          # sdirk2 diagonal


def bench_cavities(args) -> dict:
    """Batched uneven cavities: `--cavities` cavities with nx, ny drawn uniformly from
    {64, 96, ..., 512}, nu log-uniform on [1e-4, 1e-2], sdirk2_mrsav with dt = 5e-4 (mass 1,
    viscosity nu*eta*dt), two right-hand sides per cavity.  One step is one batched stage
    solve of every cavity (FP64, gate 1e-9); the baseline is a loop of single-cavity
    GpuTensorStokes solves of the same right-hand sides on the same GPU."""
    import numpy as np
    import torch
    from macsav import GpuBatchedStokes, GpuTensorStokes, MACGrid
    from macsav.fields import stream_function_velocity
    from macsav.mac.kernels import packed_convection
    rng = np.random.default_rng(2026)
    n = args.cavities
    sizes = rng.choice(np.arange(64, 513, 32), size=(n, 2))
    nu = 10.0 ** rng.uniform(-4, -2, size=n)
    visc = nu * ETA2 * 5e-4
    mass = np.ones(n)
    grids = [MACGrid(int(a), int(b)) for a, b in sizes]
    rhs = []
    for i, g in enumerate(grids):
        u = stream_function_velocity(g, seed=i, modes=8)
        rhs.append(np.column_stack([u, ETA2 * packed_convection(g, u)]))
    t0 = time.perf_counter()
    batch = GpuBatchedStokes(grids, ncols=2, tolerance=1e-9)
    setup = time.perf_counter() - t0
    flat = torch.as_tensor(batch.pack(rhs), device="cuda:0")
    batch.solve_packed(flat, mass=mass, viscosity=visc)
    timings = batch.autotune(flat, mass=mass, viscosity=visc, repeats=1)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    its = 0
    for _ in range(args.steps):
        out = batch.solve_packed(flat, mass=mass, viscosity=visc)
        its += sum(sum(c) for c in out["iterations"])
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    value = n * args.steps / elapsed
    # Baseline: the same cavities one at a time through the single-cavity library path.
    singles = [GpuTensorStokes(g, tolerance=1e-9) for g in grids]
    dev = [torch.as_tensor(r, device="cuda:0") for r in rhs]
    for s, r, m, v in zip(singles, dev, mass, visc):
        s.solve(r, mass=m, viscosity=v)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    worst = 0.0
    for s, r, m, v, k in zip(singles, dev, mass, visc, range(n)):
        res = s.solve(r, mass=m, viscosity=v)
        worst = max(worst, float(torch.max(torch.abs(res.velocity - torch.stack(
            [out["velocity"][batch.velocity_offsets[2 * k + c]:batch.velocity_offsets[2 * k + c] + grids[k].size]
             for c in (0, 1)], 1)))))
    torch.cuda.synchronize()
    loop = n / (time.perf_counter() - t0)
    return {"metric": "cavity stage solves/s, batched uneven cavities", "value": value,
            "unit": "cavity stage solves/s", "n_gpus": 1, "steps": args.steps, "warmup": 1,
            "ms_per_step": 1e3 * elapsed / args.steps, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64",
            "data": "synthetic: stream-function fields (modes<=8) and their convection, 2 RHS per cavity",
            "config": {"workload": f"{n} cavities, nx,ny uniform on {{64..512 step 32}}, nu log-uniform "
                                   f"[1e-4,1e-2], sdirk2_mrsav dt=5e-4", "total_faces": batch.total_faces},
            "per_cavity_loop": {"value": loop, "unit": "cavity stage solves/s",
                                "max_abs_difference_vs_batched": worst},
            "speedup_vs_loop": value / loop, "launch_shape": list(batch.launch),
            "autotune_seconds": {f"{k[0]},{k[1]}": v for k, v in timings.items()},
            "cg_iterations_per_rhs": its / (2 * n * args.steps), "setup_seconds": setup}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="mixed")
    ap.add_argument("--n", type=int, default=1024, help="grid size (n x n)")
    ap.add_argument("--modes", type=int, default=32)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", choices=["stage1024", "cavities"], default="stage1024")
    ap.add_argument("--cavities", type=int, default=256)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (this build has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        return 2

    if args.workload == "cavities":
        print(json.dumps(bench_cavities(args)), flush=True)
        return 0

    import numpy as np
    import torch
    from macsav import GpuTensorStokes, MACGrid

    grid = MACGrid(args.n, args.n)
    kappa = NU * DT * STAGE_DIAG
    calibration_copy()
    pairs_host = stage_pairs(grid, 4, args.modes)
    pairs = [torch.as_tensor(p, device="cuda:0") for p in pairs_host]

    def run(precision: str):
        # Both forms stop at the same bar: residual <= 1e-10 and max|D u| <= 1e-10 (absolute);
        # the FP64 form needs Schur-CG rtol 1e-14 for that at this size.
        solver = GpuTensorStokes(grid, tolerance=1e-10, krylov_tolerance=1e-14, precision=precision,
                                 refine_tolerance=1e-10)
        for k in range(args.warmup):
            solver.solve(pairs[k % 4], mass=1.0, viscosity=kappa)
        torch.cuda.synchronize()
        iters0, sweeps0 = solver.total_iterations, solver.sweeps
        worst_res, worst_div = 0.0, 0.0
        t0 = time.perf_counter()
        results = []
        for k in range(args.steps):
            res = solver.solve(pairs[k % 4], mass=1.0, viscosity=kappa)
            worst_res, worst_div = max(worst_res, res.residual), max(worst_div, res.divergence_inf)
            if k < 4:
                results.append(res)
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        iters, sweeps = solver.total_iterations - iters0, solver.sweeps - sweeps0
        # Live kernel timing of the same K solves (HIP events on the solver's stream).
        solver.set_profiling(True)
        for k in range(args.steps):
            solver.solve(pairs[k % 4], mass=1.0, viscosity=kappa)
        prof = solver.kernel_profile()
        solver.set_profiling(False)
        stats = {"stage_solves_per_s": args.steps / elapsed, "ms_per_step": 1e3 * elapsed / args.steps,
                 "cg_iterations_per_rhs": iters / (2 * args.steps),
                 "sweeps_per_solve": sweeps / args.steps,
                 "worst_residual": worst_res, "worst_divergence_inf": worst_div,
                 "gemm_share_of_step": prof["gemm_ms"] * 1e-3 / (elapsed + 1e-30)}
        return stats, prof, results

    plain, plain_prof, plain_res = run("fp64")
    if args.precision == "mixed":
        main_stats, prof, mixed_res = run("mixed")
        rel_v = max(float(torch.linalg.vector_norm(a.velocity - b.velocity) / torch.linalg.vector_norm(b.velocity))
                    for a, b in zip(mixed_res, plain_res))
        rel_p = max(float(torch.linalg.vector_norm(a.pressure - b.pressure) / torch.linalg.vector_norm(b.pressure))
                    for a, b in zip(mixed_res, plain_res))
        main_stats["rel_l2_vs_fp64_velocity"] = rel_v
        main_stats["rel_l2_vs_fp64_pressure"] = rel_p
    else:
        main_stats, prof = plain, plain_prof
    gemm_s = prof["gemm_ms"] * 1e-3
    # Flops of the MFMA work the algorithm issues: 2MNK per product, times 3 for the
    # split bf16x3 products of the mixed form (A_lo B_hi + A_hi B_lo + A_hi B_hi).
    achieved = prof["gemm_flops"] / gemm_s / 1e12
    split = 3.0 if args.precision == "mixed" else 1.0
    peak = PEAK_TFLOPS[args.precision]
    value = main_stats["stage_solves_per_s"]

    line = {
        "metric": "stage solves/s at 1024^2 MAC, mixed+learned vs plain FP64 GPU, equal residual",
        "value": value,
        "unit": "stage solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main_stats["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "bf16-gemm/f64-refine",
        "data": "synthetic: first-stage sdirk3_mrsav pairs of 4 seeded stream-function fields (modes<=32)",
        "config": {"workload": f"MAC {grid.nx}x{grid.ny} stage solve, 2 RHS, mass=1, viscosity={kappa:g} "
                               f"(nu=1e-3, dt=2e-4, sdirk3_mrsav)", "precision": args.precision,
                   "bar": "residual<=1e-10 and divergence_inf<=1e-10 (absolute)"},
        "solve": main_stats,
        "plain_fp64": plain,
        "speedup_vs_plain_fp64": value / plain["stage_solves_per_s"],
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": measured_traffic(args.precision),
                     "kernel": "k_gemm_f64" if args.precision == "fp64" else "k_gemm_bf16",
                     "launches": prof["gemm_launches"],
                     "avg_launch_us": 1e6 * gemm_s / max(prof["gemm_launches"], 1),
                     "flops_per_launch": prof["gemm_flops"] / max(prof["gemm_launches"], 1),
                     "transform_flops_per_launch": prof["gemm_flops"] / split / max(prof["gemm_launches"], 1),
                     "algorithmic_bytes_per_launch": prof["gemm_bytes"] / max(prof["gemm_launches"], 1)},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(grid, pairs_host[0], kappa, args.cpu_seconds)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
