#!/usr/bin/env python3
"""Stage-solve throughput of the MAC shifted-Stokes solve at 1024^2 on one MI355X.

Workload: MAC 1024x1024 on the unit square, nu = 1e-3,
sdirk3_mrsav with dt = 2e-4 from the seeded stream-function field with modes <= 32, so
every stage solve has mass 1 and viscosity nu*dt*D[i,i] = 5e-8 and carries the two
right-hand sides [stage rhs, convection] sharing one operator.  The stage solves of a
run are recorded (an FP64 run through the MAC model, one step of start-up first) and
replayed in order from HBM: W untimed solves, then K timed ones.  One "step" is one
two-RHS stage solve to residual <= 1e-10 and max|D u| <= 1e-10 (absolute, the
reference's definitions).

Forms, all on the same recorded right-hand sides:
  plain FP64      the reference algorithm (zero start; Schur-CG rtol 1e-14, which this
                  size needs for the 1e-10 divergence bar)
  FP64 + history  the same, started from the same-stage extrapolation of past steps
  mixed           bf16x3 MFMA transforms inside FP64 defect correction, zero start
  mixed + history the mixed form started from the extrapolated velocity and pressure
  mixed + network the mixed form started from the trained CNN's correction of the
                  extrapolated pressure (trained here on another seed's run)
`value` is the fastest mixed form's stage solves/s; the others are reported beside it.
The dominant kernel (the bf16 transform GEMM) is timed live with HIP events on the
solver's own stream over a second pass of the same K solves.  The CPU baseline times
the oracle (NumPy/SciPy restatement of the reference path) on a bounded sample.
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
STAGES = 4                                       # sdirk3_mrsav stage solves per step
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


def record_stage_solves(grid, count: int, seed: int = 0, modes: int = 32):
    """Right-hand sides (device tensors), mass and viscosity of `count` consecutive
    sdirk3_mrsav stage solves of a plain FP64 run through the MAC model, after one step
    of start-up."""
    import numpy as np
    import torch
    from macsav import GpuTensorStokes
    from macsav.fields import stream_function_velocity
    from macsav.mac_ns import MACNavierStokes
    from macsav.schemes import make_scheme
    backend = GpuTensorStokes(grid, tolerance=1e-10, krylov_tolerance=1e-14)
    model = MACNavierStokes(grid, NU, backend=backend)
    scheme = make_scheme("sdirk3_mrsav")
    state = model.state(0.0, stream_function_velocity(grid, seed=seed, modes=modes))
    state = scheme.step(model, state, DT).state
    rec = []
    original = backend.solve

    def recording(rhs, *, mass, viscosity, pressure_guess=None):
        # Held as (size, k) views of one (k, size) tensor, the layout the device-resident
        # model hands the solver (no transpose copy on the way in).
        cols = np.ascontiguousarray(np.asarray(rhs).T)
        rec.append((torch.as_tensor(cols, device="cuda:0").t(), mass, viscosity))
        return original(rhs, mass=mass, viscosity=viscosity, pressure_guess=pressure_guess)

    backend.solve = recording
    while len(rec) < count:
        state = scheme.step(model, state, DT).state
    backend.solve = original
    return rec[:count]


def train_warm_start(grid, seed: int = 7, steps: int = 6):
    """The CNN warm start trained on (b, base, p) pairs of another seed's FP64 run."""
    from macsav import GpuTensorStokes
    from macsav.fields import stream_function_velocity
    from macsav.learned import WarmStartNet, collect_pairs, fit_alpha, train
    from macsav.mac_ns import MACNavierStokes
    from macsav.schemes import make_scheme
    plain = GpuTensorStokes(grid, tolerance=1e-10, krylov_tolerance=1e-14)
    model = MACNavierStokes(grid, NU, backend=plain)
    scheme = make_scheme("sdirk3_mrsav")
    state = scheme.step(model, model.state(0.0, stream_function_velocity(grid, seed=seed, modes=32)), DT).state
    pairs = collect_pairs(model, scheme, state, steps=steps, dt=DT, period=STAGES)
    net = WarmStartNet(width=32, seed=0, alpha=fit_alpha(grid, pairs))
    losses = train(net, grid, pairs, epochs=40, lr=1e-3, batch=8)
    return net, {"pairs": len(pairs.b), "loss_first_epoch": losses[0], "loss_last_epoch": losses[-1]}


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


def replay(grid, rec, warmup: int, steps: int, precision: str, warm=None, profile: bool = False):
    """Solve rec[:warmup] untimed (they also fill the warm-start history), then time
    rec[warmup:warmup+steps]; optionally a second pass with HIP events on the timed
    solves' transform GEMMs."""
    import torch
    from macsav import GpuTensorStokes
    solver = GpuTensorStokes(grid, tolerance=1e-10, krylov_tolerance=1e-14, precision=precision,
                             refine_tolerance=1e-10)
    if warm is not None:
        solver.set_warm_start(warm, period=STAGES)

    def run_once(timed_results=None, profiled=False):
        solver.reset_history()
        for f, m, v in rec[:warmup]:
            solver.solve(f, mass=m, viscosity=v)
        torch.cuda.synchronize()
        if profiled:  # HIP events on the timed solves only
            solver.set_profiling(True)
        it0, sw0, w0 = solver.total_iterations, solver.sweeps, solver.warm_starts
        worst = [0.0, 0.0]
        t0 = time.perf_counter()
        for f, m, v in rec[warmup:warmup + steps]:
            res = solver.solve(f, mass=m, viscosity=v)
            worst = [max(worst[0], res.residual), max(worst[1], res.divergence_inf)]
            if timed_results is not None:
                timed_results.append(res.velocity)
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        return elapsed, solver.total_iterations - it0, solver.sweeps - sw0, solver.warm_starts - w0, worst

    velocities = []
    elapsed, its, sweeps, warm_used, worst = run_once(velocities)
    stats = {"stage_solves_per_s": steps / elapsed, "ms_per_step": 1e3 * elapsed / steps,
             "cg_iterations_per_rhs": its / (2 * steps), "sweeps_per_solve": sweeps / steps,
             "warm_started_solves": warm_used, "worst_residual": worst[0], "worst_divergence_inf": worst[1]}
    prof = None
    if profile:
        prof_elapsed = run_once(profiled=True)[0]
        prof = solver.kernel_profile()
        solver.set_profiling(False)
        # Share of the profiled pass's own wall time (its event pairs slow it a little).
        stats["gemm_share_of_step"] = prof["gemm_ms"] * 1e-3 / (prof_elapsed + 1e-30)
    return stats, prof, velocities


def bench_run(args) -> dict:
    """End-to-end sdirk3_mrsav run at n^2 (the stage-solve workload's run), timed over
    `--steps` steps after `--warmup` steps: the host model with the plain FP64 backend
    (the reference's arrangement, fields on the host), and the device-resident model with
    the plain FP64 and the mixed + history backends (fields stay in HBM)."""
    import numpy as np
    import torch
    from macsav import GpuTensorStokes, MACGrid
    from macsav.fields import stream_function_velocity
    from macsav.mac_ns import MACNavierStokes
    from macsav.mac_ns_device import DeviceMACNavierStokes
    from macsav.schemes import make_scheme
    grid = MACGrid(args.n, args.n)
    x0 = stream_function_velocity(grid, seed=0, modes=args.modes)
    out = {}
    finals = {}
    for form in ("host+fp64", "device+fp64", "device+mixed+history"):
        precision = "mixed" if "mixed" in form else "fp64"
        backend = GpuTensorStokes(grid, tolerance=1e-10, krylov_tolerance=1e-14, precision=precision,
                                  refine_tolerance=1e-10)
        if "history" in form:
            backend.set_warm_start("history", period=STAGES)
        model = (MACNavierStokes if form.startswith("host") else DeviceMACNavierStokes)(grid, NU, backend=backend)
        scheme = make_scheme("sdirk3_mrsav")
        state = model.state(0.0, x0)
        for _ in range(args.warmup):
            state = scheme.step(model, state, DT).state
        torch.cuda.synchronize()
        solves0 = backend.solves
        t0 = time.perf_counter()
        for _ in range(args.steps):
            state = scheme.step(model, state, DT).state
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        u = model.vector(state)
        finals[form] = u.cpu().numpy() if hasattr(u, "cpu") else np.asarray(u)
        out[form] = {"steps_per_s": args.steps / elapsed, "ms_per_step": 1e3 * elapsed / args.steps,
                     "stage_solves_per_s": (backend.solves - solves0) / elapsed}
    ref = finals["host+fp64"]
    for form, u in finals.items():
        out[form]["rel_l2_vs_host_fp64"] = float(np.linalg.norm(u - ref) / np.linalg.norm(ref))
    best = out["device+mixed+history"]
    return {"metric": "sdirk3_mrsav steps/s end to end at 1024^2 MAC", "value": best["steps_per_s"],
            "unit": "steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": best["ms_per_step"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "bf16-gemm/f64-refine",
            "data": f"seeded stream-function field (modes<={args.modes}), nu=1e-3, dt=2e-4",
            "config": {"workload": f"MAC {grid.nx}x{grid.ny} sdirk3_mrsav run, 4 two-RHS stage solves per step"},
            "forms": out, "speedup_vs_host_fp64": best["steps_per_s"] / out["host+fp64"]["steps_per_s"]}


def bench_combined(args) -> dict:
    """Combined workload at n^2 (default 2048^2): nu = 2e-4, the top wall moving at 1 plus
    the seeded stream-function perturbation (modes <= 64, amplitude 0.05), sdirk3_mrsav
    under PI step control (atol = rtol = 1e-6) to t = 0.05, fields resident on the
    device.  The plain FP64 backend and the mixed + history backend run the same
    problem; reported: accepted and attempted steps, stage solves/s, wall time and the
    final-field difference between them."""
    import numpy as np
    import torch
    from macsav import GpuTensorStokes, MACGrid
    from macsav.adaptivity import PIController, integrate_adaptive
    from macsav.fields import stream_function_velocity
    from macsav.mac_ns_device import DeviceMACNavierStokes
    from macsav.schemes import make_scheme
    n = args.n or 2048
    nu, lid, modes, t_end, first = 2e-4, 1.0, 64, args.t_end, 1e-4
    grid = MACGrid(n, n)
    x0 = stream_function_velocity(grid, seed=0, modes=modes, amplitude=0.05)
    controller = PIController(atol=1e-6, rtol=1e-6, max_steps=args.max_steps)
    out, finals = {}, {}
    for form in ("device+fp64", "device+mixed+history"):
        precision = "mixed" if "mixed" in form else "fp64"

        def backend_for():
            b = GpuTensorStokes(grid, tolerance=1e-10, krylov_tolerance=1e-14, precision=precision,
                                refine_tolerance=1e-10)
            if "history" in form:
                b.set_warm_start("history", period=STAGES)
            return b
        # One untimed step on a throwaway backend first (code paging, first-touch memory).
        scratch = DeviceMACNavierStokes(grid, nu, backend=backend_for(), lid=lid)
        make_scheme("sdirk3_mrsav").step(scratch, scratch.state(0.0, x0), first)
        del scratch
        torch.cuda.synchronize()
        backend = backend_for()
        model = DeviceMACNavierStokes(grid, nu, backend=backend, lid=lid)

        def progress(k, t, secs, form=form):
            if k % 10 == 0:
                print(f"[{form}] accepted {k} t={t:.5f} {secs:.1f}s", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        res = integrate_adaptive(model, make_scheme("sdirk3_mrsav"), model.state(0.0, x0), t_end, first,
                                 controller=controller, progress=progress, retain_history=False)
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        if res.status != "complete":
            raise RuntimeError(f"{form}: {res.status} {res.error}")
        finals[form] = model.vector(res.final).cpu().numpy()
        out[form] = {"accepted_steps": res.accepted_steps, "attempted_steps": res.attempted_steps,
                     "seconds": elapsed, "stage_solves": backend.solves,
                     "stage_solves_per_s": backend.solves / elapsed,
                     "schur_cg_iterations": backend.total_iterations,
                     "final_t": float(res.final.t), "final_divergence_inf": res.diagnostics[-1]["divergence_inf"],
                     "final_kinetic": res.diagnostics[-1]["kinetic"]}
        del model, backend
        torch.cuda.synchronize()
    ref = finals["device+fp64"]
    diff = finals["device+mixed+history"] - ref
    best, plain = out["device+mixed+history"], out["device+fp64"]
    return {"metric": f"stage solves/s, adaptive sdirk3_mrsav at {n}^2 MAC (lid + perturbation)",
            "value": best["stage_solves_per_s"], "unit": "stage solves/s", "n_gpus": 1,
            "steps": best["accepted_steps"], "warmup": 1, "ms_per_step": 1e3 * best["seconds"] / best["accepted_steps"],
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16-gemm/f64-refine",
            "data": f"seeded stream-function perturbation (modes<={modes}, amplitude 0.05) on a cavity with lid "
                    f"velocity {lid:g}",
            "config": {"workload": f"MAC {n}x{n}, nu={nu:g}, lid {lid:g}, sdirk3_mrsav with PI control "
                                   f"atol=rtol=1e-6 to t={t_end:g}, initial step {first:g}"},
            "forms": out, "speedup_vs_plain_fp64": best["stage_solves_per_s"] / plain["stage_solves_per_s"],
            "wall_speedup_vs_plain_fp64": plain["seconds"] / best["seconds"],
            "final_rel_l2_vs_fp64": float(np.linalg.norm(diff) / np.linalg.norm(ref)),
            "final_max_abs_vs_fp64": float(np.abs(diff).max())}


def bench_sizes(args) -> dict:
    """Two-RHS stage-solve time against grid size (n x n, n in 64..1024), plain FP64
    against mixed (zero start), on the recorded right-hand sides of a short
    sdirk3_mrsav run at each size (nu = 1e-3, dt = 2e-4): where the bf16 transforms start
    to pay, which sets the precision of the batched cavity path (n <= 512)."""
    import torch
    from macsav import GpuTensorStokes, MACGrid
    rows = {}
    for n in (64, 128, 256, 384, 512, 768, 1024):
        grid = MACGrid(n, n)
        rec = record_stage_solves(grid, 12, modes=min(32, n // 4))
        row = {}
        for precision in ("fp64", "mixed"):
            s = GpuTensorStokes(grid, tolerance=1e-10, krylov_tolerance=1e-14, precision=precision,
                                refine_tolerance=1e-10)
            for f, m, v in rec[:4]:
                s.solve(f, mass=m, viscosity=v)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for f, m, v in rec[4:]:
                s.solve(f, mass=m, viscosity=v)
            torch.cuda.synchronize()
            row[precision] = 1e3 * (time.perf_counter() - t0) / (len(rec) - 4)
        row["mixed_speedup"] = row["fp64"] / row["mixed"]
        rows[str(n)] = row
        print(f"[sizes] {n}: {row}", file=sys.stderr, flush=True)
    return {"metric": "ms per two-RHS stage solve against grid size, plain FP64 and mixed",
            "value": rows["1024"]["mixed_speedup"], "unit": "x (mixed over FP64 at 1024^2)", "n_gpus": 1,
            "steps": 8, "warmup": 4, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "bf16-gemm/f64-refine", "data": "recorded sdirk3_mrsav stage right-hand sides per size",
            "config": {"workload": "MAC n x n stage solves, n in 64..1024, zero start, residual 1e-10"},
            "sizes": rows}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed stage solves")
    ap.add_argument("--warmup", type=int, default=12, help="untimed stage solves before them (3 steps)")
    ap.add_argument("--n", type=int, default=None, help="grid size (n x n); 1024, or 2048 for the combined run")
    ap.add_argument("--modes", type=int, default=32)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-network", action="store_true")
    ap.add_argument("--workload", choices=["stage1024", "cavities", "run", "combined", "sizes"], default="stage1024")
    ap.add_argument("--t-end", type=float, default=0.05, help="combined run: final time")
    ap.add_argument("--max-steps", type=int, default=2000, help="combined run: accepted-step cap")
    ap.add_argument("--cavities", type=int, default=256)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (this build has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        return 2

    if args.workload == "cavities":
        print(json.dumps(bench_cavities(args)), flush=True)
        return 0
    if args.workload == "sizes":
        print(json.dumps(bench_sizes(args)), flush=True)
        return 0
    if args.workload == "combined":
        print(json.dumps(bench_combined(args)), flush=True)
        return 0
    args.n = args.n or 1024
    if args.workload == "run":
        print(json.dumps(bench_run(args)), flush=True)
        return 0

    import torch
    from macsav import MACGrid

    grid = MACGrid(args.n, args.n)
    kappa = NU * DT * STAGE_DIAG
    calibration_copy()
    t0 = time.perf_counter()
    rec = record_stage_solves(grid, args.warmup + args.steps, modes=args.modes)
    record_s = time.perf_counter() - t0
    W, K = args.warmup, args.steps

    plain, _, plain_v = replay(grid, rec, W, K, "fp64")
    forms = {"fp64": plain}
    forms["fp64+history"], _, _ = replay(grid, rec, W, K, "fp64", warm="history")
    forms["mixed"], prof_mixed, mixed_v = replay(grid, rec, W, K, "mixed", profile=True)
    forms["mixed+history"], prof_hist, hist_v = replay(grid, rec, W, K, "mixed", warm="history", profile=True)
    net_info = None
    if not args.no_network:
        net, net_info = train_warm_start(grid)
        forms["mixed+network"], _, _ = replay(grid, rec, W, K, "mixed", warm=net)

    def rel(a_list, b_list):
        return max(float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b)) for a, b in zip(a_list, b_list))

    forms["mixed"]["rel_l2_vs_fp64_velocity"] = rel(mixed_v, plain_v)
    forms["mixed+history"]["rel_l2_vs_fp64_velocity"] = rel(hist_v, plain_v)
    best = max(("mixed", "mixed+history"), key=lambda k: forms[k]["stage_solves_per_s"])
    main_stats = forms[best]
    prof = prof_hist if best == "mixed+history" else prof_mixed
    gemm_s = prof["gemm_ms"] * 1e-3
    # Flops of the MFMA work the algorithm issues: 2MNK per product, times 3 for the
    # split bf16x3 products (A_lo B_hi + A_hi B_lo + A_hi B_hi).
    achieved = prof["gemm_flops"] / gemm_s / 1e12
    split = 3.0
    value = main_stats["stage_solves_per_s"]
    for k, f in forms.items():
        f["speedup_vs_plain_fp64"] = f["stage_solves_per_s"] / plain["stage_solves_per_s"]

    line = {
        "metric": "stage solves/s at 1024^2 MAC, mixed+learned vs plain FP64 GPU, equal residual",
        "value": value,
        "unit": "stage solves/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": main_stats["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16-gemm/f64-refine",
        "data": f"synthetic: {W}+{K} consecutive stage solves of a recorded sdirk3_mrsav run from the seeded "
                f"stream-function field (modes<={args.modes}), after one start-up step",
        "config": {"workload": f"MAC {grid.nx}x{grid.ny} stage solve, 2 RHS, mass=1, viscosity={kappa:g} "
                               f"(nu=1e-3, dt=2e-4, sdirk3_mrsav)", "form": best,
                   "bar": "residual<=1e-10 and divergence_inf<=1e-10 (absolute)"},
        "solve": main_stats,
        "plain_fp64": plain,
        "speedup_vs_plain_fp64": value / plain["stage_solves_per_s"],
        "forms": forms,
        "warm_start_network": net_info,
        "record_seconds": record_s,
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_TFLOPS["mixed"], "unit": "TFLOP/s",
                     "frac": achieved / PEAK_TFLOPS["mixed"], "traffic": measured_traffic("mixed"),
                     "kernel": "k_gemm_bf16",
                     "launches": prof["gemm_launches"],
                     "avg_launch_us": 1e6 * gemm_s / max(prof["gemm_launches"], 1),
                     "flops_per_launch": prof["gemm_flops"] / max(prof["gemm_launches"], 1),
                     "transform_flops_per_launch": prof["gemm_flops"] / split / max(prof["gemm_launches"], 1),
                     "algorithmic_bytes_per_launch": prof["gemm_bytes"] / max(prof["gemm_launches"], 1)},
    }
    if not args.no_cpu:
        f0, m0, v0 = rec[W]
        line["cpu_baseline"] = cpu_baseline(grid, f0.cpu().numpy(), v0, args.cpu_seconds)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
