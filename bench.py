#!/usr/bin/env python3
"""Stage-solve throughput of the MAC shifted-Stokes solve at 1024^2 on one MI355X.

Workload: MAC 1024x1024 on the unit square, nu = 1e-3, sdirk3_mrsav with dt = 2e-4 from
the seeded stream-function field with modes <= 32, so every stage solve has mass 1 and
viscosity nu dt D[i,i] = 5e-8 and carries the two right-hand sides [stage rhs,
convection] sharing one operator.  The stage solves of a run are recorded on the
device-resident model (FP64) and replayed in order from HBM.  Each form first solves
the run's first six steps untimed (this fills the same-stage history the warm starts
extrapolate), then W untimed and K timed stage solves.  One "step" is one two-RHS stage
solve to residual <= 1e-10 and max|D u| <= 1e-10, absolute, in the reference's
definitions; each form's worst values are reported and checked against that bar.

Forms, all on the same recorded right-hand sides:
  fp64             the reference algorithm: zero start, Schur-CG rtol 1e-14.  Its
                   max|D u| stops near 2e-10 at this size (the oracle's too), above the bar
  fp64+refine      the same with rtol 1e-12 and FP64 defect correction to the bar: the
                   plain FP64 GPU solve at equal final residual (the speedup baseline)
  fp64+history     fp64+refine started from the same-stage history extrapolation
  mixed            bf16x3 MFMA transforms inside FP64 defect correction, zero start
  mixed+history    the mixed form started from the extrapolated velocity and pressure
                   (degree 5 through the last six steps)
  mixed+network    mixed+history whose early solves (fewer than three recorded steps)
                   start from the trained CNN's correction of the extrapolation
`value` is mixed+network's stage solves/s (the metric's form).  The same forms are also
timed over every stage solve of the run (cold start included: `whole_run`).  The
dominant kernel (the split bf16 transform GEMM) is timed live with HIP events on the
solver's own stream over a second pass of the K solves.  The CPU baseline times the
oracle (NumPy/SciPy restatement of the reference path) on a bounded sample, on one core
and on the host's cores.

Other workloads (--workload): learned (the learned warm start at 512^2, 40 training and
10 held-out seeds), cavities (batched uneven cavities), run (end-to-end steps),
combined (adaptive 2048^2 with lid), sizes (solve time against grid size).
"""
from __future__ import annotations

import argparse
import gc
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
for _p in (str(ROOT), str(ROOT / "mac-mr-sav-rk_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

PEAK_TFLOPS = {"fp64": 78.6, "bf16": 2500.0}     # MI355X dense: FP64 MFMA, BF16 MFMA
PEAK_HBM_GBS = 8000.0                            # MI355X HBM3E
NU, DT, STAGE_DIAG = 1e-3, 2e-4, 0.25            # sdirk3 D[i,i] = 1/4
STAGES = 4                                       # sdirk3_mrsav stage solves per step
HISTORY_STEPS = 6                                # extrapolation depth of the history forms
FILL_STEPS = 6                                   # run steps solved before the replay
BAR = 1e-10
CALIBRATION_BYTES = 512 << 20                    # one plain copy of known size, before timing
TRAFFIC_FILE = ROOT / "profiles" / "k_gemm_traffic_1024.md"
FFT_TRAFFIC_FILE = ROOT / "profiles" / "fft_traffic_1024.md"

FORMS = {
    "fp64": dict(precision="fp64", krylov=1e-14),
    "fp64+refine": dict(precision="fp64", krylov=1e-12, refine=3),
    "fp64+history": dict(precision="fp64", krylov=1e-12, refine=3, warm="history"),
    "mixed": dict(precision="mixed"),
    "mixed+history": dict(precision="mixed", warm="history"),
    "mixed+network": dict(precision="mixed", warm="network"),
    # The same history form with the split-bf16 transform GEMMs in place of the FFT passes.
    "mixed+history/gemm": dict(precision="mixed", warm="history", split=7),
}


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
        for line in (FFT_TRAFFIC_FILE if precision == "fft" else TRAFFIC_FILE).read_text().splitlines():
            parts = line.strip().strip("`").split()
            if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch" and parts[1] == precision:
                return float(parts[2])
    except OSError:
        return None
    return None


def record_run(grid, steps: int, seed: int = 0, modes: int = 32, nu: float = NU, dt: float = DT):
    """Right-hand sides ((size, 2) device views), mass and viscosity of every stage solve
    of an sdirk3_mrsav run on the device-resident model with the refined FP64 backend."""
    from macsav import GpuTensorStokes
    from macsav.fields import stream_function_velocity
    from macsav.mac_ns_device import DeviceMACNavierStokes
    from macsav.schemes import make_scheme
    backend = GpuTensorStokes(grid, tolerance=BAR, krylov_tolerance=1e-12, fp64_refine=3)
    model = DeviceMACNavierStokes(grid, nu, backend=backend)
    rec = []
    inner = backend.solve

    def recording(rhs, *, mass, viscosity, pressure_guess=None):
        rec.append((rhs.clone(), mass, viscosity))
        return inner(rhs, mass=mass, viscosity=viscosity, pressure_guess=pressure_guess)

    backend.solve = recording
    scheme = make_scheme("sdirk3_mrsav")
    state = model.state(0.0, stream_function_velocity(grid, seed=seed, modes=modes))
    for _ in range(steps):
        state = scheme.step(model, state, dt).state
    del backend.solve        # the wrapper holds the backend: break the cycle so it is freed now
    return rec


def quiesce(hold: bool = True):
    """Free what earlier work left to the cyclic collector (a solver freed there releases
    device memory, which synchronises the device) and finish queued device work, then
    (hold) keep the collector off until `resume`: nothing of that kind lands in a timed
    region.  Long runs that make garbage as they go pass hold=False."""
    import torch
    gc.collect()
    torch.cuda.synchronize()
    if hold:
        gc.disable()


def resume():
    gc.enable()


def transform_roofline(prof):
    """Roofline of the solve's transform kernels from the event profile of the timed solves.

    FFT passes (the mixed form where every length has a power-of-two FFT): HBM-bound;
    `achieved` = their algorithmic bytes (every element read once and written once per
    pass, at the pass's input and output widths) / their device time, against 8 TB/s.
    Split-bf16 GEMMs: MFMA-bound; `achieved` = the transform flops of the passes (2MNn per
    pass over a length-n transform, one product) / their device time, against 2.5 PF, with
    the MFMA work issued (three products per k-step) beside it."""
    if prof.get("fft_launches", 0) >= prof["gemm_launches"] and prof.get("fft_launches", 0) > 0:
        fft_s = prof["fft_ms"] * 1e-3
        n = prof["fft_launches"]
        achieved = prof["fft_bytes"] / fft_s / 1e9
        return {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": achieved / PEAK_HBM_GBS, "traffic": measured_traffic("fft"),
                "kernel": "k_fft_rows + k_fft_strips (the three FFT passes of each transform)",
                "work": "algorithmic bytes: each pass reads and writes every element once",
                "transform_tflops": prof["fft_transform_flops"] / fft_s / 1e12,
                "launches": n, "avg_launch_us": 1e6 * fft_s / n,
                "algorithmic_bytes_per_launch": prof["fft_bytes"] / n,
                "transform_flops_per_launch": prof["fft_transform_flops"] / n}
    gemm_s = prof["gemm_ms"] * 1e-3
    launches = max(prof["gemm_launches"], 1)
    executed = prof["gemm_flops"] / gemm_s / 1e12
    algorithmic = prof["gemm_transform_flops"] / gemm_s / 1e12
    return {"bound": "mfma", "achieved": algorithmic, "peak": PEAK_TFLOPS["bf16"], "unit": "TFLOP/s",
            "frac": algorithmic / PEAK_TFLOPS["bf16"], "traffic": measured_traffic("mixed"),
            "kernel": "k_gemm_bf16",
            "work": "transform flops 2MNn per pass (one product; n the transform length)",
            "mfma_executed_tflops": executed, "mfma_executed_frac": executed / PEAK_TFLOPS["bf16"],
            "launches": prof["gemm_launches"], "avg_launch_us": 1e6 * gemm_s / launches,
            "flops_per_launch": prof["gemm_transform_flops"] / launches,
            "algorithmic_bytes_per_launch": prof["gemm_bytes"] / launches}


def make_solver(grid, form: str, net=None, history_steps: int = HISTORY_STEPS):
    from macsav import GpuTensorStokes
    f = FORMS[form]
    s = GpuTensorStokes(grid, tolerance=BAR, krylov_tolerance=f.get("krylov", 1e-14), precision=f["precision"],
                        refine_tolerance=BAR, fp64_refine=f.get("refine", 0), mixed_split=f.get("split", 15))
    if f.get("warm") == "history":
        s.set_warm_start("history", period=STAGES, history_steps=history_steps)
    elif f.get("warm") == "network":
        s.set_warm_start(net, period=STAGES, history_steps=history_steps, network_depth=3)
    return s


def _tally(solver):
    return (solver.total_iterations, solver.sweeps, solver.warm_starts, solver.network_starts)


def _stats(solver, before, n, elapsed, worst):
    it, sw, w, nw = (a - b for a, b in zip(_tally(solver), before))
    return {"stage_solves_per_s": n / elapsed, "ms_per_step": 1e3 * elapsed / n,
            "cg_iterations_per_rhs": it / (2 * n), "sweeps_per_solve": sw / n, "warm_started_solves": w,
            "network_solves": nw, "worst_residual": worst[0], "worst_divergence_inf": worst[1],
            "meets_bar": worst[0] <= BAR and worst[1] <= BAR}


def replay(grid, rec, fill: int, warmup: int, steps: int, form: str, net=None, profile: bool = False):
    """rec[:fill + warmup] untimed (the history fill, then the warm-up), then
    rec[fill + warmup:][:steps]: once untimed, keeping the velocities, then timed;
    optionally a further pass with HIP events on the timed solves' transform launches."""
    import torch
    solver = make_solver(grid, form, net, HISTORY_STEPS)
    first = fill + warmup

    def run_once(keep=None, profiled=False):
        solver.reset_history()
        for f, m, v in rec[:first]:
            solver.solve(f, mass=m, viscosity=v)
            solver.commit_history()
        torch.cuda.synchronize()
        if profiled:
            solver.set_profiling(True)
        before = _tally(solver)
        worst = [0.0, 0.0]
        quiesce()
        t0 = time.perf_counter()
        for f, m, v in rec[first:first + steps]:
            res = solver.solve(f, mass=m, viscosity=v)
            solver.commit_history()
            worst = [max(worst[0], res.residual), max(worst[1], res.divergence_inf)]
            if keep is not None:
                keep.append(res.velocity)
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        resume()
        return _stats(solver, before, steps, elapsed, worst), elapsed

    # The velocities the forms are compared on come from an untimed pass over the same
    # records: kept inside the timed pass, every output would stay allocated, and the
    # caching allocator would call hipMalloc (about 60 us) inside every timed solve.  In the
    # timed pass each output is freed when the next replaces it, as in an integrator.
    velocities = []
    run_once(velocities)
    stats, _ = run_once()
    prof = None
    if profile:
        _, prof_elapsed = run_once(profiled=True)
        prof = solver.kernel_profile()
        solver.set_profiling(False)
        stats["transform_share_of_step"] = (prof["gemm_ms"] + prof["fft_ms"]) * 1e-3 / (prof_elapsed + 1e-30)
    return stats, prof, velocities


def whole_run(grid, rec, form: str, net=None):
    """Every recorded stage solve timed in order from a fresh solver (cold start
    included), after one untimed solve of the last record on a throwaway solver."""
    import torch
    scratch = make_solver(grid, form, net, HISTORY_STEPS)
    f, m, v = rec[-1]
    scratch.solve(f, mass=m, viscosity=v)
    del scratch
    solver = make_solver(grid, form, net, HISTORY_STEPS)
    before = _tally(solver)
    worst = [0.0, 0.0]
    quiesce()
    t0 = time.perf_counter()
    for f, m, v in rec:
        res = solver.solve(f, mass=m, viscosity=v)
        solver.commit_history()
        worst = [max(worst[0], res.residual), max(worst[1], res.divergence_inf)]
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    resume()
    return _stats(solver, before, len(rec), elapsed, worst)


def train_cold_start_network(grid, seeds=(11, 12, 13, 14, 15, 16), steps: int = 3):
    """The CNN of the mixed+network form, trained on the solves it serves: those with one
    or two recorded steps (steps 2 and 3 of a run), pairs recorded from FP64 solves of
    other seeds' runs, their base the history extrapolation through the recorded steps."""
    from macsav import GpuTensorStokes
    from macsav.fields import stream_function_velocity
    from macsav.learned import Pairs, WarmStartNet, collect_pairs, fit_alpha, train
    from macsav.mac_ns_device import DeviceMACNavierStokes
    from macsav.schemes import make_scheme
    pairs = Pairs([], [], [])
    for seed in seeds:
        backend = GpuTensorStokes(grid, tolerance=BAR, krylov_tolerance=1e-12, fp64_refine=3)
        model = DeviceMACNavierStokes(grid, NU, backend=backend)
        x0 = stream_function_velocity(grid, seed=seed, modes=32)
        got = collect_pairs(model, make_scheme("sdirk3_mrsav"), model.state(0.0, x0), steps=steps, dt=DT,
                            period=STAGES, depth=2)
        pairs.b += got.b
        pairs.base += got.base
        pairs.p += got.p
    net = WarmStartNet(width=32, seed=0, alpha=fit_alpha(grid, pairs))
    losses = train(net, grid, pairs, epochs=30, lr=1e-3, batch=8)
    return net, {"pairs": len(pairs.b), "seeds": list(seeds), "loss_first_epoch": losses[0],
                 "loss_last_epoch": losses[-1]}


def cpu_baseline(grid, pair, kappa: float, budget_s: float, workers: int) -> dict:
    """Oracle (scipy.fft transforms + the scipy-cg loop) on `workers` FFT threads."""
    import scipy.fft
    from oracle import mac_oracle as mo
    g = mo.Grid(grid.nx, grid.ny, grid.lx, grid.ly)
    solver = mo.TensorStokes(g, tolerance=1e-9, krylov_tolerance=1e-14)
    done, t0 = 0, time.perf_counter()
    with scipy.fft.set_workers(workers):
        while True:
            solver.solve(pair, mass=1.0, viscosity=kappa)
            done += 1
            if time.perf_counter() - t0 > budget_s or done >= 20:
                break
    secs = time.perf_counter() - t0
    return {"value": done / secs, "unit": "stage solves/s", "cores": workers, "kind": "port",
            "sample": f"{done} two-RHS stage solves at {grid.nx}x{grid.ny} (oracle TensorStokes, scipy.fft "
                      f"workers={workers}, host nproc={os.cpu_count()}), {secs:.1f} s"}


def bench_learned(args) -> dict:
    """The learned warm start at 512^2: sdirk3_mrsav runs with nu alternating 1e-3 / 3e-4
    (dt = 2e-4, modes <= 32); the network trained on the (b - b_prev, previous-stage
    pressure, pressure) pairs of the training seeds' FP64 solves (13 steps each, about
    100 pairs per seed), then every held-out seed's run solved with the zero start (the
    reference), the previous-stage pressure (Galerkin-scaled) and the network's
    two-term start over {previous, correction}, all by the FP64 form at the bar (Schur-CG
    rtol 1e-12 plus FP64 defect correction to 1e-10); CG iterations and wall time per
    start."""
    import numpy as np
    import torch
    from macsav import GpuTensorStokes, MACGrid
    from macsav.fields import stream_function_velocity
    from macsav.learned import Pairs, WarmStartNet, collect_pairs, fit_alpha, train
    from macsav.mac_ns_device import DeviceMACNavierStokes
    from macsav.schemes import make_scheme
    n = args.n or 512
    grid = MACGrid(n, n)
    nus = (1e-3, 3e-4)
    train_seeds = list(range(100, 100 + args.train_seeds))
    test_seeds = list(range(200, 200 + args.test_seeds))
    steps = 13

    def backend(warm=None, net=None):
        # The FP64 form at the bar: Schur-CG rtol 1e-12, then FP64 defect correction to
        # residual and max|D u| <= 1e-10 (its iterations count with the CG's).
        b = GpuTensorStokes(grid, tolerance=BAR, krylov_tolerance=1e-12, fp64_refine=3)
        if warm == "previous":
            b.set_warm_start("previous")
        elif warm == "network":
            b.set_warm_start(net)
        return b

    t0 = time.perf_counter()
    pairs = Pairs([], [], [])
    for k, seed in enumerate(train_seeds):
        model = DeviceMACNavierStokes(grid, nus[k % 2], backend=backend())
        x0 = stream_function_velocity(grid, seed=seed, modes=32)
        got = collect_pairs(model, make_scheme("sdirk3_mrsav"), model.state(0.0, x0), steps=steps, dt=DT)
        pairs.b += got.b
        pairs.base += got.base
        pairs.p += got.p
    collect_s = time.perf_counter() - t0
    n_pairs = sum(int(b.shape[1]) for b in pairs.b)
    print(f"[learned] {n_pairs} pairs from {len(train_seeds)} seeds in {collect_s:.1f}s", file=sys.stderr,
          flush=True)
    net = WarmStartNet(width=32, seed=0, alpha=fit_alpha(grid, pairs))
    t0 = time.perf_counter()
    losses = train(net, grid, pairs, epochs=args.epochs, lr=1e-3, batch=16)
    train_s = time.perf_counter() - t0
    print(f"[learned] trained {args.epochs} epochs in {train_s:.1f}s, loss {losses[0]:.3g} -> {losses[-1]:.3g}",
          file=sys.stderr, flush=True)
    del pairs
    rows = {}
    for start in ("zero", "previous", "network"):
        its = solves = 0
        secs = 0.0
        for k, seed in enumerate(test_seeds):
            b = backend(None if start == "zero" else start, net)
            model = DeviceMACNavierStokes(grid, nus[k % 2], backend=b)
            x0 = stream_function_velocity(grid, seed=seed, modes=32)
            state = model.state(0.0, x0)
            scheme = make_scheme("sdirk3_mrsav")
            torch.cuda.synchronize()
            for _ in range(steps):
                t1 = b.solve_seconds
                state = scheme.step(model, state, DT).state
                secs += b.solve_seconds - t1
            its += b.total_iterations
            solves += b.solves
        rows[start] = {"cg_iterations": its, "stage_solves": solves, "cg_iterations_per_rhs": its / (2 * solves),
                       "solve_seconds": secs, "ms_per_solve": 1e3 * secs / solves}
        print(f"[learned] {start}: {rows[start]}", file=sys.stderr, flush=True)
    gain = 1.0 - rows["network"]["cg_iterations"] / rows["previous"]["cg_iterations"]
    return {"metric": "Schur-CG iterations per RHS to 1e-10, learned warm start vs zero / previous-stage guess",
            "value": rows["network"]["cg_iterations_per_rhs"], "unit": "CG iterations per RHS", "n_gpus": 1,
            "steps": steps, "warmup": 0, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
            "dtype": "f64 (network bf16 weights, fp32 accumulation)",
            "data": f"synthetic: sdirk3_mrsav runs at {n}^2, nu alternating {nus}, seeded stream-function fields",
            "config": {"workload": f"MAC {n}x{n}, {len(train_seeds)} training seeds, {len(test_seeds)} held-out "
                                   f"seeds, {steps} steps each"},
            "starts": rows, "iteration_saving_vs_previous": gain,
            "training": {"pairs": n_pairs, "collect_seconds": collect_s, "train_seconds": train_s,
                         "epochs": args.epochs, "loss_first_epoch": losses[0], "loss_last_epoch": losses[-1]}}


ETA2 = 1.0 - 0.5 ** 0.5                          # sdirk2 diagonal


def record_cavity_stages(grids, nu, steps: int, dt: float = 5e-4, lid: float = 1.0, spinup: int = 0):
    """Per cavity, the stage solves ((size, 2) device right-hand sides, mass, viscosity) of
    `steps` sdirk2_mrsav steps after `spinup` unrecorded ones, from rest with the top wall
    moving at `lid`, each cavity on its own device-resident model with the single-cavity
    FP64 backend."""
    import numpy as np
    from macsav import GpuTensorStokes
    from macsav.mac_ns_device import DeviceMACNavierStokes
    from macsav.schemes import make_scheme
    scheme = make_scheme("sdirk2_mrsav")
    stages = []
    for g, v in zip(grids, nu):
        backend = GpuTensorStokes(g, tolerance=1e-9)
        model = DeviceMACNavierStokes(g, float(v), backend=backend, lid=lid)
        rec = []
        inner = backend.solve

        def recording(rhs, *, mass, viscosity, pressure_guess=None, inner=inner, rec=rec):
            rec.append((rhs.clone(), mass, viscosity))
            return inner(rhs, mass=mass, viscosity=viscosity, pressure_guess=pressure_guess)

        state = model.state(0.0, np.zeros(g.size))
        for _ in range(spinup):
            state = scheme.step(model, state, dt).state
        backend.solve = recording
        for _ in range(steps):
            state = scheme.step(model, state, dt).state
        del backend.solve
        stages.append(rec)
        if len(stages) % 32 == 0:
            print(f"[cavities] recorded {len(stages)} of {len(grids)} cavities", file=sys.stderr, flush=True)
    return stages


def bench_cavities(args) -> dict:
    """Batched uneven cavities: `--cavities` cavities with nx, ny drawn uniformly from
    {64, 96, ..., 512}, nu log-uniform on [1e-4, 1e-2], the top wall moving at 1,
    sdirk2_mrsav with dt = 5e-4.  The first `--record-steps` steps of every cavity are
    run from rest on its own device model and their stage solves recorded (two
    right-hand sides each).  Each recorded stage is then one batched solve of every
    cavity: the first `--fill-steps` steps untimed (they fill the history), the rest
    timed, in the mixed form with the same-stage history start (`value`), the mixed form
    from zero (FP64 CG with split-bf16 transforms under FP64 defect correction to
    residual and max|D u| <= 1e-10) and the FP64 form.  The baseline is a loop of
    single-cavity plain FP64 GPU solves (GpuTensorStokes, one cavity at a time, zero
    start) of the same timed right-hand sides."""
    import numpy as np
    import torch
    from macsav import GpuBatchedStokes, GpuTensorStokes, MACGrid
    rng = np.random.default_rng(2026)
    n = args.cavities
    sizes = rng.choice(np.arange(64, 513, 32), size=(n, 2))
    nu = 10.0 ** rng.uniform(-4, -2, size=n)
    grids = [MACGrid(int(a), int(b)) for a, b in sizes]
    t0 = time.perf_counter()
    stages = record_cavity_stages(grids, nu, args.record_steps, spinup=args.spinup_steps)
    record_s = time.perf_counter() - t0
    n_stages = len(stages[0])
    first = min(2 * args.fill_steps, n_stages - 1)
    timed = n_stages - first
    print(f"[cavities] recorded {n_stages} stages of {n} cavities in {record_s:.1f}s", file=sys.stderr, flush=True)
    masses = [np.array([stages[i][k][1] for i in range(n)]) for k in range(n_stages)]
    viscs = [np.array([stages[i][k][2] for i in range(n)]) for k in range(n_stages)]
    forms, outs = {}, {}
    flats = None
    for form in ("mixed+history", "mixed", "fp64"):
        precision = "fp64" if form == "fp64" else "mixed"
        t0 = time.perf_counter()
        batch = GpuBatchedStokes(grids, ncols=2, tolerance=1e-9, precision=precision,
                                 inner_tolerance=args.inner_tolerance)
        setup = time.perf_counter() - t0
        if flats is None:
            flats = [torch.as_tensor(batch.pack([stages[i][k][0] for i in range(n)])) for k in range(n_stages)]
        batch.solve_packed(flats[-1], mass=masses[-1], viscosity=viscs[-1])
        timings = batch.autotune(flats[-1], mass=masses[-1], viscosity=viscs[-1], repeats=1)
        if form == "mixed+history":
            batch.set_history(2, args.history_steps)
        for k in range(first):
            batch.solve_packed(flats[k], mass=masses[k], viscosity=viscs[k])
            if k % 2 == 1 and form == "mixed+history":
                batch.commit_history()
        quiesce()
        t0 = time.perf_counter()
        res, its, sweeps = [], 0, []
        for k in range(first, n_stages):
            out = batch.solve_packed(flats[k], mass=masses[k], viscosity=viscs[k])
            if k % 2 == 1 and form == "mixed+history":
                batch.commit_history()
            res.append(out)
            its += sum(sum(c) for c in out["iterations"])
            sweeps.append(batch.last_sweeps)
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        resume()
        worst_r = max(max(o["residual"]) for o in res)
        worst_d = max(max(o["divergence_inf"]) for o in res)
        forms[form] = {"cavity_stage_solves_per_s": n * timed / elapsed, "ms_per_stage": 1e3 * elapsed / timed,
                       "cg_iterations_per_rhs": its / (2 * n * timed), "sweeps_per_stage": sweeps,
                       "worst_residual": worst_r, "worst_divergence_inf": worst_d,
                       "meets_bar": worst_r <= BAR and worst_d <= BAR, "launch_shape": list(batch.launch),
                       "autotune_seconds": {f"{a},{b}": v for (a, b), v in timings.items()}, "setup_seconds": setup}
        outs[form] = (batch, res[-1])
        print(f"[cavities] {form}: {forms[form]}", file=sys.stderr, flush=True)
    del flats
    # The mixed forms against FP64 on the last stage, cavity by cavity (relative L2 of u).
    rel = {}
    bf, of = outs["fp64"]
    vo = bf.velocity_offsets
    for form in ("mixed+history", "mixed"):
        om = outs[form][1]
        worst = 0.0
        for i in range(n):
            a, b = om["velocity"][vo[2 * i]:vo[2 * i + 2]], of["velocity"][vo[2 * i]:vo[2 * i + 2]]
            worst = max(worst, float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b)))
        rel[form] = worst
    # Baseline: the same timed stage solves, one cavity at a time through the single path.
    singles = [GpuTensorStokes(g, tolerance=1e-9) for g in grids]
    for i in range(n):
        singles[i].solve(stages[i][0][0], mass=stages[i][0][1], viscosity=stages[i][0][2])
    quiesce()
    t0 = time.perf_counter()
    for k in range(first, n_stages):
        for i in range(n):
            singles[i].solve(stages[i][k][0], mass=stages[i][k][1], viscosity=stages[i][k][2])
    torch.cuda.synchronize()
    loop_s = time.perf_counter() - t0
    resume()
    loop = n * timed / loop_s
    # CPU baseline: the oracle (scipy.fft + the scipy-cg loop, one FFT worker as the
    # reference's default) on the cavities of the first timed stage in order, within a
    # bounded budget.
    import scipy.fft
    from oracle import mac_oracle as mo
    done, t0 = 0, time.perf_counter()
    with scipy.fft.set_workers(1):
        for i in range(n):
            g = grids[i]
            orc = mo.TensorStokes(mo.Grid(g.nx, g.ny, g.lx, g.ly), tolerance=1e-9)
            f, m, v = stages[i][first]
            orc.solve(f.cpu().numpy(), mass=m, viscosity=v)
            done += 1
            if time.perf_counter() - t0 > args.cpu_seconds:
                break
    cpu_s = time.perf_counter() - t0
    cpu = {"value": done / cpu_s, "unit": "cavity stage solves/s", "cores": 1, "kind": "port",
           "sample": f"cavities 0..{done - 1} of stage {first} (oracle TensorStokes, scipy.fft workers=1, host "
                     f"nproc={os.cpu_count()}), {cpu_s:.1f} s"}
    head = forms["mixed+history"]
    value = head["cavity_stage_solves_per_s"]
    return {"metric": "cavity stage solves/s, batched uneven lid-driven cavities", "value": value,
            "unit": "cavity stage solves/s", "n_gpus": 1, "steps": timed, "warmup": first,
            "ms_per_step": head["ms_per_stage"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "bf16-gemm/f64-cg",
            "data": f"recorded: sdirk2_mrsav steps {args.spinup_steps + 1}..{args.spinup_steps + args.record_steps} "
                    f"({n_stages} stage solves, 2 RHS each) of every cavity from rest with the lid moving at 1; "
                    f"stages {first}..{n_stages - 1} timed",
            "config": {"workload": f"{n} cavities, nx,ny uniform on {{64..512 step 32}}, nu log-uniform "
                                   f"[1e-4,1e-2], lid 1, sdirk2_mrsav dt=5e-4", "total_faces": bf.total_faces,
                       "form": "mixed+history", "history_steps": args.history_steps,
                       "inner_tolerance": args.inner_tolerance,
                       "bar": "residual<=1e-10 and divergence_inf<=1e-10 (absolute)"},
            "forms": forms, "max_rel_l2_velocity_vs_fp64": rel,
            "per_cavity_loop_fp64": {"value": loop, "unit": "cavity stage solves/s"},
            "speedup_vs_loop_fp64": value / loop,
            "speedup_vs_batched_fp64": value / forms["fp64"]["cavity_stage_solves_per_s"],
            "speedup_vs_batched_mixed_zero_start": value / forms["mixed"]["cavity_stage_solves_per_s"],
            "cpu_baseline": cpu, "record_seconds": record_s}


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
        solves0 = backend.solves
        quiesce()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            state = scheme.step(model, state, DT).state
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        resume()
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
            "vs_baseline": None, "dtype": "f32-fft/f64-refine (mixed form)",
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
        quiesce(hold=False)
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
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32-fft/f64-refine (mixed form)",
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
        rec = record_run(grid, 3, modes=min(32, n // 4))
        row = {}
        for precision in ("fp64", "mixed"):
            s = GpuTensorStokes(grid, tolerance=1e-10, krylov_tolerance=1e-14, precision=precision,
                                refine_tolerance=1e-10)
            for f, m, v in rec[:4]:
                s.solve(f, mass=m, viscosity=v)
            quiesce()
            t0 = time.perf_counter()
            for f, m, v in rec[4:]:
                s.solve(f, mass=m, viscosity=v)
            torch.cuda.synchronize()
            row[precision] = 1e3 * (time.perf_counter() - t0) / (len(rec) - 4)
            resume()
        row["mixed_speedup"] = row["fp64"] / row["mixed"]
        rows[str(n)] = row
        print(f"[sizes] {n}: {row}", file=sys.stderr, flush=True)
    return {"metric": "ms per two-RHS stage solve against grid size, plain FP64 and mixed",
            "value": rows["1024"]["mixed_speedup"], "unit": "x (mixed over FP64 at 1024^2)", "n_gpus": 1,
            "steps": 8, "warmup": 4, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f32-fft (power-of-two n) or bf16-gemm/f64-refine", "data": "recorded sdirk3_mrsav stage right-hand sides per size",
            "config": {"workload": "MAC n x n stage solves, n in 64..1024, zero start, residual 1e-10"},
            "sizes": rows}


def bench_stage1024(args) -> dict:
    import torch
    from macsav import MACGrid
    grid = MACGrid(args.n, args.n)
    kappa = NU * DT * STAGE_DIAG
    calibration_copy()
    W, K = args.warmup, args.steps
    fill = FILL_STEPS * STAGES
    run_steps = max(args.run_steps, FILL_STEPS + (W + K + STAGES - 1) // STAGES)
    t0 = time.perf_counter()
    rec = record_run(grid, run_steps, modes=args.modes)
    record_s = time.perf_counter() - t0
    print(f"[stage1024] recorded {len(rec)} stage solves in {record_s:.1f}s", file=sys.stderr, flush=True)
    net_info, net = None, None
    if not args.no_network:
        net, net_info = train_cold_start_network(grid)
        print(f"[stage1024] network {net_info}", file=sys.stderr, flush=True)
    forms, velocities, profs = {}, {}, {}
    wanted = set(args.forms.split(",")) if args.forms else set(FORMS)
    wanted |= {"fp64", "fp64+refine"}          # the baselines every line reports
    for form in FORMS:
        if form not in wanted or (form == "mixed+network" and net is None):
            continue
        profile = form == next(f for f in ("mixed+network", "mixed+history", "mixed", "fp64+refine")
                               if f in wanted and (f != "mixed+network" or net is not None))
        forms[form], profs[form], velocities[form] = replay(grid, rec, fill, W, K, form, net, profile)
        print(f"[stage1024] {form}: {forms[form]}", file=sys.stderr, flush=True)

    def rel(a_list, b_list):
        return max(float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))
                   for a, b in zip(a_list, b_list))
    ref_v = velocities["fp64+refine"]
    for form in forms:
        forms[form]["rel_l2_vs_fp64_refine_velocity"] = rel(velocities[form], ref_v)
    base = forms["fp64+refine"]["stage_solves_per_s"]
    for form in forms:
        forms[form]["speedup_vs_fp64_refine"] = forms[form]["stage_solves_per_s"] / base
        forms[form]["speedup_vs_fp64"] = forms[form]["stage_solves_per_s"] / forms["fp64"]["stage_solves_per_s"]
    runs = {}
    for form in ("fp64+refine", "mixed+history", "mixed+network"):
        if form in forms and not args.no_whole_run:
            runs[form] = whole_run(grid, rec, form, net)
            print(f"[stage1024] whole run {form}: {runs[form]}", file=sys.stderr, flush=True)
    head = next(f for f in ("mixed+network", "mixed+history", "mixed", "fp64+refine") if f in forms)
    main_stats, prof = forms[head], profs[head]
    line = {
        "metric": "stage solves/s at 1024^2 MAC, mixed+learned vs plain FP64 GPU, equal residual",
        "value": main_stats["stage_solves_per_s"],
        "unit": "stage solves/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": main_stats["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-fft/f64-refine" if prof.get("fft_launches", 0) > 0 else "bf16-gemm/f64-refine",
        "data": f"synthetic: stage solves {fill + W}..{fill + W + K - 1} of a recorded {run_steps}-step "
                f"sdirk3_mrsav run from the seeded stream-function field (modes<={args.modes}); the first "
                f"{fill} (history fill) and the next {W} are solved untimed",
        "config": {"workload": f"MAC {grid.nx}x{grid.ny} stage solve, 2 RHS, mass=1, viscosity={kappa:g} "
                               f"(nu=1e-3, dt=2e-4, sdirk3_mrsav)", "form": head,
                   "bar": "residual<=1e-10 and divergence_inf<=1e-10 (absolute)",
                   "baseline_form": "fp64+refine"},
        "solve": main_stats,
        "speedup_vs_plain_fp64_equal_bar": main_stats["stage_solves_per_s"] / base,
        "speedup_vs_reference_fp64": main_stats["speedup_vs_fp64"],
        "forms": forms,
        "whole_run": {"stage_solves": len(rec), "forms": runs},
        "warm_start_network": net_info,
        "record_seconds": record_s,
        "roofline": transform_roofline(prof),
    }
    if not args.no_cpu:
        f0, m0, v0 = rec[fill + W]
        pair = f0.cpu().numpy()
        line["cpu_baseline"] = cpu_baseline(grid, pair, v0, args.cpu_seconds, 1)
        cores = min(os.cpu_count() or 1, 16)
        line["cpu_baseline"]["all_cores"] = cpu_baseline(grid, pair, v0, args.cpu_seconds, cores)
    return line


def main() -> int:
    global HISTORY_STEPS
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60, help="timed stage solves")
    ap.add_argument("--warmup", type=int, default=10, help="untimed stage solves before them (after the fill)")
    ap.add_argument("--run-steps", type=int, default=12, help="steps of the recorded run (at least the fill "
                                                             "plus the warm-up and timed solves)")
    ap.add_argument("--n", type=int, default=None, help="grid size (n x n)")
    ap.add_argument("--modes", type=int, default=32)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-network", action="store_true")
    ap.add_argument("--no-whole-run", action="store_true", help="skip timing every solve of the run")
    ap.add_argument("--history-steps", type=int, default=HISTORY_STEPS,
                    help="steps the history extrapolation reaches back (1..6)")
    ap.add_argument("--forms", default="", help="comma-separated subset of the forms (the two FP64 baselines "
                                                "always run)")
    ap.add_argument("--workload", choices=["stage1024", "learned", "cavities", "run", "combined", "sizes"],
                    default="stage1024")
    ap.add_argument("--t-end", type=float, default=0.05, help="combined run: final time")
    ap.add_argument("--max-steps", type=int, default=2000, help="combined run: accepted-step cap")
    ap.add_argument("--cavities", type=int, default=256)
    ap.add_argument("--record-steps", type=int, default=8, help="cavities: sdirk2_mrsav steps recorded per cavity")
    ap.add_argument("--fill-steps", type=int, default=3, help="cavities: recorded steps solved untimed first")
    ap.add_argument("--spinup-steps", type=int, default=72, help="cavities: steps run before the recorded ones")
    ap.add_argument("--inner-tolerance", type=float, default=1e-4, help="cavities: inner CG tolerance, mixed form")
    ap.add_argument("--train-seeds", type=int, default=40)
    ap.add_argument("--test-seeds", type=int, default=10)
    ap.add_argument("--epochs", type=int, default=12)
    args = ap.parse_args()
    HISTORY_STEPS = args.history_steps
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (this build has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        return 2
    runners = {"cavities": bench_cavities, "sizes": bench_sizes, "combined": bench_combined, "learned": bench_learned,
               "run": bench_run, "stage1024": bench_stage1024}
    if args.workload in ("run", "stage1024"):
        args.n = args.n or 1024
    print(json.dumps(runners[args.workload](args)), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
