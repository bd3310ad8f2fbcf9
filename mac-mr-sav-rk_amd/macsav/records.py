"""Run results in the reference's results-file layout (schema_version 1), written as NPZ.

Restates `save_result` and `save_scalar_corrections` of the reference's
run_support/records.py:81-140: the same dataset names, shapes, dtypes and attributes,
with each HDF5 dataset path ("diagnostics/kinetic", "final/u", "roots/candidates", ...)
as an NPZ key and each file attribute as the key "attrs/<name>".  h5py is not a
dependency here, so the container is NPZ; a reader that maps keys back to HDF5 paths
sees the reference's file.  `load_result` reads one back (no pickles).

Device-model results (packed velocities in `State.u`, `State.v` None) are unpacked to
the reference's (ny, nx+1) / (ny+1, nx) face arrays through the model's grid.
"""
from __future__ import annotations

import json
from pathlib import Path
from typing import Any

import numpy as np

SCHEMA_VERSION = 1
_ROOTS = {1: 1, 2: 3, 4: 5}          # candidate-root capacity per stage count (records.py:106)


def _host(x: Any) -> np.ndarray:
    return np.asarray(x.cpu().numpy() if hasattr(x, "cpu") else x)


def _faces(state: Any, model: Any) -> tuple[np.ndarray, np.ndarray]:
    """The state's u (ny, nx+1) and v (ny+1, nx) face arrays on the host."""
    if state.v is not None:
        return _host(state.u), _host(state.v)
    if model is None:
        raise ValueError("A packed-velocity state needs the model to unpack it")
    return model.grid.unpack(_host(state.u).reshape(-1))


def scalar_corrections(stages: list[list[dict]], n_stages: int) -> tuple[dict[str, np.ndarray], str]:
    """The stage scalar and correction datasets and the `scalar_solver` attribute
    (records.py:81-101)."""
    out = {"stages/r": np.asarray([[s["r"] for s in row] for row in stages], dtype=float).reshape(-1, n_stages)}
    capacity = 2 if n_stages == 4 else 1
    counts = np.zeros((len(stages), n_stages), dtype=int)
    values = np.full((len(stages), n_stages, capacity), np.nan)
    predictors = np.full(counts.shape, np.nan)
    factors = np.full(counts.shape, np.nan)
    for n, row in enumerate(stages):
        for i, stage in enumerate(row):
            got = stage.get("corrections", [])
            counts[n, i] = len(got)
            values[n, i, :len(got)] = got
            predictors[n, i] = stage.get("correction_predictor", np.nan)
            factors[n, i] = stage.get("frozen_factor", np.nan)
    if np.any(counts):
        solver = "fixed_count_frozen_factor"
    elif any(s.get("root_count", 0) for row in stages for s in row):
        solver = "legacy_polynomial_roots"
    else:
        solver = "none"
    out.update({"corrections/count": counts, "corrections/values": values,
                "corrections/predictor": predictors, "corrections/frozen_factor": factors})
    return out, solver


def save_result(path: str | Path, result: Any, metrics: dict, n_stages: int = 2, *, n_roots: int | None = None,
                pressure_semantics: str | None = None, include_stage_pressure: bool = True,
                model: Any = None) -> Path:
    """Write `result` (an `integrate` / `integrate_adaptive` result) to `path` (NPZ,
    written to a temporary file and moved into place).  Arguments as the reference's
    `save_result` (records.py:103-140), plus `model` for device-model states."""
    path = Path(path)
    if n_roots is None:
        n_roots = _ROOTS[n_stages]
    data: dict[str, np.ndarray] = {
        "attrs/schema_version": np.asarray(SCHEMA_VERSION),
        "attrs/status": np.asarray(str(result.status)),
        "attrs/metrics_json": np.asarray(json.dumps(metrics, allow_nan=False)),
    }
    if pressure_semantics is not None:
        data["attrs/pressure_semantics"] = np.asarray(str(pressure_semantics))
    times = np.asarray(result.times, dtype=float)
    data["times"] = times
    data["dt"] = np.diff(times)
    for key in result.diagnostics[0]:
        data["diagnostics/" + key] = np.asarray([item[key] for item in result.diagnostics])
    u, v = _faces(result.final, model)
    data.update({"final/u": u, "final/v": v, "final/r": np.asarray(result.final.r),
                 "final/t": np.asarray(result.final.t)})
    if include_stage_pressure and result.final_stages:
        data["final/stage_pressure"] = np.stack([_host(s.pressure) for s in result.final_stages])
    data["snapshots/requested_times"] = np.asarray(result.snapshot_requests, dtype=float)
    data["snapshots/actual_times"] = np.asarray(result.snapshot_times, dtype=float)
    if result.snapshots:
        faces = [_faces(s, model) for s in result.snapshots]
        data["snapshots/u"] = np.stack([f[0] for f in faces])
        data["snapshots/v"] = np.stack([f[1] for f in faces])
        data["snapshots/r"] = np.asarray([s.r for s in result.snapshots])
    steps = len(result.stages)
    candidates = np.full((steps, n_stages, n_roots), np.nan)
    residuals = candidates.copy()
    counts = np.zeros((steps, n_stages), dtype=int)
    selected = np.zeros(counts.shape, dtype=float)
    for n, row in enumerate(result.stages):
        for i, stage in enumerate(row):
            count = stage["root_count"]
            counts[n, i] = count
            selected[n, i] = stage["r"]
            candidates[n, i, :count] = stage["candidates"]
            residuals[n, i, :count] = stage["root_residuals"]
    data.update({"roots/candidates": candidates, "roots/residuals": residuals, "roots/count": counts,
                 "roots/selected": selected})
    corrections, solver = scalar_corrections(result.stages, n_stages)
    data.update(corrections)
    data["attrs/scalar_solver"] = np.asarray(solver)
    for key in ("residual", "scalar_residual", "divergence_inf", "continuity_residual"):
        data["stages/" + key] = np.asarray([[s[key] for s in row] for row in result.stages],
                                           dtype=float).reshape(-1, n_stages)
    temporary = path.with_name(path.name + ".tmp")
    with open(temporary, "wb") as stream:
        np.savez_compressed(stream, **data)
    temporary.replace(path)
    return path


def load_result(path: str | Path) -> dict[str, Any]:
    """The datasets of a results file by HDF5 path, attributes under "attrs/" decoded to
    Python scalars (and `metrics` parsed from metrics_json)."""
    with np.load(Path(path), allow_pickle=False) as npz:
        out: dict[str, Any] = {}
        for key in npz.files:
            value = npz[key]
            out[key] = value.item() if key.startswith("attrs/") else value
    out["metrics"] = json.loads(out["attrs/metrics_json"])
    return out
