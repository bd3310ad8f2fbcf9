"""The results writer: the reference's results-file layout (run_support/records.py:81-140)
as NPZ keys, checked on the reference's own record smoke case
(tests/core/test_run_support.py:47-49) and on real sdirk2/3_mrsav runs, where every
dataset is rebuilt here from the run's result object."""
import json

import numpy as np
import pytest

from oracle import mac_oracle as mo


def test_reference_record_smoke_case(tmp_path):
    from macsav.core import State
    from macsav.integrate import Result
    from macsav.records import load_result, save_result
    result = Result(State(0.0, np.zeros((2, 3)), np.zeros((3, 2))), [0.0], [{"kinetic": 0.0, "r": 0.0}])
    path = save_result(tmp_path / "results.npz", result, {}, n_stages=1)
    got = load_result(path)
    assert got["attrs/schema_version"] == 1 and got["attrs/status"] == "complete"
    assert got["metrics"] == {} and got["attrs/scalar_solver"] == "none"
    np.testing.assert_array_equal(got["times"], [0.0])
    assert got["dt"].shape == (0,)
    assert got["final/u"].shape == (2, 3) and got["final/v"].shape == (3, 2)
    assert got["roots/candidates"].shape == (0, 1, 1) and got["stages/r"].shape == (0, 1)
    assert got["corrections/values"].shape == (0, 1, 1)
    assert "final/stage_pressure" not in got and "snapshots/u" not in got
    assert not (tmp_path / "results.npz.tmp").exists()


@pytest.mark.parametrize("scheme, n_stages", [("sdirk2_mrsav", 2), ("sdirk3_mrsav", 4)])
def test_run_record_datasets(tmp_path, scheme, n_stages):
    from macsav import MACGrid
    from macsav.fields import stream_function_velocity
    from macsav.integrate import integrate
    from macsav.mac_ns import MACNavierStokes
    from macsav.records import load_result, save_result
    from macsav.schemes import make_scheme
    g = MACGrid(12, 10)
    m = MACNavierStokes(g, 0.01, backend=mo.TensorStokes(mo.Grid(12, 10, 1.0, 1.0), tolerance=1e-10), lid=1.0)
    res = integrate(m, make_scheme(scheme), m.state(0.0, stream_function_velocity(g, seed=1)), [1e-3] * 4,
                    snapshots=[0.0, 2e-3])
    metrics = {"steps": 4, "scheme": scheme}
    got = load_result(save_result(tmp_path / "r.npz", res, metrics, n_stages=n_stages, pressure_semantics="p"))
    assert got["metrics"] == metrics and json.loads(got["attrs/metrics_json"]) == metrics
    assert got["attrs/pressure_semantics"] == "p"
    np.testing.assert_array_equal(got["times"], res.times)
    np.testing.assert_array_equal(got["dt"], np.diff(res.times))
    for key in res.diagnostics[0]:
        np.testing.assert_array_equal(got["diagnostics/" + key], [d[key] for d in res.diagnostics])
    np.testing.assert_array_equal(got["final/u"], res.final.u)
    np.testing.assert_array_equal(got["final/v"], res.final.v)
    assert got["final/r"] == res.final.r and got["final/t"] == res.final.t
    np.testing.assert_array_equal(got["final/stage_pressure"], np.stack([s.pressure for s in res.final_stages]))
    assert got["final/stage_pressure"].shape == (n_stages, g.ny, g.nx)
    np.testing.assert_array_equal(got["snapshots/actual_times"], res.snapshot_times)
    np.testing.assert_array_equal(got["snapshots/u"], np.stack([s.u for s in res.snapshots]))
    steps = len(res.stages)
    n_roots = {2: 3, 4: 5}[n_stages]
    assert got["roots/candidates"].shape == (steps, n_stages, n_roots)
    np.testing.assert_array_equal(got["roots/count"], [[s["root_count"] for s in row] for row in res.stages])
    np.testing.assert_array_equal(got["stages/r"], [[s["r"] for s in row] for row in res.stages])
    np.testing.assert_array_equal(got["roots/selected"], got["stages/r"])
    for key in ("residual", "scalar_residual", "divergence_inf", "continuity_residual"):
        np.testing.assert_array_equal(got["stages/" + key], [[s[key] for s in row] for row in res.stages])
    counts = np.array([[len(s["corrections"]) for s in row] for row in res.stages])
    np.testing.assert_array_equal(got["corrections/count"], counts)
    assert got["corrections/values"].shape == (steps, n_stages, 2 if n_stages == 4 else 1)
    for n, row in enumerate(res.stages):
        for i, s in enumerate(row):
            np.testing.assert_array_equal(got["corrections/values"][n, i, :counts[n, i]], s["corrections"])
            assert np.all(np.isnan(got["corrections/values"][n, i, counts[n, i]:]))
    assert got["attrs/scalar_solver"] == ("fixed_count_frozen_factor" if counts.any() else "none")


def test_packed_states_need_the_model(tmp_path):
    from macsav import MACGrid
    from macsav.core import State
    from macsav.integrate import Result
    from macsav.mac_ns import MACNavierStokes
    from macsav.records import load_result, save_result
    g = MACGrid(6, 5)
    m = MACNavierStokes(g, 0.1, backend=mo.TensorStokes(mo.Grid(6, 5, 1.0, 1.0)))
    x = np.arange(g.size, dtype=float)
    result = Result(State(0.5, x, None, 0.25), [0.0, 0.5], [{"kinetic": 1.0}, {"kinetic": 2.0}])
    with pytest.raises(ValueError):
        save_result(tmp_path / "a.npz", result, {}, n_stages=1)
    got = load_result(save_result(tmp_path / "a.npz", result, {}, n_stages=1, model=m))
    u, v = g.unpack(x)
    np.testing.assert_array_equal(got["final/u"], u)
    np.testing.assert_array_equal(got["final/v"], v)
