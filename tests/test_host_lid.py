"""The moving lid: the top wall sliding tangentially at speed U (no counterpart in the
reference's MAC model, which has fixed no-slip walls; the batched-cavity and combined
workloads name it).

It enters the MAC model as the viscous boundary load nu * 2U / hy^2 on the top row of u
faces, added to the force; the convection is unchanged because the wall's normal
velocity is zero, so no momentum is advected through it.  Pinned by
- the load against an explicitly ghosted five-point Laplacian (ghost 2U - u above the
  lid, -u below the floor, zero normal-wall faces), independent of the CSR K;
- the steady lid-driven cavity at Re = 100 against the centreline table of Ghia, Ghia
  and Shin (J. Comput. Phys. 48, 1982, Table I) and the steady energy balance
  (power of the wall load = viscous dissipation).  The table is itself about 3e-3 off
  the grid-converged centreline near its minimum (-0.2109 against about -0.214 by
  Richardson extrapolation of this discretisation's 32^2 and 64^2 values), so the bar is
  5e-3.
"""
import numpy as np
import pytest

from oracle import mac_oracle as mo


def _ghosted_minus_laplacian_u(u, lid, hx, hy):
    """-Laplacian on the interior u faces of u (ny, nx+1) with explicit ghost rows."""
    ny, nxp = u.shape
    full = np.zeros((ny + 2, nxp))
    full[1:-1] = u
    full[0] = -u[0]
    full[-1] = 2.0 * lid - u[-1]
    c = full[1:-1, 1:-1]
    lap = (full[1:-1, 2:] - 2 * c + full[1:-1, :-2]) / hx ** 2 + (full[2:, 1:-1] - 2 * c + full[:-2, 1:-1]) / hy ** 2
    return -lap


def test_lid_load_is_the_ghosted_laplacian():
    from macsav import MACGrid
    from macsav.mac.kernels import lid_load
    from macsav.mac.operators import MACOperators
    g = MACGrid(13, 9, 1.3, 0.7)
    K = MACOperators(g).K
    rng = np.random.default_rng(8)
    x = rng.normal(size=g.size)
    u, v = g.unpack(x)
    for lid in (0.0, 1.0, -2.5):
        got = g.unpack(K @ x - lid_load(g, lid))[0][:, 1:-1]
        want = _ghosted_minus_laplacian_u(u, lid, g.hx, g.hy)
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * np.abs(want).max())
    load = lid_load(g, 1.0)
    assert np.count_nonzero(load) == g.nx - 1 and np.all(load[g.nu:] == 0.0)
    assert np.all(lid_load(g, 0.0) == 0.0)


def test_model_force_carries_the_lid_load():
    from macsav import MACGrid
    from macsav.mac.kernels import lid_load
    from macsav.mac_ns import MACNavierStokes
    g = MACGrid(10, 8)
    backend = mo.TensorStokes(mo.Grid(10, 8, 1.0, 1.0), tolerance=1e-10)
    body = lambda t: np.full(g.size, t)
    m0 = MACNavierStokes(g, 0.02, force=body, backend=backend)
    m1 = MACNavierStokes(g, 0.02, force=body, backend=backend, lid=0.5)
    np.testing.assert_array_equal(m1.force(0.3), body(0.3) + 0.02 * lid_load(g, 0.5))
    np.testing.assert_array_equal(m0.force(0.3), body(0.3))
    x = np.random.default_rng(1).normal(size=g.size)
    np.testing.assert_array_equal(m1.nonlinear(x), m0.nonlinear(x))
    with pytest.raises(ValueError):
        MACNavierStokes(g, 0.02, backend=backend, lid=float("nan"))


# Ghia, Ghia & Shin (1982), Table I, Re = 100: u on the vertical centreline x = 1/2.
GHIA_Y = np.array([0.0547, 0.0703, 0.1016, 0.1719, 0.2813, 0.4531, 0.5, 0.6172, 0.7344, 0.8516,
                   0.9531, 0.9609, 0.9688, 0.9766])
GHIA_U = np.array([-0.03717, -0.04775, -0.06434, -0.10150, -0.15662, -0.21090, -0.20581, -0.13641,
                   0.00332, 0.23151, 0.68717, 0.73722, 0.78871, 0.84123])


def test_steady_lid_driven_cavity_matches_ghia_re100():
    from macsav import MACGrid
    from macsav.integrate import integrate
    from macsav.mac_ns import MACNavierStokes
    from macsav.schemes import make_scheme
    n = 32
    g = MACGrid(n, n)
    m = MACNavierStokes(g, 0.01, backend=mo.TensorStokes(mo.Grid(n, n, 1.0, 1.0), tolerance=1e-10), lid=1.0)
    res = integrate(m, make_scheme("sdirk2_mrsav"), m.state(0.0, np.zeros(g.size)), [0.04] * 400)
    assert res.status == "complete", res.error
    u, _ = g.unpack(m.vector(res.final))
    centre = np.interp(GHIA_Y, (np.arange(n) + 0.5) / n, u[:, n // 2])
    assert np.max(np.abs(centre - GHIA_U)) < 5e-3
    d = res.diagnostics[-1]
    assert abs(d["power"] - d["dissipation"]) < 1e-4 * d["dissipation"]
    assert d["divergence_inf"] < 1e-10
