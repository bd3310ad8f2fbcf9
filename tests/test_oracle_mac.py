"""Pins the CPU oracle (oracle/mac_oracle.py) to what mathematics and libraries fix.

Every check is independent of the formula it checks: library routines (scipy.fft,
scipy's cg, SuperLU, dense LAPACK), closed-form eigenvalues, exact identities of the
MAC operators and explicit loops on tiny grids.  Planted mistakes (a transposed
DST-II, a sign flip in D^T, a wrong wall diagonal) are shown to fail at the end.
"""
import numpy as np
import pytest
import scipy.sparse as sp
from scipy.fft import dct, dst, idct, idst
from scipy.sparse.linalg import LinearOperator, cg

from oracle import mac_oracle as mo


# ------------------------------------------------------------ transforms (a1, a4)
@pytest.mark.parametrize("n", [3, 5, 8, 17, 64])
@pytest.mark.parametrize("kind", ["dst1", "dst2", "dct2"])
def test_transform_matrices_match_scipy_fft(n, kind):
    x = np.random.default_rng(n).normal(size=(n, 3))
    q = mo.transform_matrix(n, kind)
    fwd, inv = {"dst1": (lambda a: dst(a, 1, axis=0, norm="ortho"), lambda a: idst(a, 1, axis=0, norm="ortho")),
                "dst2": (lambda a: dst(a, 2, axis=0, norm="ortho"), lambda a: idst(a, 2, axis=0, norm="ortho")),
                "dct2": (lambda a: dct(a, 2, axis=0, norm="ortho"), lambda a: idct(a, 2, axis=0, norm="ortho"))}[kind]
    np.testing.assert_allclose(q @ x, fwd(x), atol=1e-14 * n)
    np.testing.assert_allclose(q.T @ x, inv(x), atol=1e-14 * n)
    np.testing.assert_allclose(q @ q.T, np.eye(n), atol=1e-14 * n)


def _second_difference(n, ends):
    a = 2.0 * np.eye(n) - np.eye(n, k=1) - np.eye(n, k=-1)
    a[0, 0] = a[-1, -1] = ends
    return a


@pytest.mark.parametrize("n", [4, 9, 32])
@pytest.mark.parametrize("kind,ends", [("dst1", 2.0), ("dst2", 3.0), ("dct2", 1.0)])
def test_eigenvalues_diagonalise_the_1d_operators(n, kind, ends):
    # Dirichlet (dst1), odd half-cell wall (dst2, diagonal 3), Neumann (dct2, diagonal 1).
    h = 0.37
    q = mo.transform_matrix(n, kind)
    a = _second_difference(n, ends) / h**2
    np.testing.assert_allclose(q @ a @ q.T, np.diag(mo.eigenvalues_1d(n, h, kind)), atol=1e-11 / h**2)


# ------------------------------------------------------------ operators (a13)
def _sine_mode(g):
    xu, yu = np.arange(g.nx + 1) * g.hx, (np.arange(g.ny) + .5) * g.hy
    xv, yv = (np.arange(g.nx) + .5) * g.hx, np.arange(g.ny + 1) * g.hy
    u = np.outer(np.sin(np.pi * yu / g.ly), np.sin(np.pi * xu / g.lx))
    v = np.outer(np.sin(np.pi * yv / g.ly), np.sin(np.pi * xv / g.lx))
    u[:, [0, -1]] = 0.0
    v[[0, -1], :] = 0.0
    return g.pack(u, v)


def test_K_sine_mode_has_the_analytic_eigenvalue():
    g = mo.Grid(9, 7, 1.3, 0.8)
    x = _sine_mode(g)
    lam = (2 - 2 * np.cos(np.pi * g.hx / g.lx)) / g.hx**2 + (2 - 2 * np.cos(np.pi * g.hy / g.ly)) / g.hy**2
    np.testing.assert_allclose(mo.Operators(g).K @ x, lam * x, rtol=1e-12, atol=1e-12)


def test_G_is_minus_D_transpose_and_kills_constants_and_D_kills_curls():
    g = mo.Grid(9, 7, 1.3, 0.8)
    ops = mo.Operators(g)
    rng = np.random.default_rng(1)
    v, p = rng.normal(size=g.size), rng.normal(size=g.np)
    np.testing.assert_allclose(np.dot(ops.G @ p, v), -np.dot(p, ops.D @ v), rtol=1e-12)
    np.testing.assert_array_equal(ops.G @ np.ones(g.np), np.zeros(g.size))
    psi = rng.normal(size=(g.ny + 1, g.nx + 1))
    psi[[0, -1], :] = 0
    psi[:, [0, -1]] = 0
    u = np.diff(psi, axis=0) / g.hy
    w = -np.diff(psi, axis=1) / g.hx
    np.testing.assert_allclose(ops.D @ g.pack(u, w), 0.0, atol=1e-12)


def test_K_is_symmetric_positive_definite_with_expected_stencil():
    g = mo.Grid(6, 5, 1.0, 1.0)
    K = mo.Operators(g).K.toarray()
    np.testing.assert_allclose(K, K.T)
    assert np.linalg.eigvalsh(K).min() > 0
    # u face (j=0, i=2): tangential wall row in y -> diagonal 2/hx^2 + 3/hy^2.
    assert K[2, 2] == pytest.approx(2 / g.hx**2 + 3 / g.hy**2)
    # v face (j=1, i=0): tangential wall column in x -> 3/hx^2 + 2/hy^2.
    idx = g.nu + 1 * g.nx + 0
    assert K[idx, idx] == pytest.approx(3 / g.hx**2 + 2 / g.hy**2)


# ------------------------------------------------------- matrix-free kernels (a7)
def test_packed_divergence_and_transpose_match_csr_and_are_adjoint():
    g = mo.Grid(11, 7, 2.0, 0.7)
    ops = mo.Operators(g)
    rng = np.random.default_rng(79)
    vel, pre = rng.normal(size=(g.size, 2)), rng.normal(size=(g.np, 2))
    dv = mo.packed_divergence(vel, g)
    dtp = mo.packed_divergence_transpose(pre, g)
    np.testing.assert_allclose(dv, ops.D @ vel, atol=1e-13)
    np.testing.assert_allclose(dtp, ops.D.T @ pre, atol=1e-13)
    for c in range(2):
        a, b = np.dot(dv[:, c], pre[:, c]), np.dot(vel[:, c], dtp[:, c])
        assert abs(a - b) < 2e-12 * max(1.0, abs(a))


def _loop_divergence(vel, g):
    # The reference kernel's loop (kernels.py:46-67), term by term.
    out = np.empty(g.np)
    for j in range(g.ny):
        for i in range(g.nx):
            value = 0.0
            if i < g.nx - 1:
                value += vel[j * (g.nx - 1) + i] / g.hx
            if i > 0:
                value -= vel[j * (g.nx - 1) + i - 1] / g.hx
            if j < g.ny - 1:
                value += vel[g.nu + j * g.nx + i] / g.hy
            if j > 0:
                value -= vel[g.nu + (j - 1) * g.nx + i] / g.hy
            out[j * g.nx + i] = value
    return out


def test_vectorised_divergence_is_bitwise_the_loop():
    g = mo.Grid(7, 6, 1.3, 0.8)
    vel = np.random.default_rng(3).normal(size=g.size)
    np.testing.assert_array_equal(mo.packed_divergence(vel, g), _loop_divergence(vel, g))


def _loop_convection(u, v, hx, hy):
    ny, nx = u.shape[0], u.shape[1] - 1
    cu, cv = np.zeros_like(u), np.zeros_like(v)
    for j in range(ny):
        for i in range(1, nx):
            e = 0.5 * (u[j, i] + u[j, i + 1]); w = 0.5 * (u[j, i] + u[j, i - 1])
            n = 0.5 * (v[j + 1, i - 1] + v[j + 1, i]); s = 0.5 * (v[j, i - 1] + v[j, i])
            un = u[j + 1, i] if j < ny - 1 else -u[j, i]
            us = u[j - 1, i] if j > 0 else -u[j, i]
            cu[j, i] = (e * 0.5 * (u[j, i] + u[j, i + 1]) - w * 0.5 * (u[j, i] + u[j, i - 1])) / hx
            cu[j, i] += (n * 0.5 * (u[j, i] + un) - s * 0.5 * (u[j, i] + us)) / hy
            cu[j, i] -= 0.5 * u[j, i] * ((e - w) / hx + (n - s) / hy)
    for j in range(1, ny):
        for i in range(nx):
            e = 0.5 * (u[j - 1, i + 1] + u[j, i + 1]); w = 0.5 * (u[j - 1, i] + u[j, i])
            n = 0.5 * (v[j, i] + v[j + 1, i]); s = 0.5 * (v[j, i] + v[j - 1, i])
            ve = v[j, i + 1] if i < nx - 1 else -v[j, i]
            vw = v[j, i - 1] if i > 0 else -v[j, i]
            cv[j, i] = (e * 0.5 * (v[j, i] + ve) - w * 0.5 * (v[j, i] + vw)) / hx
            cv[j, i] += (n * 0.5 * (v[j, i] + v[j + 1, i]) - s * 0.5 * (v[j, i] + v[j - 1, i])) / hy
            cv[j, i] -= 0.5 * v[j, i] * ((e - w) / hx + (n - s) / hy)
    return cu, cv


def test_convection_matches_explicit_loop_and_is_skew():
    g = mo.Grid(11, 7, 2.0, 0.7)
    x = np.random.default_rng(73).normal(size=g.size)
    u, v = g.unpack(x)
    cu, cv = mo.convection(u, v, g.hx, g.hy)
    lu, lv = _loop_convection(u, v, g.hx, g.hy)
    np.testing.assert_array_equal(cu, lu)
    np.testing.assert_array_equal(cv, lv)
    # Skew symmetry (N(u), u)_h = 0 for any field (docs/numerics.md:28).
    assert abs(g.inner(g.pack(cu, cv), x)) < 1e-11 * g.inner(x, x) * (1 / g.hx)


# ------------------------------------------------------------------ PCG (a11)
def test_pcg_matches_scipy_cg_iterates_and_counts():
    rng = np.random.default_rng(5)
    n = 40
    a = rng.normal(size=(n, n))
    A = a @ a.T + n * np.eye(n)
    m = np.diag(1 / np.diag(A))
    b = rng.normal(size=n)
    counts = []
    xs, info = cg(LinearOperator((n, n), matvec=lambda x: A @ x), b,
                  M=LinearOperator((n, n), matvec=lambda x: m @ x), rtol=1e-12, atol=0.0, maxiter=100,
                  callback=lambda _: counts.append(1))
    mine = []
    xo, io = mo.pcg(lambda x: A @ x, lambda x: m @ x, b, 1e-12, 100, callback=lambda _: mine.append(1))
    assert info == io == 0 and len(counts) == len(mine)
    np.testing.assert_array_equal(xo, xs)
    # Exhausted loop reports info = maxiter (no final check), like scipy.
    _, i2 = cg(A, b, M=m, rtol=1e-12, atol=0.0, maxiter=3)
    _, j2 = mo.pcg(lambda x: A @ x, lambda x: m @ x, b, 1e-12, 3)
    assert i2 == j2 == 3


# ---------------------------------------------------------- the solve (a10, a18)
def test_tensor_matches_direct_and_dense_for_multi_rhs_and_limit_shifts():
    g = mo.Grid(7, 5, 2.0, 1.0)
    rhs = np.random.default_rng(19).normal(size=(g.size, 2))
    tensor = mo.TensorStokes(g, tolerance=1e-10)
    for mass, visc in ((1.0, 0.07), (1.0, 0.0), (0.0, 0.07)):
        want = mo.DirectStokes(g, tolerance=1e-10).solve(rhs, mass=mass, viscosity=visc)
        got = tensor.solve(rhs, mass=mass, viscosity=visc)
        u, p = mo.dense_saddle_solve(g, rhs, mass, visc)
        np.testing.assert_allclose(got.velocity, want.velocity, atol=3e-11)
        np.testing.assert_allclose(got.pressure, want.pressure, atol=3e-11)
        np.testing.assert_allclose(got.velocity, u, atol=3e-11)
        np.testing.assert_allclose(got.pressure.reshape(g.np, -1), p, atol=3e-11)
        assert got.residual < 1e-10 and got.divergence_inf < 1e-10
        assert abs(got.pressure.mean(axis=(0, 1))).max() < 1e-12


def test_gradient_rhs_gives_zero_velocity_and_the_pressure():
    g = mo.Grid(8, 6)
    ops = mo.Operators(g)
    p = np.random.default_rng(23).normal(size=g.np)
    p -= p.mean()
    solver = mo.TensorStokes(g, tolerance=1e-10)
    for visc in (0.02, 0.05, 0.11):
        res = solver.solve(ops.G @ p, mass=1.0, viscosity=visc)
        np.testing.assert_allclose(res.velocity, 0.0, atol=2e-11)
        np.testing.assert_allclose(res.pressure.ravel(), p, atol=2e-10)
    assert solver.solves == 3 and solver.total_iterations > 0


def test_dense_transform_form_reproduces_fft_form_and_counts():
    g = mo.Grid(11, 7, 2.0, 1.0)
    rhs = np.random.default_rng(37).normal(size=(g.size, 2))
    a, b = mo.TensorStokes(g, 1e-10, transforms="fft"), mo.TensorStokes(g, 1e-10, transforms="dense")
    for mass, visc in ((1.0, 0.03), (0.7, 0.09), (0.0, 0.05)):
        ra, rb = a.solve(rhs, mass=mass, viscosity=visc), b.solve(rhs, mass=mass, viscosity=visc)
        np.testing.assert_allclose(ra.velocity, rb.velocity, atol=2e-11)
        np.testing.assert_allclose(ra.pressure, rb.pressure, atol=2e-11)
        assert a.last_iterations == b.last_iterations


def test_multi_column_equals_single_column_solves():
    g = mo.Grid(9, 8, 1.0, 1.0)
    rhs = np.random.default_rng(41).normal(size=(g.size, 2))
    s = mo.TensorStokes(g, 1e-10)
    both = s.solve(rhs, mass=1.0, viscosity=0.01)
    its = s.last_iterations
    for c in range(2):
        one = s.solve(rhs[:, c], mass=1.0, viscosity=0.01)
        np.testing.assert_allclose(one.velocity, both.velocity[:, c], atol=1e-12)
        assert s.last_iterations == (its[c],)


def test_cg_failure_and_invalid_inputs_raise_like_the_reference():
    g = mo.Grid(8, 6)
    rhs = np.random.default_rng(2).normal(size=g.size)
    with pytest.raises(RuntimeError, match="Pressure Schur CG failed"):
        mo.TensorStokes(g, max_iterations=1).solve(rhs, mass=1.0, viscosity=0.1)
    with pytest.raises(ValueError):
        mo.TensorStokes(g).solve(rhs, mass=-1.0, viscosity=0.1)
    with pytest.raises(ValueError):
        mo.TensorStokes(g).solve(np.full(g.size, np.nan), mass=1.0, viscosity=0.1)


# --------------------------------------------------------- planted mistakes
def test_planted_mistakes_fail_the_pins(monkeypatch):
    # A transposed DST-II is no longer scipy's transform.
    q = mo.transform_matrix(8, "dst2").T
    x = np.random.default_rng(0).normal(size=8)
    assert np.abs(q @ x - dst(x, 2, norm="ortho")).max() > 1e-3
    # A sign error in D^T breaks the adjoint identity.
    g = mo.Grid(7, 5)
    vel, pre = np.random.default_rng(1).normal(size=g.size), np.random.default_rng(2).normal(size=g.np)
    bad = -mo.packed_divergence_transpose(pre, g)
    assert abs(np.dot(mo.packed_divergence(vel, g), pre) - np.dot(vel, bad)) > 1e-3
    # A wrong tangential wall diagonal (2 instead of 3) breaks the sine-mode eigenvalue.
    monkeypatch.setattr(mo, "laplacian_1d", lambda n, h, wall: sp.diags(
        [-np.ones(n - 1), np.full(n, 2.0), -np.ones(n - 1)], [-1, 0, 1], format="csr") / h**2)
    g2 = mo.Grid(9, 7, 1.3, 0.8)
    xs = _sine_mode(g2)
    lam = (2 - 2 * np.cos(np.pi * g2.hx / g2.lx)) / g2.hx**2 + (2 - 2 * np.cos(np.pi * g2.hy / g2.ly)) / g2.hy**2
    assert np.abs(mo.Operators(g2).K @ xs - lam * xs).max() > 1e-3
