"""NumPy/SciPy restatement of the reference's MAC Stokes stage solve (test oracle).

TEST INFRASTRUCTURE ONLY: imported by `tests/`, `__graft_entry__.smoke()` and the
`cpu_baseline` leg of `bench.py`, never by the product package.

What each piece follows in the reference, and what pins it (the checks live in
`tests/test_oracle_*.py`):

* `Grid`                     grid.py:8-71          pinned by layout identities (pack/unpack round trip)
* `laplacian_1d`, `Operators` operators.py:6-29    pinned by the analytic sine-mode eigenvalue of K,
                                                    G = -D^T, G 1 = 0, D curl(psi) = 0 (closed forms)
* `packed_divergence[_t]`    kernels.py:46-132     pinned against the CSR D / D^T and the adjoint identity
* `convection`               kernels.py:6-35       pinned against an explicit per-face loop on tiny grids,
                                                    skew symmetry (N(u),u)_h = 0 and zero on constants
* `transform_matrix`         tensor_stokes.py:58-101 (scipy.fft dst/dct norm="ortho") pinned against scipy.fft
* `eigenvalues_1d`           tensor_stokes.py:31-38 pinned: Q T Q^T is diagonal with these entries
* `pcg`                      scipy 1.15.3 `sparse/linalg/_isolve/iterative.py` `cg` (called at
                             tensor_stokes.py:313-321) pinned against scipy's own `cg` (iterates and
                             iteration counts)
* `TensorStokes`             tensor_stokes.py:118-364 pinned against `DirectStokes` and a dense
                             `np.linalg.solve` of the full saddle system (brute force, <= 12x10),
                             gradient-RHS special case (zero velocity), limit shifts (1,0), (0,nu)
* `DirectStokes`             stokes.py:42-86       pinned against dense `np.linalg.solve`
"""
from __future__ import annotations

from dataclasses import dataclass
from time import perf_counter

import numpy as np
import scipy.sparse as sp
from scipy.fft import dct, dst, idct, idst
from scipy.sparse.linalg import splu


# --------------------------------------------------------------------------- grid
@dataclass(frozen=True)
class Grid:
    """Uniform MAC grid; packed vector = [u interior (ny, nx-1) ; v interior (ny-1, nx)].

    Follows grid.py:8-71 (counts, spacings, pack/unpack, area-weighted inner).
    """

    nx: int
    ny: int
    lx: float = 1.0
    ly: float = 1.0

    @property
    def hx(self) -> float:
        return self.lx / self.nx

    @property
    def hy(self) -> float:
        return self.ly / self.ny

    @property
    def nu(self) -> int:
        return (self.nx - 1) * self.ny

    @property
    def nv(self) -> int:
        return self.nx * (self.ny - 1)

    @property
    def size(self) -> int:
        return self.nu + self.nv

    @property
    def np(self) -> int:
        return self.nx * self.ny

    @property
    def area(self) -> float:
        return self.hx * self.hy

    def pack(self, u: np.ndarray, v: np.ndarray) -> np.ndarray:
        return np.concatenate((u[:, 1:-1].ravel(), v[1:-1, :].ravel()))

    def unpack(self, velocity: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
        u = np.zeros((self.ny, self.nx + 1))
        v = np.zeros((self.ny + 1, self.nx))
        u[:, 1:-1] = velocity[: self.nu].reshape(self.ny, self.nx - 1)
        v[1:-1, :] = velocity[self.nu:].reshape(self.ny - 1, self.nx)
        return u, v

    def inner(self, a: np.ndarray, b: np.ndarray) -> float:
        return float(self.area * np.dot(a, b))


# ---------------------------------------------------------------------- operators
def laplacian_1d(n: int, spacing: float, half_cell_wall: bool) -> sp.csr_matrix:
    """tridiag(-1, 2, -1)/h^2 with wall diagonal 3 for a half-cell tangential wall
    (operators.py:6-9)."""
    diagonal = np.full(n, 2.0)
    if half_cell_wall:
        diagonal[[0, -1]] = 3.0
    off = -np.ones(n - 1)
    return sp.diags([off, diagonal, off], [-1, 0, 1], format="csr") / spacing**2


class Operators:
    """CSR K (block-diagonal vector Laplacian), D (cell divergence), G = -D^T
    (operators.py:11-29).  D is assembled with vectorised index maps instead of the
    reference's per-cell loop; the entries and their order per row are the same."""

    def __init__(self, grid: Grid) -> None:
        self.grid = grid
        nx, ny = grid.nx, grid.ny
        ku = sp.kron(sp.eye(ny), laplacian_1d(nx - 1, grid.hx, False)) + sp.kron(
            laplacian_1d(ny, grid.hy, True), sp.eye(nx - 1)
        )
        kv = sp.kron(sp.eye(ny - 1), laplacian_1d(nx, grid.hx, True)) + sp.kron(
            laplacian_1d(ny - 1, grid.hy, False), sp.eye(nx)
        )
        self.K = sp.block_diag([ku, kv], format="csr")
        j, i = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
        j, i = j.ravel(), i.ravel()
        cell = j * nx + i
        rows, cols, vals = [], [], []
        east = i < nx - 1
        rows.append(cell[east]); cols.append((j * (nx - 1) + i)[east]); vals.append(np.full(east.sum(), 1 / grid.hx))
        west = i > 0
        rows.append(cell[west]); cols.append((j * (nx - 1) + i - 1)[west]); vals.append(np.full(west.sum(), -1 / grid.hx))
        north = j < ny - 1
        rows.append(cell[north]); cols.append((grid.nu + j * nx + i)[north]); vals.append(np.full(north.sum(), 1 / grid.hy))
        south = j > 0
        rows.append(cell[south]); cols.append((grid.nu + (j - 1) * nx + i)[south]); vals.append(np.full(south.sum(), -1 / grid.hy))
        self.D = sp.csr_matrix(
            (np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
            shape=(grid.np, grid.size),
        )
        self.G = -self.D.T.tocsr()


# ----------------------------------------------------------- matrix-free kernels
def packed_divergence(velocity: np.ndarray, grid: Grid) -> np.ndarray:
    """D applied to packed columns (size, k) -> (np, k), kernels.py:46-67.

    Same per-cell term order as the reference loop: +u_e/hx, -u_w/hx, +v_n/hy, -v_s/hy,
    each term divided by the spacing before it is added (bitwise reproducible)."""
    nx, ny, hx, hy = grid.nx, grid.ny, grid.hx, grid.hy
    vel = velocity if velocity.ndim == 2 else velocity[:, None]
    k = vel.shape[1]
    u = vel[: grid.nu].reshape(ny, nx - 1, k)
    v = vel[grid.nu:].reshape(ny - 1, nx, k)
    out = np.zeros((ny, nx, k))
    out[:, : nx - 1] += u / hx
    out[:, 1:] -= u / hx
    out[: ny - 1, :] += v / hy
    out[1:, :] -= v / hy
    out = out.reshape(grid.np, k)
    return out if velocity.ndim == 2 else out[:, 0]


def packed_divergence_transpose(pressure: np.ndarray, grid: Grid) -> np.ndarray:
    """D^T applied to (np, k) -> (size, k), kernels.py:89-112: (p_left - p_right)/hx on
    u faces and (p_lower - p_upper)/hy on v faces."""
    nx, ny, hx, hy = grid.nx, grid.ny, grid.hx, grid.hy
    pre = pressure if pressure.ndim == 2 else pressure[:, None]
    k = pre.shape[1]
    p = pre.reshape(ny, nx, k)
    u = (p[:, :-1] - p[:, 1:]) / hx
    v = (p[:-1, :] - p[1:, :]) / hy
    out = np.concatenate((u.reshape(grid.nu, k), v.reshape(grid.nv, k)))
    return out if pressure.ndim == 2 else out[:, 0]


def convection(u: np.ndarray, v: np.ndarray, hx: float, hy: float) -> tuple[np.ndarray, np.ndarray]:
    """Skew dual-cell centred convection with odd tangential wall ghosts, kernels.py:6-35.

    u is (ny, nx+1), v is (ny+1, nx).  Vectorised over faces with the loop's own
    operation order ((a*0.5)*(b+c), then /h), so it reproduces the loop bitwise."""
    ny, nxp = u.shape
    nx = nxp - 1
    cu = np.zeros_like(u)
    cv = np.zeros_like(v)
    c = u[:, 1:nx]
    east = 0.5 * (c + u[:, 2:])
    west = 0.5 * (c + u[:, : nx - 1])
    north = 0.5 * (v[1:, : nx - 1] + v[1:, 1:nx])
    south = 0.5 * (v[:-1, : nx - 1] + v[:-1, 1:nx])
    un = np.empty_like(c)
    un[:-1] = u[1:, 1:nx]
    un[-1] = -c[-1]
    us = np.empty_like(c)
    us[1:] = u[:-1, 1:nx]
    us[0] = -c[0]
    val = (east * 0.5 * (c + u[:, 2:]) - west * 0.5 * (c + u[:, : nx - 1])) / hx
    val = val + (north * 0.5 * (c + un) - south * 0.5 * (c + us)) / hy
    val = val - 0.5 * c * ((east - west) / hx + (north - south) / hy)
    cu[:, 1:nx] = val
    c = v[1:ny, :]
    east = 0.5 * (u[: ny - 1, 1:] + u[1:ny, 1:])
    west = 0.5 * (u[: ny - 1, :nx] + u[1:ny, :nx])
    north = 0.5 * (c + v[2:, :])
    south = 0.5 * (c + v[: ny - 1, :])
    ve = np.empty_like(c)
    ve[:, :-1] = c[:, 1:]
    ve[:, -1] = -c[:, -1]
    vw = np.empty_like(c)
    vw[:, 1:] = c[:, :-1]
    vw[:, 0] = -c[:, 0]
    val = (east * 0.5 * (c + ve) - west * 0.5 * (c + vw)) / hx
    val = val + (north * 0.5 * (c + v[2:, :]) - south * 0.5 * (c + v[: ny - 1, :])) / hy
    val = val - 0.5 * c * ((east - west) / hx + (north - south) / hy)
    cv[1:ny, :] = val
    return cu, cv


# ------------------------------------------------------------------- transforms
def eigenvalues_1d(n: int, spacing: float, kind: str) -> np.ndarray:
    """(2 - 2 cos theta_k)/h^2 with theta_k = pi k/n (dct2, k=0..n-1), pi k/(n+1)
    (dst1, k=1..n), pi k/n (dst2, k=1..n); tensor_stokes.py:31-38."""
    if kind == "dct2":
        theta = np.pi * np.arange(n) / n
    elif kind == "dst1":
        theta = np.pi * np.arange(1, n + 1) / (n + 1)
    elif kind == "dst2":
        theta = np.pi * np.arange(1, n + 1) / n
    else:
        raise ValueError(kind)
    return (2.0 - 2.0 * np.cos(theta)) / spacing**2


def transform_matrix(n: int, kind: str) -> np.ndarray:
    """Orthonormal analysis matrix Q (forward = Q @ x, inverse = Q.T @ y) matching
    scipy.fft dst/dct with norm="ortho" (tensor_stokes.py:58-82)."""
    k = np.arange(n)[:, None].astype(float)
    j = np.arange(n)[None, :].astype(float)
    if kind == "dst1":
        return np.sqrt(2.0 / (n + 1)) * np.sin(np.pi * (k + 1) * (j + 1) / (n + 1))
    if kind == "dst2":
        q = np.sqrt(2.0 / n) * np.sin(np.pi * (k + 1) * (2 * j + 1) / (2 * n))
        q[n - 1] /= np.sqrt(2.0)
        return q
    if kind == "dct2":
        q = np.sqrt(2.0 / n) * np.cos(np.pi * k * (2 * j + 1) / (2 * n))
        q[0] /= np.sqrt(2.0)
        return q
    raise ValueError(kind)


def _fft_axis(values: np.ndarray, kind: str, axis: int, inverse: bool) -> np.ndarray:
    if kind == "dct2":
        return (idct if inverse else dct)(values, type=2, axis=axis, norm="ortho")
    return (idst if inverse else dst)(values, type=1 if kind == "dst1" else 2, axis=axis, norm="ortho")


@dataclass(frozen=True)
class Basis:
    """Tensor eigenbasis of one block: kinds along y and x plus the eigen-sum grid
    lambda_y[:, None] + lambda_x[None, :] (tensor_stokes.py:41-55)."""

    kind_y: str
    kind_x: str
    lam_y: np.ndarray
    lam_x: np.ndarray

    @property
    def eigenvalues(self) -> np.ndarray:
        return self.lam_y[:, None] + self.lam_x[None, :]


def make_basis(shape: tuple[int, int], spacings: tuple[float, float], kinds: tuple[str, str]) -> Basis:
    (ny, nx), (hy, hx), (ky, kx) = shape, spacings, kinds
    return Basis(ky, kx, eigenvalues_1d(ny, hy, ky), eigenvalues_1d(nx, hx, kx))


class _Transformer:
    """2-D forward/inverse transform of (n_y, n_x, k) stacks: scipy.fft (as the
    reference, tensor_stokes.py:85-101) or dense orthonormal matrices (the GEMM form
    the device path uses)."""

    def __init__(self, mode: str) -> None:
        if mode not in ("fft", "dense"):
            raise ValueError(mode)
        self.mode = mode
        self._cache: dict[tuple[int, str], np.ndarray] = {}

    def _q(self, n: int, kind: str) -> np.ndarray:
        key = (n, kind)
        if key not in self._cache:
            self._cache[key] = transform_matrix(n, kind)
        return self._cache[key]

    def __call__(self, values: np.ndarray, basis: Basis, inverse: bool) -> np.ndarray:
        if self.mode == "fft":
            return _fft_axis(_fft_axis(values, basis.kind_y, 0, inverse), basis.kind_x, 1, inverse)
        qy = self._q(values.shape[0], basis.kind_y)
        qx = self._q(values.shape[1], basis.kind_x)
        if inverse:
            t = np.einsum("ab,ajk->bjk", qy, values)       # Q_y^T Y
            return np.einsum("bjk,jn->bnk", t, qx)         # (.) Q_x
        t = np.einsum("ab,bjk->ajk", qy, values)           # Q_y X
        return np.einsum("ajk,nj->ank", t, qx)             # (.) Q_x^T


def column_means(values: np.ndarray) -> np.ndarray:
    """Per-column arithmetic mean, tensor_stokes.py:114-115."""
    return np.asarray([np.mean(values[:, c]) for c in range(values.shape[1])])


# ------------------------------------------------------------------------ PCG
def pcg(matvec, psolve, b: np.ndarray, rtol: float, maxiter: int, x0: np.ndarray | None = None,
        callback=None) -> tuple[np.ndarray, int]:
    """Preconditioned CG, the loop of scipy 1.15.3 `scipy.sparse.linalg.cg` with atol=0
    (the call at tensor_stokes.py:313-321): stop when ||r||_2 < rtol*||b||_2, tested at
    the top of each iteration before the preconditioner; info = maxiter when the loop
    is exhausted.  p <- beta*p then p += z; x += alpha*p; r -= alpha*q."""
    bnrm2 = np.linalg.norm(b)
    atol = max(0.0, rtol * bnrm2)
    if bnrm2 == 0:
        return b.copy(), 0
    x = np.zeros_like(b) if x0 is None else np.array(x0, dtype=float)
    r = b - matvec(x) if x.any() else b.copy()
    rho_prev = 0.0
    p = None
    for iteration in range(maxiter):
        if np.linalg.norm(r) < atol:
            return x, 0
        z = psolve(r)
        rho = np.dot(r, z)
        if iteration > 0:
            p *= rho / rho_prev
            p += z
        else:
            p = z.copy()
        q = matvec(p)
        alpha = rho / np.dot(p, q)
        x += alpha * p
        r -= alpha * q
        rho_prev = rho
        if callback is not None:
            callback(x)
    return x, maxiter


@dataclass
class Result:
    """velocity (size[,k]), pressure (ny,nx[,k]), residual, divergence_inf (stokes.py:32-37)."""

    velocity: np.ndarray
    pressure: np.ndarray
    residual: float
    divergence_inf: float


def _check_inputs(rhs: np.ndarray, size: int, mass: float, viscosity: float) -> None:
    if not np.isfinite([mass, viscosity]).all() or min(mass, viscosity) < 0 or mass + viscosity <= 0:
        raise ValueError("Require mass,viscosity >= 0 and nonzero total")
    if rhs.ndim not in (1, 2) or rhs.shape[0] != size or not np.isfinite(rhs).all():
        raise ValueError("Invalid Stokes RHS")


# --------------------------------------------------------------- TensorStokes
class TensorStokes:
    """CPU restatement of the reference tensor backend, tensor_stokes.py:118-364.

    Velocity blocks are inverted exactly in their DST eigenbases (u: DST-II in y x DST-I
    in x; v: DST-I x DST-II), the pressure Schur complement D L^-1 D^T is solved by `pcg`
    with preconditioner mass*(D D^T)^+ + viscosity*I (DCT-II eigenbasis, zero mode
    masked), columns one after another.  `transforms="dense"` applies the same maps as
    orthonormal-matrix products.  `pressure_guess` (np, k) is an optional CG start (the
    reference always starts from zero)."""

    def __init__(self, grid: Grid, tolerance: float = 1e-9, krylov_tolerance: float = 1e-12,
                 max_iterations: int = 100, transforms: str = "fft") -> None:
        if tolerance <= 0 or krylov_tolerance <= 0 or max_iterations < 1:
            raise ValueError("Invalid solver controls")
        start = perf_counter()
        self.grid = grid
        self.ops = Operators(grid)
        self.tolerance = tolerance
        self.krylov_tolerance = krylov_tolerance
        self.max_iterations = max_iterations
        self._t = _Transformer(transforms)
        g = grid
        self.u_basis = make_basis((g.ny, g.nx - 1), (g.hy, g.hx), ("dst2", "dst1"))
        self.v_basis = make_basis((g.ny - 1, g.nx), (g.hy, g.hx), ("dst1", "dst2"))
        self.p_basis = make_basis((g.ny, g.nx), (g.hy, g.hx), ("dct2", "dct2"))
        lam = self.p_basis.eigenvalues
        self.p_positive = lam > 100.0 * np.finfo(float).eps * float(np.max(lam))
        self.setup_seconds = perf_counter() - start
        self.solve_seconds = 0.0
        self.solves = 0
        self.right_hand_sides = 0
        self.total_iterations = 0
        self.last_iterations: tuple[int, ...] = ()

    # -- a5/a8: velocity Helmholtz inverse (tensor_stokes.py:104-111, 200-215)
    def velocity_inverse(self, values: np.ndarray, mass: float, viscosity: float) -> np.ndarray:
        g = self.grid
        k = values.shape[1]
        out = []
        for basis, block, shape in ((self.u_basis, values[: g.nu], (g.ny, g.nx - 1)),
                                    (self.v_basis, values[g.nu:], (g.ny - 1, g.nx))):
            shift = mass + viscosity * basis.eigenvalues
            hat = self._t(block.reshape(*shape, k), basis, inverse=False)
            hat = hat / shift[:, :, None]
            out.append(self._t(hat, basis, inverse=True).reshape(-1, k))
        return np.vstack(out)

    # -- a9: (D D^T)^+ on the mean-zero space (tensor_stokes.py:217-243)
    def pressure_poisson_inverse(self, values: np.ndarray) -> np.ndarray:
        g = self.grid
        cols = values.reshape(g.np, -1)
        cols = cols - column_means(cols)
        hat = self._t(cols.reshape(g.ny, g.nx, -1), self.p_basis, inverse=False)
        lam = self.p_basis.eigenvalues[:, :, None]
        hat = np.divide(hat, lam, out=np.zeros_like(hat), where=self.p_positive[:, :, None])
        res = self._t(hat, self.p_basis, inverse=True).reshape(g.np, -1)
        return res - column_means(res)

    def schur(self, pressure: np.ndarray, mass: float, viscosity: float) -> np.ndarray:
        """S p = D L^-1 D^T (p - mean), then mean removed (tensor_stokes.py:272-279)."""
        mz = pressure - column_means(pressure)
        res = packed_divergence(self.velocity_inverse(packed_divergence_transpose(mz, self.grid), mass, viscosity), self.grid)
        return res - column_means(res)

    def precondition(self, pressure: np.ndarray, mass: float, viscosity: float) -> np.ndarray:
        """mass*(D D^T)^+ r + viscosity*r on the mean-zero space (tensor_stokes.py:281-287)."""
        mz = pressure - column_means(pressure)
        res = mass * self.pressure_poisson_inverse(mz) + viscosity * mz
        return res - column_means(res)

    def solve(self, rhs: np.ndarray, *, mass: float, viscosity: float,
              pressure_guess: np.ndarray | None = None) -> Result:
        g = self.grid
        _check_inputs(rhs, g.size, mass, viscosity)
        start = perf_counter()
        two_d = rhs.ndim == 2
        force = np.asarray(rhs, dtype=float) if two_d else np.asarray(rhs, dtype=float)[:, None]
        k = force.shape[1]
        w = self.velocity_inverse(force, mass, viscosity)
        b_all = -packed_divergence(w, g)
        b_all -= b_all.mean(axis=0)
        pressures = np.empty_like(b_all)
        iterations = []
        for c in range(k):
            count = 0

            def tick(_):
                nonlocal count
                count += 1

            b = b_all[:, c]
            if np.max(np.abs(b)) == 0.0:
                x, info = np.zeros(g.np), 0
            else:
                x0 = None if pressure_guess is None else np.asarray(pressure_guess, dtype=float).reshape(g.np, -1)[:, c]
                x, info = pcg(lambda p: self.schur(p[:, None], mass, viscosity)[:, 0],
                              lambda r: self.precondition(r[:, None], mass, viscosity)[:, 0],
                              b, self.krylov_tolerance, self.max_iterations, x0=x0, callback=tick)
            if info != 0:
                raise RuntimeError(f"Pressure Schur CG failed with info={info}")
            pressures[:, c] = x - x.mean()
            iterations.append(count)
        pressures -= column_means(pressures)
        lift = packed_divergence_transpose(pressures, g)
        velocity = self.velocity_inverse(force + lift, mass, viscosity)
        defect = mass * velocity + viscosity * (self.ops.K @ velocity) - lift - force
        residual = float(np.max(np.abs(defect)) / (1 + np.max(np.abs(force))))
        divergence = float(np.max(np.abs(packed_divergence(velocity, g))))
        self.solve_seconds += perf_counter() - start
        self.solves += 1
        self.right_hand_sides += k
        self.total_iterations += sum(iterations)
        self.last_iterations = tuple(iterations)
        if (not np.isfinite(velocity).all() or not np.isfinite(pressures).all()
                or residual > self.tolerance
                or divergence > self.tolerance * (1 + np.max(np.abs(velocity)))):
            raise RuntimeError(f"Stokes residual={residual:.3e}, divergence={divergence:.3e}")
        if not two_d:
            return Result(velocity[:, 0], pressures[:, 0].reshape(g.ny, g.nx), residual, divergence)
        return Result(velocity, pressures.reshape(g.ny, g.nx, -1), residual, divergence)


# --------------------------------------------------------------- DirectStokes
class DirectStokes:
    """Gauged saddle system by sparse LU (stokes.py:42-86): one pressure gauge and one
    dependent continuity row removed, pressure shifted to mean zero afterwards."""

    def __init__(self, grid: Grid, tolerance: float = 1e-9) -> None:
        self.grid = grid
        self.ops = Operators(grid)
        self.tolerance = tolerance
        self._cache: dict[tuple[float, float], tuple] = {}

    def solve(self, rhs: np.ndarray, *, mass: float, viscosity: float) -> Result:
        g, ops = self.grid, self.ops
        _check_inputs(rhs, g.size, mass, viscosity)
        key = (float(mass), float(viscosity))
        if key not in self._cache:
            L = mass * sp.eye(g.size, format="csr") + viscosity * ops.K
            A = sp.bmat([[L, ops.G[:, :-1]], [-ops.D[:-1, :], None]], format="csc")
            self._cache[key] = (splu(A), L)
        lu, L = self._cache[key]
        two_d = rhs.ndim == 2
        force = rhs if two_d else rhs[:, None]
        sol = lu.solve(np.vstack([force, np.zeros((g.np - 1, force.shape[1]))]))
        velocity = sol[: g.size]
        pressure = np.vstack([sol[g.size:], np.zeros((1, force.shape[1]))])
        pressure -= pressure.mean(axis=0)
        defect = L @ velocity + ops.G @ pressure - force
        residual = float(np.max(np.abs(defect)) / (1 + np.max(np.abs(force))))
        divergence = float(np.max(np.abs(ops.D @ velocity)))
        if (not np.isfinite(sol).all() or residual > self.tolerance
                or divergence > self.tolerance * (1 + np.max(np.abs(velocity)))):
            raise RuntimeError(f"Stokes residual={residual:.3e}, divergence={divergence:.3e}")
        if not two_d:
            return Result(velocity[:, 0], pressure[:, 0].reshape(g.ny, g.nx), residual, divergence)
        return Result(velocity, pressure.reshape(g.ny, g.nx, -1), residual, divergence)


def dense_saddle_solve(grid: Grid, rhs: np.ndarray, mass: float, viscosity: float) -> tuple[np.ndarray, np.ndarray]:
    """Brute force: the full gauged saddle system as a dense matrix, `np.linalg.solve`.
    Independent of both backends' algorithms (only the operators are shared)."""
    ops = Operators(grid)
    L = (mass * np.eye(grid.size) + viscosity * ops.K.toarray())
    G = ops.G.toarray()
    D = ops.D.toarray()
    # Mean-zero pressure enforced by a Lagrange row instead of a dropped gauge.
    n, m = grid.size, grid.np
    A = np.zeros((n + m + 1, n + m + 1))
    A[:n, :n] = L
    A[:n, n:n + m] = G
    A[n:n + m, :n] = D
    A[n:n + m, n + m] = 1.0
    A[n + m, n:n + m] = 1.0
    f = rhs if rhs.ndim == 2 else rhs[:, None]
    b = np.vstack([f, np.zeros((m + 1, f.shape[1]))])
    sol = np.linalg.solve(A, b)
    return sol[:n], sol[n:n + m]
