"""Learned warm start for the pressure Schur CG: the network, its training, its data.

`WarmStartNet` wraps the device network of libmacstokes.so (forward, backward and Adam
written as gfx950 kernels).  For a stage column it reads the residual of a base guess
of the pressure, db = b - b_base (b the zero start's Schur right-hand side, b_base the
right-hand sides recorded with the base's pressures, combined like them), and the base
itself, both average-pooled by 4 and each normalised by its own rms.  It outputs the
correction pool4(p - base) / (alpha rms(pool4 db)); the solver adds the bilinearly
upsampled output to the base.  With one operator the correction solves
S (p - base) = db, so its scale follows db's.  The base is the same stage of the last
three steps extrapolated in time when the solver has a history period, else the
previous solve's pressure.  The solve still runs to the same tolerance and is the
judge: a poor guess costs iterations, never accuracy.

Training pairs come from the plain FP64 solver (`collect_pairs`): for every stage
column of a run that has a base, (db, base guess, pressure).  No counterpart in the
reference.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Any

import numpy as np

from . import _native


class WarmStartNet:
    def __init__(self, width: int = 32, seed: int = 0, device: int = 0, alpha: float = 1.0) -> None:
        lib = _native.load()
        handle = lib.mac_cnn_create(width, device)
        if not handle:
            raise RuntimeError(f"network setup failed: {lib.mac_cnn_create_error().decode()}")
        self._lib, self.handle, self.width, self.device = lib, handle, width, device
        self.alpha = alpha
        self.n_params = int(lib.mac_cnn_num_params(handle))
        self.set_params(self.initial_params(width, seed))

    def __del__(self) -> None:
        if getattr(self, "handle", None):
            self._lib.mac_cnn_destroy(self.handle)
            self.handle = None

    @staticmethod
    def initial_params(width: int, seed: int) -> np.ndarray:
        """He-normal weights, zero biases, flattened per layer (weights then bias)."""
        rng = np.random.default_rng(seed)
        shapes = [(width, 2)] + [(width, width)] * 5 + [(1, width)]
        parts = []
        for co, ci in shapes:
            parts.append(rng.normal(0.0, np.sqrt(2.0 / (9 * ci)), size=co * ci * 9))
            parts.append(np.zeros(co))
        return np.concatenate(parts).astype(np.float32)

    def _check(self, code: int) -> None:
        _native.raise_for(code, self._lib.mac_cnn_last_error(self.handle).decode())

    def set_params(self, flat: np.ndarray) -> None:
        flat = np.ascontiguousarray(flat, dtype=np.float32)
        if flat.size != self.n_params:
            raise ValueError(f"expected {self.n_params} parameters, got {flat.size}")
        self._check(self._lib.mac_cnn_set_params(self.handle, flat.ctypes.data))

    def get_params(self) -> np.ndarray:
        out = np.empty(self.n_params, dtype=np.float32)
        self._check(self._lib.mac_cnn_get_params(self.handle, out.ctypes.data))
        return out

    def forward(self, x: Any) -> Any:
        """x: (n, 2, H, W) float32 device tensor -> (n, 1, H, W)."""
        import torch
        x = x.contiguous()
        n, _, h, w = x.shape
        y = torch.empty((n, 1, h, w), dtype=torch.float32, device=x.device)
        torch.cuda.current_stream(x.device).synchronize()
        self._check(self._lib.mac_cnn_forward(self.handle, x.data_ptr(), y.data_ptr(), n, h, w))
        return y

    def loss_grad(self, x: Any, target: Any) -> tuple[float, Any]:
        import torch
        x, target = x.contiguous(), target.contiguous()
        n, _, h, w = x.shape
        grad = torch.empty(self.n_params, dtype=torch.float32, device=x.device)
        loss = ctypes.c_float()
        torch.cuda.current_stream(x.device).synchronize()
        self._check(self._lib.mac_cnn_loss_grad(self.handle, x.data_ptr(), target.data_ptr(), n, h, w,
                                                grad.data_ptr(), ctypes.byref(loss)))
        return float(loss.value), grad

    def train_step(self, x: Any, target: Any, lr: float) -> float:
        import torch
        x, target = x.contiguous(), target.contiguous()
        n, _, h, w = x.shape
        loss = ctypes.c_float()
        torch.cuda.current_stream(x.device).synchronize()
        self._check(self._lib.mac_cnn_train_step(self.handle, x.data_ptr(), target.data_ptr(), n, h, w, lr,
                                                 ctypes.byref(loss)))
        return float(loss.value)


def prepare(grid, b: Any, pprev: Any | None, target: Any | None, alpha: float) -> tuple[Any, Any, Any]:
    """Network inputs (n, 2, ny/4, nx/4), targets (n, 1, ny/4, nx/4) and the b scales
    (n,) for columns of b / base / target, each (np, n) device tensors, exactly as the
    solver prepares them at inference (the base plane is normalised by its own rms)."""
    import torch
    lib = _native.load()
    dev = b.device
    n = b.shape[1]
    bc = b.t().contiguous()
    pc = None if pprev is None else pprev.t().contiguous()
    tc = None if target is None else target.t().contiguous()
    h, w = grid.ny // 4, grid.nx // 4
    x = torch.empty((n, 2, h, w), dtype=torch.float32, device=dev)
    t = torch.empty((n, 1, h, w), dtype=torch.float32, device=dev) if target is not None else None
    s = torch.empty(2 * n, dtype=torch.float64, device=dev)
    torch.cuda.current_stream(dev).synchronize()
    code = lib.mac_warm_prepare(grid.nx, grid.ny, bc.data_ptr(), grid.np, None if pc is None else pc.data_ptr(),
                                grid.np, None if tc is None else tc.data_ptr(), grid.np, n, alpha, x.data_ptr(),
                                None if t is None else t.data_ptr(), s.data_ptr())
    _native.raise_for(code, "mac_warm_prepare failed")
    return x, t, s[:n]


@dataclass
class Pairs:
    """Stage columns recorded from plain FP64 solves: the base's residual b - b_base (in
    `b`), base guess, pressure."""
    b: list
    base: list
    p: list


def history_weights(steps: int) -> list[float]:
    """Equal-step extrapolation weights through the last `steps` values:
    (-1)^(k+1) C(steps, k), k = 1..steps (2, -1; 3, -3, 1; ...)."""
    from math import comb
    return [(-1.0) ** (k + 1) * comb(steps, k) for k in range(1, steps + 1)]


def base_guess(history: list, period: int | None, depth: int = 3) -> Any:
    """The solver's base guess from the recorded pressures (oldest first) of solves with
    the same column count: the same stage of the last min(depth, recorded) steps
    extrapolated with `history_weights`, zero before (P = period), or the previous
    pressure when there is no period."""
    if period:
        steps = min(len(history) // period, depth)
        if steps == 0:
            return None
        w = history_weights(steps)
        out = w[0] * history[-period]
        for k in range(1, steps):
            out = out + w[k] * history[-(k + 1) * period]
        return out
    return history[-1] if history else None


def collect_pairs(model, scheme, initial, steps: int, dt: float, period: int | None = None,
                  depth: int = 3) -> Pairs:
    """Run `scheme` on `model` (whose backend must be an FP64 GpuTensorStokes) and record,
    for every stage solve that has a base guess, (b - b_base, base guess, pressure) per
    column as device tensors of shape (np, ncols); `period` as in
    `GpuTensorStokes.set_warm_start`."""
    import torch
    backend = model.backend
    backend.reset_history()
    pairs = Pairs([], [], [])
    original = backend.solve
    history: dict[int, list] = {}
    rhs_history: dict[int, list] = {}

    def recording(rhs, *, mass, viscosity, pressure_guess=None):
        res = original(rhs, mass=mass, viscosity=viscosity, pressure_guess=pressure_guess)
        k = 1 if len(rhs.shape) == 1 else int(rhs.shape[1])
        b = backend.last_schur_rhs(k)
        p = torch.as_tensor(res.pressure, device=b.device).reshape(model.grid.np, k)
        past, past_b = history.setdefault(k, []), rhs_history.setdefault(k, [])
        base = base_guess(past, period, depth)
        if base is not None:
            # Solves without a base guess are not learned: the solver then starts from zero.
            pairs.b.append(b - base_guess(past_b, period, depth))
            pairs.base.append(base)
            pairs.p.append(p.clone())
        past.append(p.clone())
        past_b.append(b.clone())
        del past[:-depth * max(period or 1, 1)]
        del past_b[:-depth * max(period or 1, 1)]
        return res

    backend.solve = recording
    try:
        state = initial
        for _ in range(steps):
            state = scheme.step(model, state, dt).state
    finally:
        backend.solve = original
    return pairs


def fit_alpha(grid, pairs: Pairs) -> float:
    """alpha = median over columns of rms(pool4 (p - base)) / rms(pool4 (b - b_base))."""
    ratios = []
    for b, pp, p in zip(pairs.b, pairs.base, pairs.p):
        xb, _, sb = prepare(grid, b, None, None, 1.0)
        xp, _, sp = prepare(grid, p - pp, None, None, 1.0)
        ratios.extend((sp / sb).cpu().numpy().tolist())
    return float(np.median(ratios))


def train(net: WarmStartNet, grid, pairs: Pairs, epochs: int = 40, lr: float = 1e-3, batch: int = 8,
          seed: int = 0) -> list[float]:
    """Adam on the mean squared error of the normalised pooled pressure; returns the
    per-epoch mean loss."""
    import torch
    xs, ts = [], []
    for b, pp, p in zip(pairs.b, pairs.base, pairs.p):
        x, t, _ = prepare(grid, b, pp, p - pp, net.alpha)
        xs.append(x)
        ts.append(t)
    X, T = torch.cat(xs), torch.cat(ts)
    rng = np.random.default_rng(seed)
    history = []
    for _ in range(epochs):
        order = rng.permutation(X.shape[0])
        losses = []
        for k in range(0, len(order), batch):
            idx = torch.as_tensor(order[k:k + batch], device=X.device)
            losses.append(net.train_step(X[idx], T[idx], lr))
        history.append(float(np.mean(losses)))
    return history
