"""NumPy forward/backward of the pressure warm-start network -- TEST INFRASTRUCTURE ONLY.

No counterpart in the reference (the learned warm start is new in this build); this
module is the CPU oracle that the HIP training and inference kernels are checked
against.  It is pinned by finite differences of its own loss (every parameter tensor)
and by closed forms: a zero network gives zero output, avg-pool / nearest-upsample are
adjoint, and a 3x3 conv with a centre-only kernel is a per-pixel channel mix.

Network (4 levels, `width` channels, 3x3 convs with zero padding, ReLU after every conv
but the last):
    e0 = relu(conv0(x))            x: (2, H, W)  -> (w, H, W)
    e1 = relu(conv1(pool(e0)))     (w, H/2, W/2)
    e2 = relu(conv2(pool(e1)))     (w, H/4, W/4)
    e3 = relu(conv3(pool(e2)))     (w, H/8, W/8)
    g2 = relu(conv4(up(e3) + e2))
    g1 = relu(conv5(up(g2) + e1))
    y  = conv6(up(g1) + e0)        (1, H, W)
pool = 2x2 average, up = 2x nearest.  Parameters are a list of (weight (co, ci, 3, 3),
bias (co,)) pairs in the order conv0..conv6.  The forward pass rounds weights to bf16
(round to nearest even), as the device kernels store them; accumulation is float64 here.
"""
from __future__ import annotations

import numpy as np

LAYERS = ("c0", "c1", "c2", "c3", "c4", "c5", "c6")


def layer_shapes(width: int = 32, cin: int = 2) -> list[tuple[int, int]]:
    """(co, ci) per conv."""
    return [(width, cin), (width, width), (width, width), (width, width), (width, width), (width, width),
            (1, width)]


def init_params(seed: int, width: int = 32, cin: int = 2) -> list[tuple[np.ndarray, np.ndarray]]:
    rng = np.random.default_rng(seed)
    params = []
    for co, ci in layer_shapes(width, cin):
        std = np.sqrt(2.0 / (9 * ci))
        params.append((rng.normal(0.0, std, size=(co, ci, 3, 3)).astype(np.float32),
                       np.zeros(co, dtype=np.float32)))
    return params


def bf16_round(a: np.ndarray) -> np.ndarray:
    """float32 -> nearest bf16 (ties to even) -> float32."""
    u = np.asarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    rounded = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return rounded.astype(np.uint32).view(np.float32)


def _im2col(x: np.ndarray) -> np.ndarray:
    c, h, w = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1)))
    cols = np.empty((c, 9, h, w), dtype=x.dtype)
    for dy in range(3):
        for dx in range(3):
            cols[:, dy * 3 + dx] = xp[:, dy:dy + h, dx:dx + w]
    return cols.reshape(c * 9, h * w)


def _col2im(cols: np.ndarray, c: int, h: int, w: int) -> np.ndarray:
    cols = cols.reshape(c, 9, h, w)
    xp = np.zeros((c, h + 2, w + 2), dtype=cols.dtype)
    for dy in range(3):
        for dx in range(3):
            xp[:, dy:dy + h, dx:dx + w] += cols[:, dy * 3 + dx]
    return xp[:, 1:-1, 1:-1]


def conv(x, weight, bias):
    co = weight.shape[0]
    h, w = x.shape[1:]
    return (weight.reshape(co, -1) @ _im2col(x) + bias[:, None]).reshape(co, h, w)


def conv_backward(x, weight, dout):
    co, ci = weight.shape[:2]
    h, w = x.shape[1:]
    d = dout.reshape(co, h * w)
    dw = (d @ _im2col(x).T).reshape(weight.shape)
    db = d.sum(axis=1)
    dx = _col2im(weight.reshape(co, -1).T @ d, ci, h, w)
    return dx, dw, db


def pool(x):
    return 0.25 * (x[:, ::2, ::2] + x[:, 1::2, ::2] + x[:, ::2, 1::2] + x[:, 1::2, 1::2])


def pool_backward(d):
    return 0.25 * up(d)


def up(x):
    return x.repeat(2, axis=1).repeat(2, axis=2)


def up_backward(d):
    return d[:, ::2, ::2] + d[:, 1::2, ::2] + d[:, ::2, 1::2] + d[:, 1::2, 1::2]


def forward(params, x, keep: bool = False, round_weights: bool = True):
    """x: (2, H, W) float64 -> y (1, H, W); with keep, also the activations.  The weight
    gradient of the rounded weight is used for the master weight (straight-through)."""
    cast = bf16_round if round_weights else (lambda a: a)
    w = [(cast(W).astype(np.float64), b.astype(np.float64)) for W, b in params]
    relu = lambda a: np.maximum(a, 0.0)  # noqa: E731
    e0 = relu(conv(x, *w[0]))
    p1 = pool(e0)
    e1 = relu(conv(p1, *w[1]))
    p2 = pool(e1)
    e2 = relu(conv(p2, *w[2]))
    p3 = pool(e2)
    e3 = relu(conv(p3, *w[3]))
    u2 = up(e3) + e2
    g2 = relu(conv(u2, *w[4]))
    u1 = up(g2) + e1
    g1 = relu(conv(u1, *w[5]))
    u0 = up(g1) + e0
    y = conv(u0, *w[6])
    if not keep:
        return y
    return y, dict(x=x, e0=e0, p1=p1, e1=e1, p2=p2, e2=e2, p3=p3, e3=e3, u2=u2, g2=g2, u1=u1, g1=g1, u0=u0, w=w)


def loss_and_grads(params, x, target, round_weights: bool = True):
    """Mean squared error over the H*W outputs and its gradient per parameter."""
    y, a = forward(params, x, keep=True, round_weights=round_weights)
    n = y.size
    diff = y - target
    loss = float(np.sum(diff * diff) / n)
    dy = 2.0 * diff / n
    w = a["w"]
    grads = [None] * 7
    du0, dW, db = conv_backward(a["u0"], w[6][0], dy)
    grads[6] = (dW, db)
    dg1 = up_backward(du0) * (a["g1"] > 0)
    de0 = du0.copy()
    du1, dW, db = conv_backward(a["u1"], w[5][0], dg1)
    grads[5] = (dW, db)
    dg2 = up_backward(du1) * (a["g2"] > 0)
    de1 = du1.copy()
    du2, dW, db = conv_backward(a["u2"], w[4][0], dg2)
    grads[4] = (dW, db)
    de3 = up_backward(du2) * (a["e3"] > 0)
    de2 = du2.copy()
    dp3, dW, db = conv_backward(a["p3"], w[3][0], de3)
    grads[3] = (dW, db)
    de2 = (de2 + pool_backward(dp3)) * (a["e2"] > 0)
    dp2, dW, db = conv_backward(a["p2"], w[2][0], de2)
    grads[2] = (dW, db)
    de1 = (de1 + pool_backward(dp2)) * (a["e1"] > 0)
    dp1, dW, db = conv_backward(a["p1"], w[1][0], de1)
    grads[1] = (dW, db)
    de0 = (de0 + pool_backward(dp1)) * (a["e0"] > 0)
    _, dW, db = conv_backward(a["x"], w[0][0], de0)
    grads[0] = (dW, db)
    return loss, grads


def flatten(params) -> np.ndarray:
    """Parameters in device order: per layer the weights (co, ci, 3, 3) then the bias."""
    return np.concatenate([np.concatenate([W.ravel(), b.ravel()]) for W, b in params]).astype(np.float32)


def unflatten(flat: np.ndarray, width: int = 32, cin: int = 2):
    out, k = [], 0
    for co, ci in layer_shapes(width, cin):
        nw = co * ci * 9
        out.append((flat[k:k + nw].reshape(co, ci, 3, 3).copy(), flat[k + nw:k + nw + co].copy()))
        k += nw + co
    return out


def adam_step(w, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba, 2015, Algorithm 1) in float64: the moments, their
    bias corrections 1 - b^step, and w - lr m_hat / (sqrt(v_hat) + eps).  Returns the new
    (w, m, v)."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    m_hat = m / (1 - b1 ** step)
    v_hat = v / (1 - b2 ** step)
    return w - lr * m_hat / (np.sqrt(v_hat) + eps), m, v
