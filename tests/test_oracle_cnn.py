"""Pins the warm-start network oracle (oracle/cnn_oracle.py): finite differences of
its own loss for every parameter tensor, and closed forms of its building blocks."""
import numpy as np
import pytest

from oracle import cnn_oracle as co


def _small(seed=0, width=4):
    rng = np.random.default_rng(seed)
    params = [(W.astype(np.float64), b.astype(np.float64) + 0.05 * rng.normal(size=b.shape))
              for W, b in co.init_params(seed, width=width)]
    x = rng.normal(size=(2, 16, 16))
    t = rng.normal(size=(1, 16, 16))
    return params, x, t


def test_gradients_match_central_differences_for_every_tensor():
    params, x, t = _small()
    _, grads = co.loss_and_grads(params, x, t, round_weights=False)
    rng = np.random.default_rng(3)
    for layer in range(7):
        for which in (0, 1):
            arr = params[layer][which]
            for _ in range(3):
                idx = tuple(rng.integers(0, s) for s in arr.shape)
                h = 1e-6
                old = arr[idx]
                arr[idx] = old + h
                lp = co.loss_and_grads(params, x, t, round_weights=False)[0]
                arr[idx] = old - h
                lm = co.loss_and_grads(params, x, t, round_weights=False)[0]
                arr[idx] = old
                fd = (lp - lm) / (2 * h)
                an = grads[layer][which][idx]
                assert abs(fd - an) <= 1e-6 * max(1.0, abs(fd)), (layer, which, idx, fd, an)


def test_pool_and_upsample_are_adjoint_up_to_the_quarter():
    rng = np.random.default_rng(1)
    a, b = rng.normal(size=(3, 8, 8)), rng.normal(size=(3, 4, 4))
    assert abs(np.sum(co.pool(a) * b) - 0.25 * np.sum(a * co.up(b))) < 1e-12
    np.testing.assert_allclose(co.up_backward(co.up(b)), 4 * b)


def test_centre_tap_conv_is_a_channel_mix_and_zero_net_gives_bias_only():
    rng = np.random.default_rng(2)
    x = rng.normal(size=(3, 5, 6))
    W = np.zeros((2, 3, 3, 3))
    M = rng.normal(size=(2, 3))
    W[:, :, 1, 1] = M
    np.testing.assert_allclose(co.conv(x, W, np.zeros(2)), np.einsum("oc,chw->ohw", M, x))
    params = [(np.zeros_like(W_), np.zeros_like(b_)) for W_, b_ in co.init_params(0, width=4)]
    params[6] = (params[6][0], np.array([0.5]))
    np.testing.assert_allclose(co.forward(params, rng.normal(size=(2, 16, 16))), 0.5)


def test_bf16_rounding_is_round_to_nearest_even():
    vals = np.array([1.0, 1.0 + 2**-8, 1.0 + 3 * 2**-9, -2.5, 3.1415926], dtype=np.float32)
    r = co.bf16_round(vals)
    assert r[0] == 1.0 and r[1] == 1.0          # tie -> even (mantissa ...0)
    assert r[2] == np.float32(1.0 + 2**-7)      # 1.5 ulp -> rounds up
    assert r[3] == -2.5 and abs(r[4] - 3.140625) < 1e-7


def test_flatten_roundtrip():
    params = co.init_params(5, width=8)
    back = co.unflatten(co.flatten(params), width=8)
    for (a, b), (c, d) in zip(params, back):
        np.testing.assert_array_equal(a, c)
        np.testing.assert_array_equal(b, d)
