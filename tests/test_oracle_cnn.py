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


def test_adam_first_step_is_the_signed_learning_rate_and_moments_follow_the_recurrence():
    # Closed forms: with zero moments, one step gives m_hat = g and v_hat = g^2, so the
    # update is lr g / (|g| + eps) ~ lr sign(g); a second step with gradient g2 has
    # m = 0.9 (0.1 g) + 0.1 g2 and v = 0.999 (0.001 g^2) + 0.001 g2^2.
    g = np.array([0.3, -2.0, 1e-3])
    w0 = np.array([1.0, 2.0, -1.0])
    w1, m1, v1 = co.adam_step(w0, g, np.zeros(3), np.zeros(3), 1, lr=0.01)
    np.testing.assert_allclose(w1, w0 - 0.01 * g / (np.abs(g) + 1e-8), rtol=1e-12)
    g2 = np.array([-0.1, 0.5, 0.2])
    w2, m2, v2 = co.adam_step(w1, g2, m1, v1, 2, lr=0.01)
    np.testing.assert_allclose(m2, 0.9 * 0.1 * g + 0.1 * g2, rtol=1e-12)
    np.testing.assert_allclose(v2, 0.999 * 0.001 * g * g + 0.001 * g2 * g2, rtol=1e-12)
    mh, vh = m2 / (1 - 0.81), v2 / (1 - 0.999 ** 2)
    np.testing.assert_allclose(w2, w1 - 0.01 * mh / (np.sqrt(vh) + 1e-8), rtol=1e-12)


def test_network_buffer_plan_covers_every_pass_extent():
    """The host-side plan the network sizes its buffers by (mac_cnn_plan, no GPU call):
    each buffer holds the largest extent a pass writes or reads in it, for the shapes the
    solver and the training use.  The extents are derived here from the topology (4
    levels, 2x2 pooling, `width` channels; weight-gradient partials per 16x16 tile)."""
    import ctypes
    from macsav import _native
    lib = _native.load()
    for width, n, H, W in [(32, 2, 256, 256), (32, 16, 128, 128), (8, 3, 8, 8), (32, 1, 512, 256), (32, 8, 32, 64)]:
        counts = (ctypes.c_int64 * 29)()
        assert lib.mac_cnn_plan(width, n, H, W, counts) == 0
        c = list(counts)
        full = n * width * H * W
        level = [full, full // 4, full // 16, full // 64]
        # e0 p1 e1 p2 e2 p3 e3 u2 g2 u1 g1 u0 y dy du0 dg1 du1 dg2 du2 de3 dp3 de2 dp2 de1 dp1 de0
        want = [level[0], level[1], level[1], level[2], level[2], level[3], level[3], level[2], level[2], level[1],
                level[1], level[0], n * H * W, n * H * W, level[0], level[1], level[1], level[2], level[2], level[3],
                level[3], level[2], level[2], level[1], level[1], level[0]]
        assert c[:26] == want
        tiles = -(-H // 16) * -(-W // 16) * n
        layers = [(width, 2)] + [(width, width)] * 5 + [(1, width)]
        for l, (cout, cin) in enumerate(layers):
            scale = 2 ** min(l, 6 - l)                  # resolution of layer l: H/1, /2, /4, /8, /4, /2, /1
            tl = -(-(H // scale) // 16) * -(-(W // scale) // 16) * n
            assert tl * (cout * cin * 9 + cout) <= c[26]
        assert c[26] == tiles * (width * width * 9 + width)
        assert c[27] >= 256 and 2 * c[28] >= width * width * 9
    bad = (ctypes.c_int64 * 29)()
    assert lib.mac_cnn_plan(32, 1, 12, 16, bad) == _native.MAC_INVALID_ARGUMENT
