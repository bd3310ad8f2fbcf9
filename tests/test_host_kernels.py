"""Product host-side convection against the oracle's explicit loop (bitwise)."""
import numpy as np

from oracle import mac_oracle as mo
from tests.test_oracle_mac import _loop_convection


def test_product_convection_is_bitwise_the_reference_loop():
    from macsav.mac.kernels import convection
    g = mo.Grid(13, 9, 2.0, 0.7)
    u, v = g.unpack(np.random.default_rng(7).normal(size=g.size))
    cu, cv = convection(u, v, g.hx, g.hy)
    lu, lv = _loop_convection(u, v, g.hx, g.hy)
    np.testing.assert_array_equal(cu, lu)
    np.testing.assert_array_equal(cv, lv)


def test_stream_function_field_is_discretely_divergence_free():
    from macsav import MACGrid
    from macsav.fields import stream_function_velocity
    grid = MACGrid(40, 32, 1.3, 0.8)
    x = stream_function_velocity(grid, seed=3, modes=8)
    g = mo.Grid(40, 32, 1.3, 0.8)
    assert np.abs(mo.packed_divergence(x, g)).max() < 1e-11
    assert np.abs(x).max() == 1.0
    assert not np.array_equal(x, stream_function_velocity(grid, seed=4, modes=8))
