"""The allocation-free inner product against mac_dot, through the C ABI: bitwise equal
results on two calls in a row, the second on a smaller input (the shared file-scope
scratch keeps the first call's partials), and the size limit rejected as an error."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def test_dot_noalloc_equals_dot_on_consecutive_calls():
    from macsav import _native
    lib = _native.load()
    lib.mac_dot_noalloc.argtypes = [ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.POINTER(ctypes.c_double)]
    lib.mac_dot_noalloc.restype = ctypes.c_int32
    lib.mac_noalloc_last_error.restype = ctypes.c_char_p
    gen = torch.Generator(device="cuda:0").manual_seed(3)
    for n in (4_194_311, 70_001, 5):
        a = torch.randn(n, dtype=torch.float64, device="cuda:0", generator=gen)
        b = torch.randn(n, dtype=torch.float64, device="cuda:0", generator=gen)
        torch.cuda.synchronize()
        want, got = ctypes.c_double(), ctypes.c_double()
        assert lib.mac_dot(n, a.data_ptr(), b.data_ptr(), ctypes.byref(want)) == 0
        assert lib.mac_dot_noalloc(n, a.data_ptr(), b.data_ptr(), ctypes.byref(got)) == 0
        print(f"n={n}: mac_dot {want.value!r}, mac_dot_noalloc {got.value!r}", flush=True)
        assert got.value == want.value
        assert got.value == pytest.approx(float((a * b).sum()), rel=1e-12, abs=1e-12)
    bad = ctypes.c_double()
    assert lib.mac_dot_noalloc(0, a.data_ptr(), b.data_ptr(), ctypes.byref(bad)) == _native.MAC_INVALID_ARGUMENT
    assert lib.mac_dot_noalloc(1 << 31, a.data_ptr(), b.data_ptr(), ctypes.byref(bad)) == _native.MAC_INVALID_ARGUMENT
    assert b"2^31" in lib.mac_noalloc_last_error()
