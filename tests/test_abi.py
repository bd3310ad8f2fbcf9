"""The C-ABI library loads and exports exactly what include/macstokes.h declares.

Runs without a GPU: only symbol lookup and host-side argument validation (which
returns before any HIP call) are exercised.
"""
import ctypes
import re
from pathlib import Path

import pytest

from macsav import _native

HEADER = Path(__file__).resolve().parents[1] / "include" / "macstokes.h"


def declared_symbols():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(mac_[A-Za-z_0-9]+)\s*\(", text)))


def test_header_and_binding_list_the_same_entry_points():
    assert declared_symbols() == sorted(_native.EXPORTED)


def test_library_exports_every_declared_symbol():
    lib = _native.load()
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_default_opts_are_the_reference_defaults():
    opts = _native.default_opts()
    assert opts.tolerance == 1e-9 and opts.krylov_tolerance == 1e-12 and opts.max_iterations == 100
    assert opts.precision == _native.MAC_PRECISION_FP64


@pytest.mark.parametrize("nx,ny,lx,field,value", [
    (2, 5, 1.0, None, None), (5, 5, -1.0, None, None),
    (5, 5, 1.0, "tolerance", 0.0), (5, 5, 1.0, "max_iterations", 0), (5, 5, 1.0, "precision", 7),
])
def test_create_rejects_invalid_grids_and_controls_before_touching_the_gpu(nx, ny, lx, field, value):
    lib = _native.load()
    opts = _native.default_opts()
    if field:
        setattr(opts, field, value)
    assert not lib.mac_stokes_create(nx, ny, lx, 1.0, ctypes.byref(opts))
    msg = lib.mac_stokes_create_error().decode()
    assert msg.startswith("Require") or msg.startswith("Invalid solver controls")


def test_null_handle_is_an_invalid_argument():
    lib = _native.load()
    assert lib.mac_stokes_last_error(None) == b""
    assert lib.mac_apply_divergence(None, None, 0, 1, None, 0) == _native.MAC_INVALID_ARGUMENT


MIXED = _native.MAC_PRECISION_MIXED


@pytest.mark.parametrize("count,nx,ncols,fields", [
    (0, 8, 2, {}), (2, 2, 2, {}), (2, 8, 0, {}), (2, 8, 9, {}), (2, 8, 2, {"tolerance": -1.0}),
    (2, 8, 2, {"precision": 7}), (2, 8, 2, {"precision": MIXED, "inner_tolerance": 1.5}),
    (2, 8, 2, {"precision": MIXED, "refine_tolerance": 0.0}), (2, 8, 2, {"precision": MIXED, "max_sweeps": 0}),
])
def test_batch_create_rejects_invalid_cavities_and_controls_before_touching_the_gpu(count, nx, ncols, fields):
    lib = _native.load()
    opts = _native.default_opts()
    for field, value in fields.items():
        setattr(opts, field, value)
    n = max(count, 1)
    arr_i = (ctypes.c_int32 * n)(*([nx] * n))
    arr_d = (ctypes.c_double * n)(*([1.0] * n))
    assert not lib.mac_batch_create(count, arr_i, arr_i, arr_d, arr_d, ncols, ctypes.byref(opts))
    msg = lib.mac_batch_create_error().decode()
    assert msg.startswith("Require") or msg.startswith("Invalid solver controls")


def test_batch_null_handle_is_an_invalid_argument():
    lib = _native.load()
    assert lib.mac_batch_last_error(None) == b"null handle"
    assert lib.mac_batch_last_sweeps(None) == -1
    assert lib.mac_batch_set_history(None, 2, 3) == _native.MAC_INVALID_ARGUMENT
    assert lib.mac_batch_history_mark(None, 0) == _native.MAC_INVALID_ARGUMENT
    assert lib.mac_stokes_solve_batched(None, None, None, None, None, None, None, None, None, 0) == \
        _native.MAC_INVALID_ARGUMENT
    assert lib.mac_batch_set_launch(None, 1, 1) == _native.MAC_INVALID_ARGUMENT
