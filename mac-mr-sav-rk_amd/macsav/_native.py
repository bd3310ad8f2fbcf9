"""ctypes binding of libmacstokes.so (the gfx950 HIP library, C ABI in include/macstokes.h).

The library is loaded from this package directory.  There is no fallback: when the
shared object is missing, or no GPU is visible, the solver raises instead of
computing anything on the CPU.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libmacstokes.so"

MAC_OK = 0
MAC_INVALID_ARGUMENT = 1
MAC_NOT_CONVERGED = 2
MAC_GATE_FAILED = 3
MAC_HIP_ERROR = 4

MAC_WARM_OFF = 0
MAC_WARM_NETWORK = 1
MAC_WARM_PREVIOUS = 2
MAC_WARM_HISTORY = 3

MAC_PRECISION_FP64 = 0
MAC_PRECISION_MIXED = 1

# Every symbol include/macstokes.h declares, checked by the CPU test suite.
EXPORTED = (
    "mac_stokes_default_opts",
    "mac_stokes_create",
    "mac_stokes_create_error",
    "mac_stokes_destroy",
    "mac_stokes_last_error",
    "mac_stokes_solve",
    "mac_stokes_set_stream",
    "mac_stokes_set_profiling",
    "mac_stokes_get_profile",
    "mac_cnn_create",
    "mac_cnn_create_error",
    "mac_cnn_destroy",
    "mac_cnn_last_error",
    "mac_cnn_num_params",
    "mac_cnn_set_params",
    "mac_cnn_get_params",
    "mac_cnn_plan",
    "mac_cnn_forward",
    "mac_cnn_loss_grad",
    "mac_cnn_train_step",
    "mac_stokes_set_warm_start",
    "mac_stokes_reset_history",
    "mac_stokes_set_warm_history",
    "mac_stokes_last_schur_rhs",
    "mac_stokes_set_history_steps",
    "mac_stokes_history_mark",
    "mac_stokes_set_network_depth",
    "mac_warm_prepare",
    "mac_apply_divergence",
    "mac_apply_divergence_transpose",
    "mac_velocity_inverse",
    "mac_pressure_poisson_inverse",
    "mac_model_last_error",
    "mac_convection",
    "mac_combine",
    "mac_apply_K",
    "mac_grid_gradient",
    "mac_grid_divergence",
    "mac_dot",
    "mac_max_abs",
    "mac_reduce_into",
    "mac_dot_noalloc",
    "mac_noalloc_last_error",
    "mac_batch_create",
    "mac_batch_create_error",
    "mac_batch_destroy",
    "mac_batch_last_error",
    "mac_batch_layout",
    "mac_batch_set_launch",
    "mac_batch_last_sweeps",
    "mac_batch_set_history",
    "mac_batch_history_mark",
    "mac_stokes_solve_batched",
)


class StokesOpts(ctypes.Structure):
    _fields_ = [
        ("tolerance", ctypes.c_double),
        ("krylov_tolerance", ctypes.c_double),
        ("max_iterations", ctypes.c_int32),
        ("precision", ctypes.c_int32),
        ("refine_tolerance", ctypes.c_double),
        ("inner_tolerance", ctypes.c_double),
        ("max_sweeps", ctypes.c_int32),
        ("device", ctypes.c_int32),
        ("mixed_split", ctypes.c_int32),
        ("fp64_refine", ctypes.c_int32),
    ]


class SolveReport(ctypes.Structure):
    _fields_ = [
        ("residual", ctypes.c_double),
        ("divergence_inf", ctypes.c_double),
        ("velocity_max", ctypes.c_double),
        ("seconds", ctypes.c_double),
        ("sweeps", ctypes.c_int32),
        ("total_iterations", ctypes.c_int32),
        ("warm_started", ctypes.c_int32),
        ("pad0", ctypes.c_int32),
        ("first_residual", ctypes.c_double),
        ("first_divergence", ctypes.c_double),
    ]


class KernelProfile(ctypes.Structure):
    _fields_ = [
        ("gemm_ms", ctypes.c_double),
        ("gemm_flops", ctypes.c_double),
        ("gemm_bytes", ctypes.c_double),
        ("gemm_launches", ctypes.c_int64),
        ("gemm_transform_flops", ctypes.c_double),
        ("fft_ms", ctypes.c_double),
        ("fft_bytes", ctypes.c_double),
        ("fft_launches", ctypes.c_int64),
        ("fft_transform_flops", ctypes.c_double),
    ]


_lib: ctypes.CDLL | None = None

_P = ctypes.c_void_p
_D = ctypes.c_double
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64


def load() -> ctypes.CDLL:
    """Load libmacstokes.so once and declare every entry point's signature."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(
            f"{LIB_PATH} is not built: run `python -c 'import __graft_entry__ as g; g.build()'` "
            "(hipcc --offload-arch=gfx950); there is no CPU fallback")
    lib = ctypes.CDLL(os.fspath(LIB_PATH))
    lib.mac_stokes_default_opts.argtypes = [ctypes.POINTER(StokesOpts)]
    lib.mac_stokes_default_opts.restype = None
    lib.mac_stokes_create.argtypes = [_I32, _I32, _D, _D, ctypes.POINTER(StokesOpts)]
    lib.mac_stokes_create.restype = _P
    lib.mac_stokes_create_error.argtypes = []
    lib.mac_stokes_create_error.restype = ctypes.c_char_p
    lib.mac_stokes_destroy.argtypes = [_P]
    lib.mac_stokes_destroy.restype = None
    lib.mac_stokes_last_error.argtypes = [_P]
    lib.mac_stokes_last_error.restype = ctypes.c_char_p
    lib.mac_stokes_solve.argtypes = [_P, _P, _I64, _I32, _D, _D, _P, _I64, _P, _I64, _P, _I64,
                                     ctypes.POINTER(_I32), ctypes.POINTER(SolveReport), _I32]
    lib.mac_stokes_solve.restype = _I32
    lib.mac_stokes_set_stream.argtypes = [_P, _P, _I32]
    lib.mac_stokes_set_stream.restype = _I32
    lib.mac_stokes_set_profiling.argtypes = [_P, _I32]
    lib.mac_stokes_set_profiling.restype = None
    lib.mac_stokes_get_profile.argtypes = [_P, ctypes.POINTER(KernelProfile)]
    lib.mac_stokes_get_profile.restype = _I32
    F = ctypes.POINTER(ctypes.c_float)
    lib.mac_cnn_create.argtypes = [_I32, _I32]
    lib.mac_cnn_create.restype = _P
    lib.mac_cnn_create_error.argtypes = []
    lib.mac_cnn_create_error.restype = ctypes.c_char_p
    lib.mac_cnn_destroy.argtypes = [_P]
    lib.mac_cnn_destroy.restype = None
    lib.mac_cnn_last_error.argtypes = [_P]
    lib.mac_cnn_last_error.restype = ctypes.c_char_p
    lib.mac_cnn_num_params.argtypes = [_P]
    lib.mac_cnn_num_params.restype = _I64
    lib.mac_cnn_set_params.argtypes = [_P, _P]
    lib.mac_cnn_set_params.restype = _I32
    lib.mac_cnn_get_params.argtypes = [_P, _P]
    lib.mac_cnn_get_params.restype = _I32
    lib.mac_cnn_plan.argtypes = [_I32, _I32, _I32, _I32, _P]
    lib.mac_cnn_plan.restype = _I32
    lib.mac_cnn_forward.argtypes = [_P, _P, _P, _I32, _I32, _I32]
    lib.mac_cnn_forward.restype = _I32
    lib.mac_cnn_loss_grad.argtypes = [_P, _P, _P, _I32, _I32, _I32, _P, F]
    lib.mac_cnn_loss_grad.restype = _I32
    lib.mac_cnn_train_step.argtypes = [_P, _P, _P, _I32, _I32, _I32, ctypes.c_float, F]
    lib.mac_cnn_train_step.restype = _I32
    lib.mac_stokes_set_warm_start.argtypes = [_P, _I32, _P, _D]
    lib.mac_stokes_set_warm_start.restype = _I32
    lib.mac_stokes_reset_history.argtypes = [_P]
    lib.mac_stokes_reset_history.restype = _I32
    lib.mac_stokes_set_warm_history.argtypes = [_P, _I32]
    lib.mac_stokes_set_warm_history.restype = _I32
    lib.mac_stokes_set_history_steps.argtypes = [_P, _I32]
    lib.mac_stokes_set_history_steps.restype = _I32
    lib.mac_stokes_history_mark.argtypes = [_P, _I32]
    lib.mac_stokes_history_mark.restype = _I32
    lib.mac_stokes_set_network_depth.argtypes = [_P, _I32]
    lib.mac_stokes_set_network_depth.restype = _I32
    lib.mac_stokes_last_schur_rhs.argtypes = [_P, _P, _I64, _I32]
    lib.mac_stokes_last_schur_rhs.restype = _I32
    lib.mac_warm_prepare.argtypes = [_I32, _I32, _P, _I64, _P, _I64, _P, _I64, _I32, _D, _P, _P, _P]
    lib.mac_warm_prepare.restype = _I32
    lib.mac_apply_divergence.argtypes = [_P, _P, _I64, _I32, _P, _I64]
    lib.mac_apply_divergence.restype = _I32
    lib.mac_apply_divergence_transpose.argtypes = [_P, _P, _I64, _I32, _P, _I64]
    lib.mac_apply_divergence_transpose.restype = _I32
    lib.mac_velocity_inverse.argtypes = [_P, _P, _I64, _I32, _D, _D, _P, _I64]
    lib.mac_velocity_inverse.restype = _I32
    lib.mac_pressure_poisson_inverse.argtypes = [_P, _P, _I64, _I32, _P, _I64]
    lib.mac_pressure_poisson_inverse.restype = _I32
    lib.mac_model_last_error.argtypes = []
    lib.mac_model_last_error.restype = ctypes.c_char_p
    lib.mac_convection.argtypes = [_I32, _I32, _D, _D, _P, _I64, _I32, _P, _I64]
    lib.mac_convection.restype = _I32
    lib.mac_combine.argtypes = [_I64, _I32, _P, _P, _P]
    lib.mac_combine.restype = _I32
    for name in ("mac_apply_K", "mac_grid_gradient", "mac_grid_divergence"):
        getattr(lib, name).argtypes = [_I32, _I32, _D, _D, _P, _P]
        getattr(lib, name).restype = _I32
    lib.mac_dot.argtypes = [_I64, _P, _P, ctypes.POINTER(_D)]
    lib.mac_dot.restype = _I32
    lib.mac_max_abs.argtypes = [_I64, _P, ctypes.POINTER(_D)]
    lib.mac_max_abs.restype = _I32
    lib.mac_reduce_into.argtypes = [_I64, _P, _P, _I32, _P]
    lib.mac_reduce_into.restype = _I32
    lib.mac_dot_noalloc.argtypes = [_I64, _P, _P, ctypes.POINTER(_D)]
    lib.mac_dot_noalloc.restype = _I32
    lib.mac_noalloc_last_error.argtypes = []
    lib.mac_noalloc_last_error.restype = ctypes.c_char_p
    lib.mac_batch_create.argtypes = [_I32, _P, _P, _P, _P, _I32, ctypes.POINTER(StokesOpts)]
    lib.mac_batch_create.restype = _P
    lib.mac_batch_create_error.argtypes = []
    lib.mac_batch_create_error.restype = ctypes.c_char_p
    lib.mac_batch_destroy.argtypes = [_P]
    lib.mac_batch_destroy.restype = None
    lib.mac_batch_last_error.argtypes = [_P]
    lib.mac_batch_last_error.restype = ctypes.c_char_p
    lib.mac_batch_layout.argtypes = [_P, _P, _P]
    lib.mac_batch_layout.restype = _I32
    lib.mac_batch_set_launch.argtypes = [_P, _I32, _I32]
    lib.mac_batch_set_launch.restype = _I32
    lib.mac_batch_last_sweeps.argtypes = [_P]
    lib.mac_batch_last_sweeps.restype = _I32
    lib.mac_batch_set_history.argtypes = [_P, _I32, _I32]
    lib.mac_batch_set_history.restype = _I32
    lib.mac_batch_history_mark.argtypes = [_P, _I32]
    lib.mac_batch_history_mark.restype = _I32
    lib.mac_stokes_solve_batched.argtypes = [_P, _P, _P, _P, _P, _P, _P, _P, _P, _I32]
    lib.mac_stokes_solve_batched.restype = _I32
    _lib = lib
    return lib


def default_opts() -> StokesOpts:
    opts = StokesOpts()
    load().mac_stokes_default_opts(ctypes.byref(opts))
    return opts


def raise_for(code: int, message: str) -> None:
    """Map a C status to the reference's exception types (tensor_stokes.py:246-257,
    322-323, 345-351)."""
    if code == MAC_OK:
        return
    if code == MAC_INVALID_ARGUMENT:
        raise ValueError(message)
    raise RuntimeError(message)
