"""ctypes binding of libparsmooth_hip.so (the C ABI declared in include/psm.h).

Only plain pointers and sizes cross the boundary; torch tensors are the device
arrays on this side.  Every compute call goes to the HIP library: when the
library is missing, or when no HIP device is visible, calls raise instead of
computing anything elsewhere.
"""
from __future__ import annotations

import ctypes
import os
import threading

import torch

PSM_F32, PSM_F64 = 0, 1
PSM_F32_TANGENT, PSM_F64_TANGENT = 2, 3
SCAN_K_MASK, SCAN_TEAM, SCAN_HALF, SCAN_QUAD = 0xFF, 0x100, 0x200, 0x400   # psm_problem_t.scan_k bits
SCAN_TOP_SHIFT, SCAN_TOP_MASK = 16, 0xFF0000
FORM_SQRT, FORM_STD = 0, 1

MODEL_LINEAR = 0
MODEL_COORDINATED_TURN = 1
MODEL_BEARINGS = 2
MODEL_SIN = 3
MODEL_COS = 4
MODEL_EXP = 5
MODEL_ARCTAN = 6
MODEL_TABULATED = 7
MODEL_CONDITIONAL = 8

LIB_NAME = "libparsmooth_hip.so"
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), LIB_NAME)

# Exported symbols (kept in sync with include/psm.h; checked by the tests).
EXPORTS = (
    "psm_version", "psm_last_error", "psm_supported", "psm_linearize", "psm_filter_elements",
    "psm_filter_combine", "psm_smoother_elements", "psm_smoother_combine", "psm_workspace_bytes",
    "psm_filter", "psm_smoother", "psm_filter_smoother", "psm_mean_sq_diff", "psm_clear_graphs",
    "psm_timing_enable", "psm_timing_report", "psm_chunk_size", "psm_autotune",
    "psm_sample_workspace_bytes", "psm_sample", "psm_seq_filter", "psm_seq_smoother",
    "psm_iterated_smoother_graph", "psm_filter_smoother_noalloc", "psm_adjoint_workspace_bytes",
    "psm_adjoint_prepare", "psm_filter_smoother_vjp", "psm_filter_vjp", "psm_math_check",
    "psm_tria", "psm_cholesky_update", "psm_mvn_loglikelihood",
)


class PsmModel(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("nin", ctypes.c_int32), ("nout", ctypes.c_int32),
                ("pad_", ctypes.c_int32), ("params", ctypes.c_void_p), ("q_mean", ctypes.c_void_p),
                ("q_chol", ctypes.c_void_p)]


class PsmSigma(ctypes.Structure):
    _fields_ = [("npts", ctypes.c_int32), ("pad_", ctypes.c_int32), ("xi", ctypes.c_void_p),
                ("wm", ctypes.c_void_p), ("wc", ctypes.c_void_p)]


class PsmProblem(ctypes.Structure):
    _fields_ = [("dtype", ctypes.c_int32), ("form", ctypes.c_int32), ("nx", ctypes.c_int32),
                ("ny", ctypes.c_int32), ("n_seq", ctypes.c_int32), ("chunk", ctypes.c_int32),
                ("n_states", ctypes.c_int64), ("state_offsets", ctypes.c_void_p),
                ("fallback_ratio", ctypes.c_float), ("scan_k", ctypes.c_int32),
                ("fallback_count", ctypes.c_void_p)]


_lib = None
_lock = threading.Lock()

_P = ctypes.c_void_p
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64
_SIGS = {
    "psm_version": (ctypes.c_int, []),
    "psm_last_error": (ctypes.c_char_p, []),
    "psm_supported": (ctypes.c_int, [_I32, _I32, _I32, _I32]),
    "psm_linearize": (ctypes.c_int, [_I32, _I32, ctypes.POINTER(PsmModel), ctypes.POINTER(PsmSigma),
                                     _I64, _P, _P, _P, _P, _P, _P]),
    "psm_filter_elements": (ctypes.c_int, [_I32, _I32, _I32, _I32, _I64] + [_P] * 15),
    "psm_filter_combine": (ctypes.c_int, [_I32, _I32, _I32, _I64, _P, _P, _P, _P]),
    "psm_smoother_elements": (ctypes.c_int, [_I32, _I32, _I32, _I64] + [_P] * 9),
    "psm_smoother_combine": (ctypes.c_int, [_I32, _I32, _I32, _I64, _P, _P, _P, _P]),
    "psm_workspace_bytes": (ctypes.c_size_t, [ctypes.POINTER(PsmProblem)]),
    "psm_filter": (ctypes.c_int, [ctypes.POINTER(PsmProblem), ctypes.POINTER(PsmModel),
                                  ctypes.POINTER(PsmModel), ctypes.POINTER(PsmSigma)]
                   + [_P] * 8 + [_P, ctypes.c_size_t, _P]),
    "psm_smoother": (ctypes.c_int, [ctypes.POINTER(PsmProblem), ctypes.POINTER(PsmModel),
                                    ctypes.POINTER(PsmSigma)] + [_P] * 6 + [_P, ctypes.c_size_t, _P]),
    "psm_filter_smoother": (ctypes.c_int, [ctypes.POINTER(PsmProblem), ctypes.POINTER(PsmModel),
                                           ctypes.POINTER(PsmModel), ctypes.POINTER(PsmSigma)]
                            + [_P] * 9 + [_P, ctypes.c_size_t, _I32, _P]),
    "psm_iterated_smoother_graph": (ctypes.c_int, [ctypes.POINTER(PsmProblem), ctypes.POINTER(PsmModel),
                                                   ctypes.POINTER(PsmModel), ctypes.POINTER(PsmSigma)]
                                    + [_P] * 5 + [_I32] + [_P] * 7 + [_P, ctypes.c_size_t, _I32, _P]),
    "psm_filter_smoother_noalloc": (ctypes.c_int, [ctypes.POINTER(PsmProblem), ctypes.POINTER(PsmModel),
                                                   ctypes.POINTER(PsmModel), ctypes.POINTER(PsmSigma)]
                                    + [_P] * 10),
    "psm_adjoint_workspace_bytes": (ctypes.c_size_t, [ctypes.POINTER(PsmProblem)]),
    "psm_adjoint_prepare": (ctypes.c_int, [ctypes.POINTER(PsmProblem), ctypes.POINTER(PsmModel),
                                           ctypes.POINTER(PsmModel), ctypes.POINTER(PsmSigma)]
                            + [_P] * 9 + [_P, ctypes.c_size_t, _P, ctypes.c_size_t, _P]),
    "psm_filter_smoother_vjp": (ctypes.c_int, [ctypes.POINTER(PsmProblem), ctypes.POINTER(PsmModel),
                                               ctypes.POINTER(PsmModel), ctypes.POINTER(PsmSigma)]
                                + [_P] * 14 + [_P, ctypes.c_size_t, _P]),
    "psm_filter_vjp": (ctypes.c_int, [ctypes.POINTER(PsmProblem), ctypes.POINTER(PsmModel),
                                      ctypes.POINTER(PsmModel), ctypes.POINTER(PsmSigma)]
                       + [_P] * 10 + [_P, ctypes.c_size_t, _P]),
    "psm_mean_sq_diff": (ctypes.c_int, [_I32, _I64, _P, _P, _P, _P]),
    "psm_math_check": (ctypes.c_int, [_I64, _P, _P, _P, _P]),
    "psm_tria": (ctypes.c_int, [_I32, _I64, _I32, _I32, _P, _P]),
    "psm_cholesky_update": (ctypes.c_int, [_I32, _I64, _I32, _I32, _P, _P, ctypes.c_double, _P, _P]),
    "psm_mvn_loglikelihood": (ctypes.c_int, [_I32, _I64, _I32, _P, _P, _P, _P, _P]),
    "psm_clear_graphs": (None, []),
    "psm_timing_enable": (None, [_I32]),
    "psm_chunk_size": (ctypes.c_int, [ctypes.POINTER(PsmProblem)]),
    "psm_timing_report": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t]),
    "psm_sample_workspace_bytes": (ctypes.c_size_t, [ctypes.POINTER(PsmProblem), _I32]),
    "psm_sample": (ctypes.c_int, [ctypes.POINTER(PsmProblem), ctypes.POINTER(PsmModel),
                                  ctypes.POINTER(PsmSigma)] + [_P] * 5 + [_I32, _P, _P, ctypes.c_size_t, _P]),
    "psm_seq_filter": (ctypes.c_int, [ctypes.POINTER(PsmProblem), ctypes.POINTER(PsmModel),
                                      ctypes.POINTER(PsmModel), ctypes.POINTER(PsmSigma)]
                       + [_P] * 8 + [_P]),
    "psm_seq_smoother": (ctypes.c_int, [ctypes.POINTER(PsmProblem), ctypes.POINTER(PsmModel),
                                        ctypes.POINTER(PsmSigma)] + [_P] * 6 + [_P]),
    "psm_autotune": (ctypes.c_int, [ctypes.POINTER(PsmProblem), ctypes.POINTER(PsmModel),
                                    ctypes.POINTER(PsmModel), ctypes.POINTER(PsmSigma)]
                     + [_P] * 10 + [ctypes.c_size_t, _P, ctypes.POINTER(_I32), ctypes.POINTER(_I32),
                                    ctypes.POINTER(ctypes.c_float)]),
}


def load_library(path: str = LIB_PATH):
    """Load the shared library and declare every exported signature."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(path):
            raise RuntimeError(
                f"{LIB_NAME} is not built ({path} missing); run __graft_entry__.build() or "
                f"`make -C sqrt-parallel-smoothers_amd/csrc`")
        lib = ctypes.CDLL(path)
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


def lib():
    return _lib if _lib is not None else load_library()


def device():
    """The HIP device every compute call runs on (fails loudly without one)."""
    if not torch.cuda.is_available():
        raise RuntimeError("parsmooth needs a HIP device (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


def stream_ptr():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def check(rc: int, what: str):
    if rc != 0:
        msg = lib().psm_last_error().decode(errors="replace")
        raise RuntimeError(f"{what} failed ({rc}): {msg}")


def ptr(t):
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def dtype_code(dtype: torch.dtype) -> int:
    if dtype == torch.float64:
        return PSM_F64
    if dtype == torch.float32:
        return PSM_F32
    raise TypeError(f"unsupported dtype {dtype}; use float32 or float64")


def ptr_array(tensors):
    arr = (ctypes.c_void_p * len(tensors))(*[t.data_ptr() for t in tensors])
    return arr


def timing_enable(on: bool = True):
    lib().psm_timing_enable(1 if on else 0)


def timing_report() -> dict:
    """Per-kernel {name: (total_ms, launches)} since the last report."""
    import json
    buf = ctypes.create_string_buffer(1 << 16)
    lib().psm_timing_report(buf, len(buf))
    return {k: tuple(v) for k, v in json.loads(buf.value.decode()).items()}
