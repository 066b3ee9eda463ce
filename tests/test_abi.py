"""CPU checks of the C ABI: the library loads, exports every entry point that
include/psm.h declares, reports its supported slices and rejects bad
arguments before touching a device."""
import ctypes
import os
import re

import pytest

from parsmooth import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "psm.h")).read()
    return sorted(set(re.findall(r"\b(psm_[a-z_]+)\s*\(", txt)))


def test_header_declares_what_binding_binds():
    assert _declared() == sorted(N.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = N.load_library()
    for name in _declared():
        assert hasattr(lib, name), name
    assert lib.psm_version() == 1


def test_supported_slices():
    lib = N.load_library()
    for dtype in (N.PSM_F32, N.PSM_F64):
        for form in (N.FORM_SQRT, N.FORM_STD):
            for nx in range(1, 6):
                for ny in range(1, 4):
                    assert lib.psm_supported(dtype, form, nx, ny) == 1
    assert lib.psm_supported(N.PSM_F64, N.FORM_SQRT, 6, 2) == 0
    assert lib.psm_supported(7, N.FORM_SQRT, 5, 2) == 0


def test_workspace_bytes_and_argument_errors():
    lib = N.load_library()
    prob = N.PsmProblem(N.PSM_F64, N.FORM_SQRT, 5, 2, 1, 0, 1025, 1)  # dummy non-null offsets pointer
    ws = lib.psm_workspace_bytes(ctypes.byref(prob))
    # linearized model + per-state log-likelihood + smoother elements at least
    assert ws >= 1025 * 8 * ((2 * 25 + 5 + 10 + 4 + 2) + 1 + (2 * 25 + 5 + 1))
    bad = N.PsmProblem(N.PSM_F64, N.FORM_SQRT, 5, 2, 0, 0, 10, 1)
    assert lib.psm_workspace_bytes(ctypes.byref(bad)) == 0
    assert b"sequence" in lib.psm_last_error()
    # model checks happen before any launch
    model = N.PsmModel(N.MODEL_COORDINATED_TURN, 4, 4, 0, 1, 1, 1)
    rc = lib.psm_filter(ctypes.byref(prob), ctypes.byref(model), ctypes.byref(model), None,
                        *([None] * 8), None, 0, None)
    assert rc == -1
    assert b"model dims" in lib.psm_last_error() or b"coordinated" in lib.psm_last_error()
    rc = lib.psm_filter_combine(N.PSM_F64, N.FORM_SQRT, 9, 1, None, None, None, None)
    assert rc == -2 and b"unsupported" in lib.psm_last_error()
