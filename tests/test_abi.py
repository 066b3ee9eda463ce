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


def test_header_constants_match_binding():
    txt = open(os.path.join(ROOT, "include", "psm.h")).read()
    defs = {k: int(v, 0) for k, v in re.findall(r"#define (PSM_[A-Z0-9_]+) (0x[0-9a-fA-F]+|\d+)", txt)}
    assert defs["PSM_SCAN_K_MASK"] == N.SCAN_K_MASK and defs["PSM_SCAN_TEAM"] == N.SCAN_TEAM
    assert defs["PSM_SCAN_HALF"] == N.SCAN_HALF and defs["PSM_SCAN_QUAD"] == N.SCAN_QUAD
    assert defs["PSM_SCAN_TOP_SHIFT"] == N.SCAN_TOP_SHIFT and defs["PSM_SCAN_TOP_MASK"] == N.SCAN_TOP_MASK
    assert defs["PSM_F32_TANGENT"] == N.PSM_F32_TANGENT and defs["PSM_F64_TANGENT"] == N.PSM_F64_TANGENT


def test_library_exports_every_declared_symbol():
    lib = N.load_library()
    for name in _declared():
        assert hasattr(lib, name), name
    assert lib.psm_version() == 1


def test_supported_slices():
    lib = N.load_library()
    for dtype in (N.PSM_F32, N.PSM_F64):
        for form in (N.FORM_SQRT, N.FORM_STD):
            for nx in range(1, 9):
                for ny in range(1, 5):
                    assert lib.psm_supported(dtype, form, nx, ny) == 1
    assert lib.psm_supported(N.PSM_F64, N.FORM_SQRT, 9, 2) == 0
    assert lib.psm_supported(N.PSM_F64, N.FORM_SQRT, 5, 5) == 0
    assert lib.psm_supported(N.PSM_F64, N.FORM_SQRT, 0, 1) == 0
    assert lib.psm_supported(7, N.FORM_SQRT, 5, 2) == 0
    # tangent (forward-mode) slices: square-root form only
    for nx, ny in ((1, 1), (2, 1), (2, 2), (3, 2), (4, 2), (5, 2)):
        assert lib.psm_supported(N.PSM_F64_TANGENT, N.FORM_SQRT, nx, ny) == 1
        assert lib.psm_supported(N.PSM_F64_TANGENT, N.FORM_STD, nx, ny) == 0
    assert lib.psm_supported(N.PSM_F32_TANGENT, N.FORM_SQRT, 5, 2) == 1
    assert lib.psm_supported(N.PSM_F32_TANGENT, N.FORM_SQRT, 2, 1) == 0
    assert lib.psm_supported(N.PSM_F64_TANGENT, N.FORM_SQRT, 5, 3) == 0


def test_tangent_workspace_is_larger():
    """Tangent problems store (value, tangent) pairs and FP64 pair log-likelihood terms."""
    lib = N.load_library()
    p64 = N.PsmProblem(N.PSM_F64, N.FORM_SQRT, 5, 2, 1, 0, 1025, 1)
    pd = N.PsmProblem(N.PSM_F64_TANGENT, N.FORM_SQRT, 5, 2, 1, 0, 1025, 1)
    w64, wd = lib.psm_workspace_bytes(ctypes.byref(p64)), lib.psm_workspace_bytes(ctypes.byref(pd))
    assert 1.5 * w64 <= wd <= 2.5 * w64


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


def test_schedule_environment_is_validated(monkeypatch):
    """PSM_CHUNK / PSM_SCAN_K that are empty, malformed or out of range fall back
    to the defaults instead of reaching the planner (chunk 0 would divide by
    zero)."""
    lib = N.load_library()
    p = N.PsmProblem(N.PSM_F64, N.FORM_SQRT, 5, 2, 1, 0, 131073, 1)
    want = lib.psm_chunk_size(ctypes.byref(p))
    ws = lib.psm_workspace_bytes(ctypes.byref(p))
    for bad in ("", "abc", "0", "-3", "4x", "99999999999"):
        monkeypatch.setenv("PSM_CHUNK", bad)
        monkeypatch.setenv("PSM_SCAN_K", bad)
        assert lib.psm_chunk_size(ctypes.byref(p)) == want, bad
        assert lib.psm_workspace_bytes(ctypes.byref(p)) == ws, bad
    monkeypatch.setenv("PSM_CHUNK", "7")
    assert lib.psm_chunk_size(ctypes.byref(p)) == 7


def test_schedule_fields_change_the_plan():
    """chunk and scan_k of the problem select the launch schedule; the workspace
    follows it (interleaved buffers are padded to whole 64-chunk groups)."""
    lib = N.load_library()
    sizes = {}
    for chunk in (2, 4, 8, 16, 32):
        for k in (2, 4, 8):
            p = N.PsmProblem(N.PSM_F64, N.FORM_SQRT, 5, 2, 1, chunk, 100001, 1)
            p.scan_k = k
            assert lib.psm_chunk_size(ctypes.byref(p)) == chunk
            sizes[chunk, k] = lib.psm_workspace_bytes(ctypes.byref(p))
            assert sizes[chunk, k] > 100001 * 8 * (71 + 56)
    assert len(set(sizes.values())) > 1
    # default chunk: 4, larger to keep the aggregates within one round (<= 32768), at most 16
    for M, want in ((1025, 4), (131072, 4), (131073, 5), (262145, 9), (1 << 21, 16), (1 << 24, 16)):
        p = N.PsmProblem(N.PSM_F64, N.FORM_SQRT, 5, 2, 1, 0, M, 1)
        assert lib.psm_chunk_size(ctypes.byref(p)) == want, (M, lib.psm_chunk_size(ctypes.byref(p)))
    # a larger top-level fold limit can only remove levels (less workspace)
    p = N.PsmProblem(N.PSM_F64, N.FORM_SQRT, 5, 2, 1, 4, 4093, 1)  # 1024 chunk aggregates
    p.scan_k = 2
    w2 = lib.psm_workspace_bytes(ctypes.byref(p))
    p.scan_k = 2 | (16 << N.SCAN_TOP_SHIFT)
    assert lib.psm_workspace_bytes(ctypes.byref(p)) < w2
    # autotuner argument checks happen before any launch
    p = N.PsmProblem(N.PSM_F64, N.FORM_SQRT, 5, 2, 1, 0, 1025, 1)
    m = N.PsmModel(N.MODEL_LINEAR, 5, 5, 0, 1, 1, 1)
    rc = lib.psm_autotune(ctypes.byref(p), ctypes.byref(m), ctypes.byref(m), None, *([None] * 10), 0, None,
                          None, None, None)
    assert rc != 0 and b"null" in lib.psm_last_error()


def test_sequential_entry_points_check_arguments():
    """psm_seq_filter / psm_seq_smoother refuse, before any launch: a nominal
    mean without its chol (or the reverse), tangent dtypes, caller-evaluated
    models without a nominal, and null arrays."""
    lib = N.load_library()
    p = N.PsmProblem(N.PSM_F64, N.FORM_SQRT, 5, 2, 1, 0, 11, 1)  # dummy non-null offsets pointer
    lin = N.PsmModel(N.MODEL_LINEAR, 5, 5, 0, 1, 1, 1)
    obs = N.PsmModel(N.MODEL_LINEAR, 5, 2, 0, 1, 1, 1)
    tab = N.PsmModel(N.MODEL_TABULATED, 5, 5, 0, 1, 1, 1)
    one = ctypes.c_void_p(1)
    rc = lib.psm_seq_filter(ctypes.byref(p), ctypes.byref(lin), ctypes.byref(obs), None, one, one, one,
                            one, None, one, one, None, None)
    assert rc == -1 and b"nom_m and nom_L" in lib.psm_last_error()
    rc = lib.psm_seq_filter(ctypes.byref(p), ctypes.byref(tab), ctypes.byref(obs), None, one, one, one,
                            None, None, one, one, None, None)
    assert rc == -1 and b"nominal" in lib.psm_last_error()
    rc = lib.psm_seq_filter(ctypes.byref(p), ctypes.byref(lin), ctypes.byref(obs), None, None, one, one,
                            None, None, one, one, None, None)
    assert rc == -1 and b"null" in lib.psm_last_error()
    rc = lib.psm_seq_smoother(ctypes.byref(p), ctypes.byref(tab), None, one, one, None, None, one, one, None)
    assert rc == -1 and b"nominal" in lib.psm_last_error()
    pd = N.PsmProblem(N.PSM_F64_TANGENT, N.FORM_SQRT, 5, 2, 1, 0, 11, 1)
    rc = lib.psm_seq_smoother(ctypes.byref(pd), ctypes.byref(lin), None, one, one, one, one, one, one, None)
    assert rc == -1 and b"PSM_F32 or PSM_F64" in lib.psm_last_error()


def test_are_inputs_compatible_follows_the_reference():
    """parsmooth/_base.py:25-30: TypeError unless some consecutive pair of
    arguments shares its type -- so also with fewer than two arguments."""
    from parsmooth import MVNSqrt, MVNStandard
    from parsmooth._base import are_inputs_compatible
    a, b = MVNSqrt(1.0, 2.0), MVNStandard(1.0, 2.0)
    are_inputs_compatible(a, a)
    are_inputs_compatible(a, b, b)
    for args in ((), (a,), (a, b), (b, a, b)):
        with pytest.raises(TypeError):
            are_inputs_compatible(*args)


def test_fixed_count_criterion_probe():
    """iterated_smoothing recognises criteria that depend on the iteration count
    only (run as one launch sequence) without calling them, and leaves every
    other criterion to the host loop."""
    from parsmooth.methods import FixedIterations, _default_criterion, _fixed_count
    n_iter, calls = 6, []

    def counted(i, *_):
        return i < n_iter

    def stateful(i, prev, cur):
        calls.append(i)
        return i < 3

    assert _fixed_count(lambda i, *_: i < 10) == 10
    assert _fixed_count(lambda i, prev, cur: i < 3) == 3
    assert _fixed_count(lambda i, *_: i < n_iter) == 6          # closure variable
    assert _fixed_count(lambda i, *_: i <= 4) == 5
    assert _fixed_count(counted) == 6
    assert _fixed_count(lambda i, *_: i < 0) == 1               # the first application always runs
    assert _fixed_count(FixedIterations(7)) == 7
    assert _fixed_count(stateful) is None and calls == []       # never called
    assert _fixed_count(_default_criterion) is None
    assert _fixed_count(lambda i, *_: False) is None
    assert _fixed_count(lambda i, prev, cur: i < 5 and prev.mean is not None) is None
    assert _fixed_count(lambda i, prev, cur: bool((prev[0] - cur[0]).sum() > 0)) is None
    assert _fixed_count(lambda i, *_: i < 2.5) is None
    assert _fixed_count(lambda i, *_: 3 > i) is None


def test_noalloc_refuses_problems_beyond_the_static_workspace():
    """psm_filter_smoother_noalloc rejects a problem whose workspace exceeds
    PSM_STATIC_WORKSPACE_BYTES with PSM_E_WORKSPACE before touching a device
    (here: a single FP64 sequence of 2^18 steps; the pointers are never read)."""
    lib = N.load_library()
    txt = open(os.path.join(ROOT, "include", "psm.h")).read()
    m = re.search(r"#define PSM_STATIC_WORKSPACE_BYTES \((\d+)u << (\d+)\)", txt)
    limit = int(m.group(1)) << int(m.group(2))
    dummy = ctypes.c_void_p(0x1000)
    prob = N.PsmProblem(N.PSM_F64, N.FORM_SQRT, 5, 2, 1, 0, (1 << 18) + 1, dummy, 0.0, 0, None)
    assert lib.psm_workspace_bytes(ctypes.byref(prob)) > limit
    trans = N.PsmModel(N.MODEL_COORDINATED_TURN, 5, 5, 0, dummy, dummy, dummy)
    obs = N.PsmModel(N.MODEL_BEARINGS, 5, 2, 0, dummy, dummy, dummy)
    rc = lib.psm_filter_smoother_noalloc(ctypes.byref(prob), ctypes.byref(trans), ctypes.byref(obs), None,
                                         *([dummy] * 9), None)
    assert rc == -4
    assert b"static workspace" in lib.psm_last_error()
