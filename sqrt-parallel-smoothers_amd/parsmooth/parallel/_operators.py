"""Batched associative operators on the HIP device
(parsmooth/parallel/_operators.py:7-125).  Elements are tuples of tensors with
an optional leading batch axis."""
from __future__ import annotations

import torch

from .. import _native as N
from .._engine import to_dev


def _prep(elems, nparts):
    ref = elems[0]
    dtype = ref.dtype if isinstance(ref, torch.Tensor) and ref.dtype in (torch.float32, torch.float64) else torch.float64
    dev = N.device()
    return [to_dev(e, dtype, dev) for e in elems], dtype, dev


def _filter_op(e1, e2, form):
    a, dtype, dev = _prep(list(e1), 5)
    b, _, _ = _prep(list(e2), 5)
    nx = a[1].shape[-1]
    batch = a[1].shape[:-1]
    n = int(torch.tensor(batch).prod().item()) if len(batch) else 1
    out = [torch.empty_like(t) for t in b]
    rc = N.lib().psm_filter_combine(N.dtype_code(dtype), form, nx, n, N.ptr_array(a), N.ptr_array(b),
                                    N.ptr_array(out), N.stream_ptr())
    N.check(rc, "psm_filter_combine")
    return tuple(out)


def _smoother_op(e1, e2, form):
    a, dtype, dev = _prep(list(e1), 3)
    b, _, _ = _prep(list(e2), 3)
    nx = a[0].shape[-1]
    batch = a[0].shape[:-1]
    n = int(torch.tensor(batch).prod().item()) if len(batch) else 1
    out = [torch.empty_like(t) for t in b]
    rc = N.lib().psm_smoother_combine(N.dtype_code(dtype), form, nx, n, N.ptr_array(a), N.ptr_array(b),
                                      N.ptr_array(out), N.stream_ptr())
    N.check(rc, "psm_smoother_combine")
    return tuple(out)


def sqrt_filtering_operator(elem1, elem2):
    """(A, b, U, eta, Z) of elem1 (+) elem2, elem1 earlier (_operators.py:43-77)."""
    return _filter_op(elem1, elem2, N.FORM_SQRT)


def standard_filtering_operator(elem1, elem2):
    """(A, b, C, eta, J) of elem1 (+) elem2 (_operators.py:7-40)."""
    return _filter_op(elem1, elem2, N.FORM_STD)


def sqrt_smoothing_operator(elem1, elem2):
    """(g, E, D), elem1 the later aggregate (_operators.py:104-125)."""
    return _smoother_op(elem1, elem2, N.FORM_SQRT)


def standard_smoothing_operator(elem1, elem2):
    """(g, E, L), elem1 the later aggregate (_operators.py:80-101)."""
    return _smoother_op(elem1, elem2, N.FORM_STD)
