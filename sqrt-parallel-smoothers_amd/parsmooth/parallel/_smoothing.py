"""Parallel smoother (parsmooth/parallel/_smoothing.py:14-85) on the HIP device."""
from __future__ import annotations

from typing import Callable, Union

import torch

from .. import _native as N
from .._base import MVNSqrt, MVNStandard, are_inputs_compatible
from .._engine import Models, Problem, smoother_pass, to_dev
from ..linearization._methods import linearize, resolve


def smoothing(transition_model, filter_trajectory: Union[MVNSqrt, MVNStandard],
              linearization_method: Callable, nominal_trajectory=None):
    if nominal_trajectory is not None:
        are_inputs_compatible(filter_trajectory, nominal_trajectory)
    sqrt = isinstance(filter_trajectory, MVNSqrt)
    fm0 = filter_trajectory[0]
    dtype = fm0.dtype if isinstance(fm0, torch.Tensor) and fm0.dtype in (torch.float32, torch.float64) \
        else torch.float64
    dev = N.device()
    fm = to_dev(filter_trajectory[0], dtype, dev)
    n_states, nx = fm.shape
    fL = to_dev(filter_trajectory[1], dtype, dev).reshape(n_states, nx, nx)
    method = resolve(linearization_method)
    prob = Problem(nx, 1, [n_states - 1], dtype, sqrt, dev)
    mods = Models(prob, transition_model, None, method)
    if nominal_trajectory is None:
        nom_m, nom_L = prob.default_nominal()
    else:
        nom_m = to_dev(nominal_trajectory[0], dtype, dev).reshape(n_states, nx)
        nom_L = to_dev(nominal_trajectory[1], dtype, dev).reshape(n_states, nx, nx)
    sm, sL = smoother_pass(prob, mods, fm, fL, nom_m, nom_L)
    return type(filter_trajectory)(sm, sL)


def _elements(linearization_method, transition_model, n_k_1, m, chol_P, form):
    F, cq, bt = linearize(transition_model, n_k_1, resolve(linearization_method))
    dtype, dev = F.dtype, F.device
    nx = F.shape[-1]
    batch = F.shape[:-2]
    n = int(torch.tensor(batch).prod().item()) if len(batch) else 1
    m = to_dev(m, dtype, dev).reshape(batch + (nx,))
    P = to_dev(chol_P, dtype, dev).reshape(batch + (nx, nx))
    g = torch.empty(batch + (nx,), dtype=dtype, device=dev)
    E = torch.empty(batch + (nx, nx), dtype=dtype, device=dev)
    D = torch.empty_like(E)
    rc = N.lib().psm_smoother_elements(N.dtype_code(dtype), form, nx, n, *[N.ptr(t) for t in
                                       (F, cq, bt, m, P, g, E, D)], N.stream_ptr())
    N.check(rc, "psm_smoother_elements")
    return g, E, D


def _sqrt_associative_params(linearization_method, transition_model, n_k_1, m, chol_P):
    """Smoother element (g, E, D) (_smoothing.py:72-85)."""
    return _elements(linearization_method, transition_model, n_k_1, m, chol_P, N.FORM_SQRT)


def _standard_associative_params(linearization_method, transition_model, n_k_1, m, P):
    """Covariance-form smoother element (g, E, L) (_smoothing.py:60-69)."""
    return _elements(linearization_method, transition_model, n_k_1, m, P, N.FORM_STD)
