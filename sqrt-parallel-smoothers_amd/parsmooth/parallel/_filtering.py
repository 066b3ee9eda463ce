"""Parallel filter (parsmooth/parallel/_filtering.py:13-163) on the HIP device."""
from __future__ import annotations

from typing import Callable, Optional, Union

import torch

from .. import _native as N
from .._base import MVNSqrt, MVNStandard, are_inputs_compatible
from .._engine import Models, Problem, compute_dtype, filter_pass, to_dev
from ..linearization._methods import linearize, resolve


def filtering(observations, x0: Union[MVNSqrt, MVNStandard], transition_model, observation_model,
              linearization_method: Callable, nominal_trajectory=None, return_loglikelihood=False):
    """Filtered trajectory (T+1 states, x0 first) at the nominal trajectory
    (zero mean / identity when None), and optionally the log-likelihood."""
    if nominal_trajectory is not None:
        are_inputs_compatible(x0, nominal_trajectory)
    sqrt = isinstance(x0, MVNSqrt)
    dtype = compute_dtype(x0)
    dev = N.device()
    method = resolve(linearization_method)
    m0 = to_dev(x0[0], dtype, dev)
    nx = m0.shape[-1]
    ys = to_dev(observations, dtype, dev)
    T = ys.shape[0]
    ys = ys.reshape(T, -1)
    ny = ys.shape[1]
    prob = Problem(nx, ny, [T], dtype, sqrt, dev)
    mods = Models(prob, transition_model, observation_model, method)
    L0 = to_dev(x0[1], dtype, dev).reshape(1, nx, nx)
    if nominal_trajectory is None:
        nom_m, nom_L = prob.default_nominal()
    else:
        nom_m = to_dev(nominal_trajectory[0], dtype, dev).reshape(T + 1, nx)
        nom_L = to_dev(nominal_trajectory[1], dtype, dev).reshape(T + 1, nx, nx)
    fm, fL, ell = filter_pass(prob, mods, ys, m0.reshape(1, nx), L0, nom_m, nom_L, return_loglikelihood)
    res = type(x0)(fm, fL)
    if return_loglikelihood:
        return res, ell[0]
    return res


def _associative_params_one(linearization_method, transition_model, observation_model, n_k_1, n_k,
                            m0, L0, y, form):
    t_lin = linearize(transition_model, n_k_1, resolve(linearization_method))
    o_lin = linearize(observation_model, n_k, resolve(linearization_method))
    F, cq, bt = t_lin
    H, cr, c = o_lin
    dtype = F.dtype
    dev = F.device
    nx = F.shape[-1]
    ny = H.shape[-2]
    batch = F.shape[:-2]
    n = int(torch.tensor(batch).prod().item()) if len(batch) else 1
    m0 = to_dev(m0, dtype, dev).expand(batch + (nx,)).contiguous()
    L0 = to_dev(L0, dtype, dev).expand(batch + (nx, nx)).contiguous()
    y = to_dev(y, dtype, dev).reshape(batch + (ny,)).contiguous()
    A = torch.empty(batch + (nx, nx), dtype=dtype, device=dev)
    b = torch.empty(batch + (nx,), dtype=dtype, device=dev)
    U = torch.empty_like(A)
    eta = torch.empty_like(b)
    Z = torch.empty_like(A)
    rc = N.lib().psm_filter_elements(N.dtype_code(dtype), form, nx, ny, n, *[N.ptr(t) for t in
                                     (F, cq, bt, H, cr, c, m0, L0, y, A, b, U, eta, Z)], N.stream_ptr())
    N.check(rc, "psm_filter_elements")
    return (A, b, U, eta, Z), (F, cq, bt, H, cr, c)


def _sqrt_associative_params_one(linearization_method, transition_model, observation_model,
                                 n_k_1, n_k, m0, L0, y):
    """One filter element and the linearized model (_filtering.py:115-146).
    Z is returned triangularised (Z Z^T equals the reference's padded Z Z^T)."""
    return _associative_params_one(linearization_method, transition_model, observation_model, n_k_1,
                                   n_k, m0, L0, y, N.FORM_SQRT)


def _standard_associative_params_one(linearization_method, transition_model, observation_model,
                                     n_k_1, n_k, m, P, y):
    """Covariance-form element (_filtering.py:78-97)."""
    return _associative_params_one(linearization_method, transition_model, observation_model, n_k_1,
                                   n_k, m, P, y, N.FORM_STD)
