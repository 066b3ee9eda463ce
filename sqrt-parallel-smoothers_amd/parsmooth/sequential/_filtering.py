"""Sequential filter (parsmooth/sequential/_filtering.py:11-108) on the HIP device."""
from __future__ import annotations

from typing import Callable, Union

import torch

from .. import _native as N
from .._base import MVNSqrt, MVNStandard, are_inputs_compatible
from .._engine import Models, Problem, compute_dtype, seq_filter_pass, to_dev
from ..linearization._methods import resolve
from .._utils import _step_arrays, mvn_loglikelihood, tria


def filtering(observations, x0: Union[MVNSqrt, MVNStandard], transition_model, observation_model,
              linearization_method: Callable, nominal_trajectory=None, return_loglikelihood=False):
    """Filtered trajectory (T+1 states, x0 first) by the sequential Kalman
    recursion, and optionally the log-likelihood.  Without a nominal trajectory
    the transition is linearized at the filtered state and the observation at
    the predicted state (an extended / sigma-point Kalman filter); with one,
    both at the nominal trajectory."""
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
    nom_m = nom_L = None
    if nominal_trajectory is not None:
        nom_m = to_dev(nominal_trajectory[0], dtype, dev).reshape(T + 1, nx)
        nom_L = to_dev(nominal_trajectory[1], dtype, dev).reshape(T + 1, nx, nx)
    fm, fL, ell = seq_filter_pass(prob, mods, ys, m0.reshape(1, nx), L0, nom_m, nom_L,
                                  return_loglikelihood)
    res = type(x0)(fm, fL)
    if return_loglikelihood:
        return res, ell[0]
    return res


# ---------------------------------------------------------------- single steps
# One prediction or update at a time (parsmooth/sequential/_filtering.py:56-108),
# for callers and tests that use the steps one by one; the sequential passes
# run them inside k_seq_filter.  Arguments are device tensors or arrays (moved
# to the device as float64); the factorisations and log-densities run in the
# library (psm_tria, psm_mvn_loglikelihood), the small products in torch.
def _standard_predict(F, Q, b, x):
    """(F m + b, F P F^T + Q)."""
    F, Q, b, m, P = _step_arrays(F, Q, b, x[0], x[1])
    return MVNStandard(F @ m + b, Q + F @ P @ F.T)


def _standard_update(H, R, c, x, y):
    """Kalman update of (m, P) by y: (posterior, log N(y; H m + c, S))."""
    H, R, c, m, P, y = _step_arrays(H, R, c, x[0], x[1], y)
    y_diff = y - (H @ m + c)
    S = R + H @ P @ H.T
    chol_S = torch.linalg.cholesky(S)
    G = torch.cholesky_solve(H, chol_S).T
    G = P @ G
    ell = mvn_loglikelihood(y_diff, chol_S)
    return MVNStandard(m + G @ y_diff, P - G @ S @ G.T), ell


def _sqrt_predict(F, cholQ, b, x):
    """(F m + b, tria([F chol P, chol Q]))."""
    F, cQ, b, m, L = _step_arrays(F, cholQ, b, x[0], x[1])
    return MVNSqrt(F @ m + b, tria(torch.cat([F @ L, cQ], dim=1)))


def _sqrt_update(H, cholR, c, x, y):
    """Square-root update: tria of [[H chol P, chol R], [chol P, 0]] gives the
    innovation factor, the gain and the posterior factor."""
    H, cR, c, m, L, y = _step_arrays(H, cholR, c, x[0], x[1], y)
    nx, ny = m.shape[0], y.shape[0]
    y_diff = y - (H @ m + c)
    M = torch.cat([torch.cat([H @ L, cR], dim=1),
                   torch.cat([L, torch.zeros(nx, ny, dtype=L.dtype, device=L.device)], dim=1)], dim=0)
    chol_S = tria(M)
    G = chol_S[ny:, :ny]
    I = chol_S[:ny, :ny]
    step = torch.linalg.solve_triangular(I, y_diff[:, None], upper=False)[:, 0]
    ell = mvn_loglikelihood(y_diff, I)
    return MVNSqrt(m + G @ step, chol_S[ny:, ny:].contiguous()), ell
