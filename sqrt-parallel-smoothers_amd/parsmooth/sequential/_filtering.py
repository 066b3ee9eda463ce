"""Sequential filter (parsmooth/sequential/_filtering.py:11-108) on the HIP device."""
from __future__ import annotations

from typing import Callable, Union

from .. import _native as N
from .._base import MVNSqrt, MVNStandard, are_inputs_compatible
from .._engine import Models, Problem, compute_dtype, seq_filter_pass, to_dev
from ..linearization._methods import resolve


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
