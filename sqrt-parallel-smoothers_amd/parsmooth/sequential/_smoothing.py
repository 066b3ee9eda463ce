"""Sequential (Rauch-Tung-Striebel) smoother (parsmooth/sequential/_smoothing.py:12-77)
on the HIP device."""
from __future__ import annotations

from typing import Callable, Union

import torch

from .. import _native as N
from .._base import MVNSqrt, MVNStandard, are_inputs_compatible
from .._engine import Models, Problem, seq_smoother_pass, to_dev
from ..linearization._methods import resolve


def smoothing(transition_model, filter_trajectory: Union[MVNSqrt, MVNStandard],
              linearization_method: Callable, nominal_trajectory=None):
    """Smoothed trajectory from a filtered one, backwards from its last state.
    Without a nominal trajectory the transition out of state k is linearized
    at the smoothed state k+1; with one, at the nominal state k."""
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
    nom_m = nom_L = None
    if nominal_trajectory is not None:
        nom_m = to_dev(nominal_trajectory[0], dtype, dev).reshape(n_states, nx)
        nom_L = to_dev(nominal_trajectory[1], dtype, dev).reshape(n_states, nx, nx)
    sm, sL = seq_smoother_pass(prob, mods, fm, fL, nom_m, nom_L)
    return type(filter_trajectory)(sm, sL)
