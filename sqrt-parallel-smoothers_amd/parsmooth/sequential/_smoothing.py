"""Sequential (Rauch-Tung-Striebel) smoother (parsmooth/sequential/_smoothing.py:12-77)
on the HIP device."""
from __future__ import annotations

from typing import Callable, Union

import torch

from .. import _native as N
from .._base import MVNSqrt, MVNStandard, are_inputs_compatible
from .._engine import Models, Problem, seq_smoother_pass, to_dev
from ..linearization._methods import resolve
from .._utils import _step_arrays, tria


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


# ---------------------------------------------------------------- single steps
# One smoothing step at a time (parsmooth/sequential/_smoothing.py:45-77), for
# callers and tests that use the steps one by one; the sequential pass runs them
# inside k_seq_smoother.  Arguments are device tensors or arrays (moved to the
# device as float64); the factorisations run in the library (psm_tria), the
# small products and solves in torch.
def _standard_smooth(F, Q, b, xf, xs):
    """RTS step: gain P_f F^T S^-1 with S = F P_f F^T + Q."""
    F, Q, b, mf, Pf, ms, Ps = _step_arrays(F, Q, b, xf[0], xf[1], xs[0], xs[1])
    S = F @ Pf @ F.T + Q
    gain = Pf @ torch.cholesky_solve(F, torch.linalg.cholesky(S)).T
    return MVNStandard(mf + gain @ (ms - (b + F @ mf)), Pf + gain @ (Ps - S) @ gain.T)


def _sqrt_smooth(F, cholQ, b, xf, xs):
    """Square-root RTS step: tria of [[F chol P_f, chol Q], [chol P_f, 0]] gives
    the gain Phi21 Phi11^-1 and the factor tria([Phi22, gain chol P_s])."""
    F, cQ, b, mf, Lf, ms, Ls = _step_arrays(F, cholQ, b, xf[0], xf[1], xs[0], xs[1])
    nx = F.shape[0]
    Phi = torch.cat([torch.cat([F @ Lf, cQ], dim=1), torch.cat([Lf, torch.zeros_like(F)], dim=1)], dim=0)
    t = tria(Phi)
    gain = torch.linalg.solve_triangular(t[:nx, :nx], t[nx:, :nx], upper=False, left=False)
    mean = mf + gain @ (ms - (b + F @ mf))
    return MVNSqrt(mean, tria(torch.cat([t[nx:, nx:], gain @ Ls], dim=1)))
