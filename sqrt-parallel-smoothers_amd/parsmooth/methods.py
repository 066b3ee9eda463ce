"""Public API (parsmooth/methods.py:14-91 of the reference) on the HIP device.

``parallel=True`` (the default) runs the associative-scan formulation;
``parallel=False`` runs the sequential recursion, one device lane per
sequence, which without a nominal trajectory linearizes at the running
estimate as the reference's sequential methods do.  The fixed-point loop of
``iterated_smoothing`` runs here on the host, exactly like
``parsmooth/_utils.py:136-146``: the carry starts at ``(1, x0, f(x0))`` and
``f`` is applied again while ``criterion(i, previous, current)`` holds.  Each
application of ``f`` is one call of ``psm_filter_smoother`` (filter, then
smoother, at the same nominal trajectory), or of ``psm_seq_filter`` and
``psm_seq_smoother``.
"""
from __future__ import annotations

from typing import Callable, Optional, Union

import torch

from . import _native as N
from ._base import MVNSqrt, MVNStandard, are_inputs_compatible
from ._engine import (Models, Problem, compute_dtype, filter_pass, filter_smoother_pass, mean_sq_diff,
                      seq_filter_pass, seq_smoother_pass, to_dev)
from .linearization._methods import resolve
from .parallel._filtering import filtering as _par_filtering
from .parallel._smoothing import smoothing as _par_smoothing
from .sampling import sampling  # noqa: F401  (parsmooth/methods.py:79-91)
from .sequential._filtering import filtering as _seq_filtering
from .sequential._smoothing import smoothing as _seq_smoothing


def filtering(observations, x0: Union[MVNSqrt, MVNStandard], transition_model, observation_model,
              linearization_method: Callable, nominal_trajectory=None, parallel: bool = True,
              return_loglikelihood: bool = False):
    """methods.py:14-26."""
    run = _par_filtering if parallel else _seq_filtering
    return run(observations, x0, transition_model, observation_model, linearization_method,
               nominal_trajectory, return_loglikelihood)


def smoothing(transition_model, filter_trajectory, linearization_method: Callable,
              nominal_trajectory=None, parallel: bool = True):
    """methods.py:29-35."""
    run = _par_smoothing if parallel else _seq_smoothing
    return run(transition_model, filter_trajectory, linearization_method, nominal_trajectory)


def filter_smoother(observations, x0, transition_model, observation_model, linearization_method,
                    nominal_trajectory=None, parallel: bool = True):
    """methods.py:38-47: filter, then smoother, at the same nominal trajectory."""
    it = _Iteration(observations, x0, transition_model, observation_model, linearization_method,
                    parallel)
    nom = it.nominal(nominal_trajectory)
    return it.step(nom)


def _default_criterion(_i, nominal_traj_prev, curr_nominal_traj):
    """Continue while mean((prev.mean - cur.mean)^2) > 1e-6 (methods.py:50-51);
    the reduction runs on the device."""
    return bool(mean_sq_diff(nominal_traj_prev.mean, curr_nominal_traj.mean).item() > 1e-6)


class _Iteration:
    """The fixed-point map of iterated smoothing for one sequence (parallel or
    sequential passes)."""

    def __init__(self, observations, x0, transition_model, observation_model, linearization_method,
                 parallel: bool = True):
        self.parallel = bool(parallel)
        self.cls = type(x0)
        self.sqrt = isinstance(x0, MVNSqrt)
        self.dtype = compute_dtype(x0)
        self.dev = N.device()
        self.method = resolve(linearization_method)
        m0 = to_dev(x0[0], self.dtype, self.dev)
        self.nx = m0.shape[-1]
        ys = to_dev(observations, self.dtype, self.dev)
        self.T = ys.shape[0]
        self.ys = ys.reshape(self.T, -1)
        self.ny = self.ys.shape[1]
        self.m0 = m0.reshape(1, self.nx)
        self.L0 = to_dev(x0[1], self.dtype, self.dev).reshape(1, self.nx, self.nx)
        self.x0 = x0
        self.prob = Problem(self.nx, self.ny, [self.T], self.dtype, self.sqrt, self.dev)
        self.mods = Models(self.prob, transition_model, observation_model, self.method)

    def nominal(self, traj):
        """Device (mean, chol/cov) of a nominal trajectory.  None: the parallel
        passes' default (zero mean, identity); the sequential passes keep None
        (linearization at the running estimate)."""
        if traj is None:
            return self.prob.default_nominal() if self.parallel else (None, None)
        are_inputs_compatible(self.x0, traj)
        return (to_dev(traj[0], self.dtype, self.dev).reshape(self.T + 1, self.nx),
                to_dev(traj[1], self.dtype, self.dev).reshape(self.T + 1, self.nx, self.nx))

    def step(self, nom):
        filt = self.prob.new_traj()
        out = self.prob.new_traj()
        if self.parallel:
            filter_smoother_pass(self.prob, self.mods, self.ys, self.m0, self.L0, nom[0], nom[1], filt,
                                 out)
        else:
            seq_filter_pass(self.prob, self.mods, self.ys, self.m0, self.L0, nom[0], nom[1], False,
                            out=filt)
            seq_smoother_pass(self.prob, self.mods, filt[0], filt[1], nom[0], nom[1], out=out)
        return self.cls(*out)

    def loglikelihood(self, nom):
        run = filter_pass if self.parallel else seq_filter_pass
        _, _, ell = run(self.prob, self.mods, self.ys, self.m0, self.L0, nom[0], nom[1], True)
        return ell[0]


def iterated_smoothing(observations, x0: Union[MVNSqrt, MVNStandard], transition_model,
                       observation_model, linearization_method: Callable,
                       init_nominal_trajectory=None, parallel: bool = True,
                       criterion: Callable = _default_criterion, return_loglikelihood: bool = False):
    """IEKS (extended) / IPLS (sigma-point) smoothing (methods.py:54-76)."""
    it = _Iteration(observations, x0, transition_model, observation_model, linearization_method,
                    parallel)
    if init_nominal_trajectory is None:
        cur = it.step(it.nominal(None))
    else:
        cur = it.cls(*it.nominal(init_nominal_trajectory))
    i, prev, cur = 1, cur, it.step(tuple(cur))
    while criterion(i, prev, cur):
        i, prev, cur = i + 1, cur, it.step(tuple(cur))
    if return_loglikelihood:
        return cur, it.loglikelihood(tuple(cur))
    return cur
