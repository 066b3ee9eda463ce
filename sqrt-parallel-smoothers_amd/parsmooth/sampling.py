"""Pathwise sampling from the smoothing distribution
(parsmooth/methods.py:79-91, parsmooth/_pathwise_sampler.py:13-38, 64-127)
on the HIP device through ``psm_sample``.

The reference draws its normals from a JAX key; here ``key`` is an int seed,
a ``torch.Generator`` (on the device) or the normals themselves, a tensor of
shape (T + 1, n_samples, nx) whose row k drives state k (row T: the last
state).  Samples are returned with shape (T + 1, n_samples, nx), as the
reference's.

``parallel`` selects how the default nominal trajectory (the smoothed one) is
computed.  The reference's sequential and parallel samplers share their gains,
increments and normals (_pathwise_sampler.py:41-61 and :13-38 both start from
``_sampling_common``) and differ only in evaluating the same backward affine
recursion step by step or as a scan; the device evaluates it with its chunked
scan either way.
"""
from __future__ import annotations

import ctypes
from typing import Callable, Optional, Union

import torch

from . import _native as N
from ._base import MVNSqrt, MVNStandard, are_inputs_compatible
from ._engine import Models, Problem, to_dev
from .linearization._methods import resolve
from .parallel._smoothing import smoothing as _par_smoothing
from .sequential._smoothing import smoothing as _seq_smoothing


def _normals(key, shape, dtype, dev) -> torch.Tensor:
    if isinstance(key, torch.Tensor):
        eps = key.to(device=dev, dtype=dtype)
        if tuple(eps.shape) != tuple(shape):
            raise ValueError(f"normals must have shape {tuple(shape)}, got {tuple(eps.shape)}")
        return eps.contiguous()
    if isinstance(key, torch.Generator):
        # draw on the generator's own device (a CPU generator is accepted) and move
        return torch.randn(shape, generator=key, device=key.device, dtype=dtype).to(dev)
    gen = torch.Generator(device=dev)
    gen.manual_seed(int(key))
    return torch.randn(shape, generator=gen, device=dev, dtype=dtype)


def sampling(key, n_samples: int, transition_model, filter_trajectory: Union[MVNSqrt, MVNStandard],
             linearization_method: Callable, nominal_trajectory=None, parallel: bool = True):
    # the reference's default nominal is the smoothed trajectory (methods.py:86-87)
    if nominal_trajectory is None:
        smooth = _par_smoothing if parallel else _seq_smoothing
        nominal_trajectory = smooth(transition_model, filter_trajectory, linearization_method, None)
    are_inputs_compatible(filter_trajectory, nominal_trajectory)
    sqrt = isinstance(filter_trajectory, MVNSqrt)
    fm0 = filter_trajectory[0]
    dtype = fm0.dtype if isinstance(fm0, torch.Tensor) and fm0.dtype in (torch.float32, torch.float64) \
        else torch.float64
    dev = N.device()
    fm = to_dev(filter_trajectory[0], dtype, dev)
    n_states, nx = fm.shape
    fL = to_dev(filter_trajectory[1], dtype, dev).reshape(n_states, nx, nx)
    nom_m = to_dev(nominal_trajectory[0], dtype, dev).reshape(n_states, nx)
    nom_L = to_dev(nominal_trajectory[1], dtype, dev).reshape(n_states, nx, nx)
    n_samples = int(n_samples)
    eps = _normals(key, (n_states, n_samples, nx), dtype, dev)
    prob = Problem(nx, 1, [n_states - 1], dtype, sqrt, dev)
    mods = Models(prob, transition_model, None, resolve(linearization_method))
    mods.prepare(nom_m, nom_L, obs=False)
    nbytes = int(N.lib().psm_sample_workspace_bytes(ctypes.byref(prob.struct), n_samples))
    if nbytes == 0:
        N.check(-1, "psm_sample_workspace_bytes")
    ws = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    out = torch.empty((n_states, n_samples, nx), dtype=dtype, device=dev)
    rc = N.lib().psm_sample(ctypes.byref(prob.struct), ctypes.byref(mods.trans), mods.sigma_ref(),
                            N.ptr(fm), N.ptr(fL), N.ptr(nom_m), N.ptr(nom_L), N.ptr(eps), n_samples,
                            N.ptr(out), N.ptr(ws), nbytes, N.stream_ptr())
    N.check(rc, "psm_sample")
    return out
