"""Ragged batches: many independent sequences of uneven length smoothed by one
launch sequence per pass.

The reference handles one sequence per call; here sequences are concatenated
(sequence s owns states [state_offsets[s], state_offsets[s+1]), its prior
state first) and the filter / smoother scans run over the concatenation in
single launches.  ``IteratedSmoother`` owns every device buffer, so repeated
runs reuse them and each pass can be replayed from a captured hipGraph.
"""
from __future__ import annotations

import ctypes
from typing import Callable, Optional, Sequence

import numpy as np
import torch

from . import _native as N
from ._base import ConditionalMomentsModel, MVNSqrt, MVNStandard
from ._engine import (Models, Problem, filter_pass, filter_smoother_pass, seq_filter_pass,
                      seq_smoother_pass, to_dev)
from .linearization._methods import resolve


def _noise_dim(model) -> int:
    """Output dimension of a FunctionalModel: the size of its noise mean."""
    if isinstance(model, ConditionalMomentsModel):
        raise TypeError("pass nx (and ny) explicitly for conditional-moments models")
    q = model[1][0]
    return int(q.numel()) if isinstance(q, torch.Tensor) else int(np.prod(np.shape(q)))


# Lower bound of the sequential filter + smoother time per step of one lane
# (measured: ~18 us in FP64 on MI355X for the bearings model).
SEQ_STEP_MS_MIN = 0.004


class IteratedSmoother:
    """Iterated (IEKS / IPLS) parallel smoother for a ragged batch.

    Parameters
    ----------
    lengths: number of observations of each sequence.
    transition_model, observation_model: FunctionalModel (device functions).
    linearization_method: parsmooth.linearization.extended / cubature / ...
    dtype: torch.float64 or torch.float32.
    sqrt: square-root (MVNSqrt) or covariance (MVNStandard) form.
    sequential: run each pass as the sequential recursion, one lane per
        sequence (psm_seq_filter / psm_seq_smoother at the same nominal
        trajectory: the same posterior as the parallel passes, to rounding).
        Faster for batches of many short sequences; ``autotune`` chooses.
        These passes are two plain launches each (``use_graph`` does not apply).
    """

    def __init__(self, lengths: Sequence[int], transition_model, observation_model,
                 linearization_method: Callable, dtype=torch.float64, sqrt: bool = True, chunk: int = 0,
                 nx: int = None, ny: int = None, sequential: bool = False):
        self.sequential = bool(sequential)
        self.tuned = None
        self.dev = N.device()
        self.dtype = dtype
        self.sqrt = sqrt
        nx = int(nx) if nx is not None else _noise_dim(transition_model)
        ny = int(ny) if ny is not None else _noise_dim(observation_model)
        self.prob = Problem(nx, ny, lengths, dtype, sqrt, self.dev, chunk)
        self.mods = Models(self.prob, transition_model, observation_model, resolve(linearization_method))
        self.nx, self.ny = nx, ny
        self._bufs = [self.prob.new_traj(), self.prob.new_traj()]
        self._filt = self.prob.new_traj()
        self._default = self.prob.default_nominal()
        self._init = self.prob.new_traj()
        self.ys = None
        self.m0 = None
        self.L0 = None

    def fallback_count(self) -> int:
        """Number of FP64 recomputations since construction (FP32 square-root form)."""
        return int(self.prob.fallback_counter.item())

    def autotune(self, nominal=None, reps: int = 3):
        """Pick the fastest launch schedule (chunk, scan fan-in) for this batch
        by timing one filter-smoother pass per candidate at ``nominal`` (the
        default nominal when None); returns (chunk, scan_k, ms per pass) of the
        parallel schedule.  The sequential passes (one lane per sequence) are
        timed at the same nominal too and replace the parallel ones when faster;
        ``self.tuned`` records both times and the choice.  Cached graphs are
        dropped, since the schedule changes the launch sequence."""
        if self.ys is None:
            raise RuntimeError("call set_data first")
        if nominal is None:
            nominal = self._default
        else:
            nominal = (to_dev(nominal[0], self.dtype, self.dev).reshape(self._init[0].shape),
                       to_dev(nominal[1], self.dtype, self.dev).reshape(self._init[1].shape))
        res = self.prob.autotune(self.mods, self.ys, self.m0, self.L0, nominal[0], nominal[1], self._filt,
                                 self._bufs[1])
        N.lib().psm_clear_graphs()
        par_ms = self._time_pass(nominal, False, reps)
        # the sequential passes cost at least SEQ_STEP_MS_MIN per step of the
        # longest sequence; skip timing them when that alone loses
        seq_ms = None
        if int(self.prob.lengths.max()) * SEQ_STEP_MS_MIN < par_ms:
            seq_ms = self._time_pass(nominal, True, reps)
        self.sequential = seq_ms is not None and seq_ms < par_ms
        self.tuned = {"chunk": int(res[0]), "scan_k": int(res[1]), "parallel_ms": par_ms,
                      "sequential_ms": seq_ms, "sequential": self.sequential}
        return res

    def _time_pass(self, nominal, sequential: bool, reps: int) -> float:
        """Mean ms of one filter-smoother pass at ``nominal`` (after one untimed pass)."""
        keep = self.sequential
        self.sequential = sequential
        try:
            self.pass_once(nominal, self._bufs[1])
            start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            start.record()
            for _ in range(reps):
                self.pass_once(nominal, self._bufs[1])
            end.record()
            end.synchronize()
            return start.elapsed_time(end) / reps
        finally:
            self.sequential = keep

    def chunk_size(self) -> int:
        """States per lane in the chunk passes (as resolved by the library)."""
        import ctypes
        return int(N.lib().psm_chunk_size(ctypes.byref(self.prob.struct)))

    @property
    def state_offsets(self) -> np.ndarray:
        return self.prob.state_offsets_host

    def set_data(self, ys, m0, L0):
        """ys: (n_obs, ny) concatenated observations; m0: (n_seq, nx);
        L0: (n_seq, nx, nx) prior chol (sqrt form) or covariance."""
        p = self.prob
        self.ys = to_dev(ys, self.dtype, self.dev).reshape(p.n_obs, self.ny)
        self.m0 = to_dev(m0, self.dtype, self.dev).reshape(p.n_seq, self.nx)
        self.L0 = to_dev(L0, self.dtype, self.dev).reshape(p.n_seq, self.nx, self.nx)

    def pass_once(self, nominal, out, use_graph=False):
        if self.sequential:
            seq_filter_pass(self.prob, self.mods, self.ys, self.m0, self.L0, nominal[0], nominal[1], False,
                            out=self._filt)
            seq_smoother_pass(self.prob, self.mods, self._filt[0], self._filt[1], nominal[0], nominal[1],
                              out=out)
            return out
        filter_smoother_pass(self.prob, self.mods, self.ys, self.m0, self.L0, nominal[0], nominal[1],
                             self._filt, out, use_graph)
        return out

    def run(self, n_iter: int, init_nominal=None, use_graph: bool = True,
            return_loglikelihood: bool = False):
        """Same passes as ``iterated_smoothing(..., init_nominal, criterion=lambda i, *_: i < n_iter)``:
        one pass at the default nominal when ``init_nominal`` is None, then
        ``n_iter`` passes.  Returns (mean, chol) of all states (views of internal
        buffers, valid until the next run) and optionally the per-sequence
        log-likelihood at the final nominal trajectory.

        Parallel passes over device-evaluated models run as one launch
        sequence (``psm_iterated_smoother_graph``): every pass and the final
        log-likelihood pass are enqueued without host work in between, and with
        ``use_graph`` the whole sequence is one hipGraph, captured on the first
        run and replayed on later runs with the same buffers."""
        if self.ys is None:
            raise RuntimeError("call set_data first")
        if init_nominal is None:
            start, n = self._default, int(n_iter) + 1
        else:
            self._init[0].copy_(to_dev(init_nominal[0], self.dtype, self.dev).reshape(self._init[0].shape))
            self._init[1].copy_(to_dev(init_nominal[1], self.dtype, self.dev).reshape(self._init[1].shape))
            start, n = self._init, int(n_iter)
        if n == 0:
            cur = self._init
        elif self.sequential or self.mods.tabulated:
            cur = start
            for k in range(n):
                cur = self.pass_once(cur, self._bufs[k % 2], use_graph)
        else:
            ell = torch.empty(self.prob.n_seq, dtype=torch.float64, device=self.dev) \
                if return_loglikelihood else None
            rc = N.lib().psm_iterated_smoother_graph(
                ctypes.byref(self.prob.struct), ctypes.byref(self.mods.trans), ctypes.byref(self.mods.obs),
                self.mods.sigma_ref(), N.ptr(self.ys), N.ptr(self.m0), N.ptr(self.L0), N.ptr(start[0]),
                N.ptr(start[1]), n, N.ptr(self._filt[0]), N.ptr(self._filt[1]), N.ptr(self._bufs[0][0]),
                N.ptr(self._bufs[0][1]), N.ptr(self._bufs[1][0]), N.ptr(self._bufs[1][1]), N.ptr(ell),
                N.ptr(self.prob.workspace), self.prob.ws_bytes, int(bool(use_graph)), N.stream_ptr())
            N.check(rc, "psm_iterated_smoother_graph")
            cur = self._bufs[(n - 1) % 2]
            if return_loglikelihood:
                return cur[0], cur[1], ell
        if return_loglikelihood:
            run = seq_filter_pass if self.sequential else filter_pass
            _, _, ell = run(self.prob, self.mods, self.ys, self.m0, self.L0, cur[0], cur[1], True,
                            out=self._filt)
            return cur[0], cur[1], ell
        return cur[0], cur[1]

    def split(self, traj: torch.Tensor):
        """Per-sequence views of a (n_states, ...) trajectory tensor."""
        o = self.prob.state_offsets_host
        return [traj[o[s]:o[s + 1]] for s in range(self.prob.n_seq)]


def iterated_smoothing(observations: Sequence, x0, transition_model, observation_model,
                       linearization_method: Callable, n_iter: int, init_nominal=None,
                       return_loglikelihood: bool = False):
    """Batched iterated smoothing with a fixed iteration count.

    observations: list of (T_s, ny) arrays; x0: MVNSqrt / MVNStandard with
    leading axis n_seq.  Returns a list of per-sequence trajectories of the same
    type as x0 (and the per-sequence log-likelihoods)."""
    sqrt = isinstance(x0, MVNSqrt)
    m0 = x0[0]
    dtype = m0.dtype if isinstance(m0, torch.Tensor) and m0.dtype in (torch.float32, torch.float64) \
        else torch.float64
    lengths = [int(np.shape(y)[0]) for y in observations]
    runner = IteratedSmoother(lengths, transition_model, observation_model, linearization_method,
                              dtype=dtype, sqrt=sqrt)
    dev = runner.dev
    ys = torch.cat([to_dev(y, dtype, dev).reshape(-1, runner.ny) for y in observations])
    runner.set_data(ys, x0[0], x0[1])
    out = runner.run(n_iter, init_nominal, use_graph=False, return_loglikelihood=return_loglikelihood)
    cls = type(x0)
    trajs = [cls(m.clone(), L.clone()) for m, L in zip(runner.split(out[0]), runner.split(out[1]))]
    if return_loglikelihood:
        return trajs, out[2]
    return trajs
