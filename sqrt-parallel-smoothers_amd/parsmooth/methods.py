"""Public API (parsmooth/methods.py:14-91 of the reference) on the HIP device.

``parallel=True`` (the default) runs the associative-scan formulation;
``parallel=False`` runs the sequential recursion, one device lane per
sequence, which without a nominal trajectory linearizes at the running
estimate as the reference's sequential methods do.

The fixed-point loop of ``iterated_smoothing`` follows
``parsmooth/_utils.py:136-146``: the carry starts at ``(1, x0, f(x0))`` and
``f`` (filter, then smoother, at the same nominal trajectory) is applied
again while ``criterion(i, previous, current)`` holds.  A criterion that
depends on the iteration count only (``lambda i, *_: i < n``, or
``FixedIterations(n)``) fixes the number of passes in advance; then, for the
parallel passes over models the kernels evaluate, the whole loop and the
final log-likelihood pass run as one hipGraph (``psm_iterated_smoother_graph``)
that is captured once per problem shape and replayed by later calls.  Any other
criterion runs on the host between passes, each pass one call of
``psm_filter_smoother`` (or of ``psm_seq_filter`` and ``psm_seq_smoother``).
"""
from __future__ import annotations

import dis
from collections import OrderedDict
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


class FixedIterations:
    """Criterion of exactly ``n`` fixed-point applications: ``i < n``."""

    def __init__(self, n: int):
        self.n = int(n)

    def __call__(self, i, *_):
        return i < self.n


_SKIP_OPS = frozenset(("RESUME", "NOP", "CACHE", "COPY_FREE_VARS", "MAKE_CELL", "PRECALL"))


def _fixed_count(criterion) -> Optional[int]:
    """The number of fixed-point applications a criterion allows when it
    depends on the iteration count only, else None.  Recognised without
    calling it (a criterion may keep state): ``FixedIterations(n)``, and a
    function whose whole body is ``i < n`` or ``i <= n`` on its first
    argument, with ``n`` an int constant, closure variable or global."""
    if isinstance(criterion, FixedIterations):
        return max(1, criterion.n)
    fn = criterion
    code = getattr(fn, "__code__", None)
    if code is None or code.co_argcount < 1:
        return None
    try:
        ins = [x for x in dis.get_instructions(fn) if x.opname not in _SKIP_OPS]
    except TypeError:
        return None
    if len(ins) != 4:
        return None
    first, bound, cmp, ret = ins
    if first.opname != "LOAD_FAST" or first.argval != code.co_varnames[0]:
        return None
    if cmp.opname != "COMPARE_OP" or ret.opname != "RETURN_VALUE":
        return None
    op = str(cmp.argrepr).replace("bool(", "").rstrip(")")
    if op not in ("<", "<="):
        return None
    if bound.opname == "LOAD_CONST":
        n = bound.argval
    elif bound.opname == "LOAD_DEREF" and bound.argval in code.co_freevars and fn.__closure__:
        try:
            n = fn.__closure__[code.co_freevars.index(bound.argval)].cell_contents
        except ValueError:   # empty cell
            return None
    elif bound.opname == "LOAD_GLOBAL" and bound.argval in fn.__globals__:
        n = fn.__globals__[bound.argval]
    else:
        return None
    if isinstance(n, bool) or not isinstance(n, int):
        return None
    # the carry starts after one application; criterion(i) is asked with i = 1, 2, ...
    return max(1, n if op == "<" else n + 1)


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
    count = _fixed_count(criterion) if parallel else None
    if count is not None:
        out = _whole_loop(observations, x0, transition_model, observation_model, linearization_method,
                          init_nominal_trajectory, count, return_loglikelihood)
        if out is not None:
            return out
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


# Batch runners of recent problem shapes (device buffers and the captured
# whole-loop graphs that reference them), so that repeated fixed-count calls
# replay one graph.  clear_caches() releases them.
_RUNNERS: "OrderedDict[tuple, object]" = OrderedDict()
_MAX_RUNNERS = 2


def clear_caches():
    """Release the buffers of cached runners and every captured hipGraph."""
    _RUNNERS.clear()
    N.lib().psm_clear_graphs()


def _whole_loop(observations, x0, transition_model, observation_model, linearization_method,
                init_nominal_trajectory, count: int, return_loglikelihood: bool):
    """Fixed-count iterated smoothing as one launch sequence; None when the
    models are evaluated by torch (their tables are refilled between passes)."""
    from .batch import IteratedSmoother
    from .models import DeviceFunction
    from ._base import FunctionalModel
    if not all(isinstance(m, FunctionalModel) and isinstance(m[0], DeviceFunction)
               for m in (transition_model, observation_model)):
        return None
    dtype = compute_dtype(x0)
    dev = N.device()
    sqrt = isinstance(x0, MVNSqrt)
    ys = to_dev(observations, dtype, dev)
    T = ys.shape[0]
    ys = ys.reshape(T, -1)
    m0 = to_dev(x0[0], dtype, dev)
    nx, ny = m0.shape[-1], ys.shape[1]
    method = resolve(linearization_method)
    key = (T, nx, ny, dtype, sqrt, method.name, tuple(sorted(method.kwargs.items())))
    runner = _RUNNERS.pop(key, None)
    if runner is None:
        runner = IteratedSmoother([T], transition_model, observation_model, method, dtype=dtype, sqrt=sqrt,
                                  nx=nx, ny=ny)
    else:
        runner.mods = Models(runner.prob, transition_model, observation_model, method)
    _RUNNERS[key] = runner
    while len(_RUNNERS) > _MAX_RUNNERS:
        _RUNNERS.popitem(last=False)
    runner.set_data(ys, m0.reshape(1, nx), to_dev(x0[1], dtype, dev).reshape(1, nx, nx))
    init = None
    if init_nominal_trajectory is not None:
        are_inputs_compatible(x0, init_nominal_trajectory)
        init = (init_nominal_trajectory[0], init_nominal_trajectory[1])
    # IteratedSmoother.run adds the pass at the default nominal when init is None
    out = runner.run(count, init, use_graph=True, return_loglikelihood=return_loglikelihood)
    cls = type(x0)
    res = cls(out[0].clone(), out[1].clone())
    if return_loglikelihood:
        return res, out[2][0].clone()
    return res
