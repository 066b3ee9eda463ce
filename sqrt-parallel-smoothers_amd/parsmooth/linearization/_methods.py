"""The linearization plugins (parsmooth/linearization/__init__.py:1-5 of the
reference): ``method(model, x) -> (F, chol_or_cov, b)``, evaluated by the HIP
library at every point of ``x`` (leading batch axes allowed)."""
from __future__ import annotations

import ctypes
import functools

import numpy as np
import torch

from .. import _native as N
from .._base import ConditionalMomentsModel, FunctionalModel, MVNSqrt, MVNStandard, are_inputs_compatible
from ..models import ConditionalMoments, TorchFunction, as_device_function
from . import _tables


def _to_dev(x, dtype, dev):
    if isinstance(x, torch.Tensor):
        return x.to(device=dev, dtype=dtype).contiguous()
    return torch.as_tensor(np.asarray(x, dtype=np.float64)).to(device=dev, dtype=dtype).contiguous()


class Method:
    """Resolved linearization: extended (table None) or a sigma-point table."""

    def __init__(self, name, table_fn=None, kwargs=None):
        self.name = name
        self.table_fn = table_fn
        self.kwargs = dict(kwargs or {})
        self._cache = {}

    @property
    def uses_sigma(self):
        return self.table_fn is not None

    def sigma(self, n, dtype, dev):
        """ctypes PsmSigma for state dimension n (device tensors cached).  The
        table is FP64 whatever the compute dtype: FP32-rounded weights would not
        sum to one, which the regression's F P F^T subtraction cannot absorb."""
        if self.table_fn is None:
            return None, None
        key = (n, str(dev))
        hit = self._cache.get(key)
        if hit is None:
            xi, wm, wc = self.table_fn(n, **self.kwargs)
            t = [torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64)).to(device=dev) for a in (xi, wm, wc)]
            s = N.PsmSigma(xi.shape[0], 0, t[0].data_ptr(), t[1].data_ptr(), t[2].data_ptr())
            hit = (s, t)
            self._cache[key] = hit
        return hit


_METHODS = {}


def resolve(method) -> Method:
    """Map a linearization callable (or a functools.partial of one with keyword
    arguments such as ``order=5``) to its device implementation.  A resolved
    method is returned as it is."""
    if isinstance(method, Method):
        return method
    kwargs = {}
    fn = method
    while isinstance(fn, functools.partial):
        if fn.args:
            raise TypeError("positional arguments in a partial linearization are not supported")
        kwargs = {**fn.keywords, **kwargs}
        fn = fn.func
    name = getattr(fn, "_psm_method", None)
    if name is None:
        raise TypeError(f"{method!r} is not a parsmooth linearization method "
                        "(extended, cubature, gauss_hermite, unscented)")
    key = (name, tuple(sorted(kwargs.items())))
    m = _METHODS.get(key)
    if m is None:
        table = {"extended": None, "cubature": _tables.cubature_table,
                 "gauss_hermite": _tables.gauss_hermite_table,
                 "unscented": _tables.unscented_table}[name]
        m = Method(name, table, kwargs)
        _METHODS[key] = m
    return m


def _conditional_function(model: ConditionalMomentsModel, nin: int, dtype, dev) -> ConditionalMoments:
    probe = model.conditional_mean(torch.zeros(nin, dtype=dtype, device=dev))
    nout = int(np.prod(np.shape(probe)))
    return ConditionalMoments(model.conditional_mean, model.conditional_covariance_or_cholesky, nin, nout)


def model_function(model: FunctionalModel, nin: int):
    """The device function of a FunctionalModel: a parsmooth.models function as
    is, any other callable wrapped in a TorchFunction (input dimension nin,
    output dimension that of the noise mean)."""
    if not isinstance(model, FunctionalModel):
        raise TypeError("model must be a FunctionalModel or a ConditionalMomentsModel")
    f, q = model
    nout = int(np.prod(np.shape(q[0]))) if not isinstance(q[0], torch.Tensor) else int(q[0].numel())
    return as_device_function(f, nin, nout)


def model_struct(model: FunctionalModel, dtype, dev, sqrt: bool, nin: int = None):
    """PsmModel for a FunctionalModel plus the tensors it points to (and the
    device function).  For a TorchFunction the value table is attached per call
    (``params`` is set by whoever tabulates)."""
    if isinstance(model, ConditionalMomentsModel):
        # no additive noise: zero mean and factor, the conditional factors come with the table
        f = _conditional_function(model, int(nin), dtype, dev)
        qm = torch.zeros(f.nout, dtype=dtype, device=dev)
        qc = torch.zeros((f.nout, f.nout), dtype=dtype, device=dev)
        s = N.PsmModel(f.kind, f.nin, f.nout, 0, 0, qm.data_ptr(), qc.data_ptr())
        return s, (None, qm, qc, f)
    f = model_function(model, nin if nin is not None else getattr(model.function, "nin", 0))
    q = model[1]
    if sqrt != isinstance(q, MVNSqrt):
        raise TypeError("the model noise must be MVNSqrt in square-root form and MVNStandard otherwise")
    qm = _to_dev(q[0], dtype, dev).reshape(-1)
    qc = _to_dev(q[1], dtype, dev).reshape(f.nout, f.nout)
    if qm.numel() != f.nout:
        raise ValueError(f"noise mean has {qm.numel()} entries, the function outputs {f.nout}")
    if isinstance(f, TorchFunction):
        params = None
        s = N.PsmModel(f.kind, f.nin, f.nout, 0, 0, qm.data_ptr(), qc.data_ptr())
    else:
        params = f.device_params(dtype, dev)
        s = N.PsmModel(f.kind, f.nin, f.nout, 0, params.data_ptr(), qm.data_ptr(), qc.data_ptr())
    return s, (params, qm, qc, f)


def linearize(model, x, method: Method):
    if isinstance(model, FunctionalModel):
        are_inputs_compatible(x, model[1])
    sqrt = isinstance(x, MVNSqrt)
    mean = x[0]
    dtype = mean.dtype if isinstance(mean, torch.Tensor) else torch.float64
    if dtype not in (torch.float32, torch.float64):
        dtype = torch.float64
    dev = N.device()
    m = _to_dev(mean, dtype, dev)
    n = m.shape[-1]
    batch = m.shape[:-1]
    L = _to_dev(x[1], dtype, dev).reshape(batch + (n, n))
    ms, st = model_struct(model, dtype, dev, sqrt, nin=n)
    f = st[3]
    if f.nin != n:
        raise ValueError(f"function expects dimension {f.nin}, x has {n}")
    cnt = int(np.prod(batch)) if len(batch) else 1
    F = torch.empty(batch + (f.nout, n), dtype=dtype, device=dev)
    C = torch.empty(batch + (f.nout, f.nout), dtype=dtype, device=dev)
    b = torch.empty(batch + (f.nout,), dtype=dtype, device=dev)
    sig, _keep = method.sigma(n, dtype, dev)
    table = None
    if isinstance(f, TorchFunction):
        mm_ = m.reshape(cnt, n)
        LL = L.reshape(cnt, n, n)
        if sig is not None and not sqrt:
            LL = torch.linalg.cholesky(LL)   # the kernel factors the covariance the same way
        table = f.tabulate(mm_, LL, _keep[0] if sig is not None else None)
        ms.params = table.data_ptr()
    rc = N.lib().psm_linearize(N.dtype_code(dtype), N.FORM_SQRT if sqrt else N.FORM_STD,
                               ctypes.byref(ms), ctypes.byref(sig) if sig is not None else None,
                               cnt, N.ptr(m), N.ptr(L), N.ptr(F), N.ptr(C), N.ptr(b), N.stream_ptr())
    N.check(rc, "psm_linearize")
    return F, C, b


def extended(model, x):
    """First-order Taylor linearization at the mean of x
    (parsmooth/linearization/_extended.py:9-40)."""
    return linearize(model, x, resolve(extended))


def cubature(model, x):
    """Spherical cubature SLR (parsmooth/linearization/_cubature.py:10-34)."""
    return linearize(model, x, resolve(cubature))


def gauss_hermite(model, x, order: int = 3):
    """Gauss-Hermite SLR with order^n points (parsmooth/linearization/_gh.py:13-41)."""
    return linearize(model, x, resolve(functools.partial(gauss_hermite, order=order)))


def unscented(model, x, alpha: float = 1.0, beta: float = 0.0, kappa=None):
    """Unscented SLR (parsmooth/linearization/_unscented.py:10-39)."""
    return linearize(model, x, resolve(functools.partial(unscented, alpha=alpha, beta=beta, kappa=kappa)))


extended._psm_method = "extended"
cubature._psm_method = "cubature"
gauss_hermite._psm_method = "gauss_hermite"
unscented._psm_method = "unscented"
