"""Device-executable model functions for ``FunctionalModel.function``.

A ``FunctionalModel`` in the reference holds an arbitrary JAX callable.  The
HIP kernels evaluate the model (and its Jacobian) per time step on the GPU, so
the function must be one the library knows how to evaluate there.  Each class
below is callable on torch tensors or numpy arrays (leading batch axes allowed)
for convenience, and carries the identifiers the kernels need.

* ``LinearFunction(A)``: x -> A x            (tests/_lgssm.py:5-39)
* ``CoordinatedTurnTransition(dt)``          (tests/bearings/bearings_utils.py:7-44)
* ``BearingsObservation(s1, s2)``            (tests/bearings/bearings_utils.py:47-69)
* ``Elementwise("sin"|"cos"|"exp"|"arctan", dim)`` (tests/test_linearization.py:157)
* ``TorchFunction(fn, nin, nout)``: any torch-callable model.  The host
  evaluates ``fn`` (and, for the extended linearization, its Jacobian with
  ``torch.func``) at the points each linearization needs and hands the values
  to the kernels as a table; the regression, the elements and the scans stay
  on the device.  A plain callable in ``FunctionalModel.function`` is wrapped
  in one automatically.

``make_parameters`` mirrors tests/bearings/bearings_utils.py:72-115 and
notebooks/bearing_data.py:93-134.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _native as N


def _as_tensor(x, dtype=None, device=None):
    if isinstance(x, torch.Tensor):
        return x.to(dtype=dtype or x.dtype, device=device or x.device)
    return torch.as_tensor(np.asarray(x), dtype=dtype, device=device)


class DeviceFunction:
    """Base class: ``kind`` (model id of include/psm.h), ``nin``, ``nout`` and the
    parameter vector the kernels read."""

    kind: int = -1
    nin: int = 0
    nout: int = 0

    def __init__(self):
        self._cache = {}

    def _param_values(self) -> np.ndarray:
        return np.zeros(1)

    def device_params(self, dtype: torch.dtype, device: torch.device) -> torch.Tensor:
        key = (dtype, str(device))
        t = self._cache.get(key)
        if t is None:
            t = torch.as_tensor(np.ascontiguousarray(self._param_values(), dtype=np.float64)).to(
                dtype=dtype, device=device)
            self._cache[key] = t
        return t

    def __call__(self, x):  # pragma: no cover - exercised through subclasses
        raise NotImplementedError


class LinearFunction(DeviceFunction):
    """x -> A x."""

    kind = N.MODEL_LINEAR

    def __init__(self, A):
        super().__init__()
        self.A = np.asarray(A.detach().cpu() if isinstance(A, torch.Tensor) else A, dtype=np.float64)
        if self.A.ndim != 2:
            raise ValueError("A must be a matrix")
        self.nout, self.nin = self.A.shape

    def _param_values(self):
        return self.A.reshape(-1)

    def __call__(self, x):
        if isinstance(x, torch.Tensor):
            A = torch.as_tensor(self.A, dtype=x.dtype, device=x.device)
            return x @ A.T
        return np.asarray(x) @ self.A.T


class CoordinatedTurnTransition(DeviceFunction):
    """Coordinated-turn transition on (x, y, vx, vy, w) with step dt; at
    |w| < 1e-6 the terms sin(w dt)/w and (cos(w dt) - 1)/w are replaced by dt and
    0 (tests/bearings/bearings_utils.py:24-37)."""

    kind = N.MODEL_COORDINATED_TURN
    nin = 5
    nout = 5

    def __init__(self, dt):
        super().__init__()
        self.dt = float(dt)

    def _param_values(self):
        return np.array([self.dt])

    def __call__(self, x):
        lib = torch if isinstance(x, torch.Tensor) else np
        w = x[..., 4]
        c = lib.cos(w * self.dt)
        s = lib.sin(w * self.dt)
        small = lib.abs(w) < 1e-6
        safe_w = lib.where(small, lib.ones_like(w), w)
        a = lib.where(small, self.dt * lib.ones_like(w), s / safe_w)
        b = lib.where(small, lib.zeros_like(w), (c - 1.0) / safe_w)
        out = [x[..., 0] + a * x[..., 2] - b * x[..., 3],
               x[..., 1] + b * x[..., 2] + a * x[..., 3],
               c * x[..., 2] + s * x[..., 3],
               -s * x[..., 2] + c * x[..., 3],
               x[..., 4]]
        return lib.stack(out, -1) if lib is torch else np.stack(out, -1)


class BearingsObservation(DeviceFunction):
    """Angles (arctan2) of the position (x[0], x[1]) seen from two sensors, with
    no angle wrapping."""

    kind = N.MODEL_BEARINGS
    nout = 2

    def __init__(self, s1, s2, nin: int = 5):
        super().__init__()
        self.s1 = np.asarray(s1, dtype=np.float64).reshape(2)
        self.s2 = np.asarray(s2, dtype=np.float64).reshape(2)
        self.nin = int(nin)

    def _param_values(self):
        return np.concatenate([self.s1, self.s2])

    def __call__(self, x):
        lib = torch if isinstance(x, torch.Tensor) else np
        out = [lib.arctan2(x[..., 1] - self.s1[1], x[..., 0] - self.s1[0]),
               lib.arctan2(x[..., 1] - self.s2[1], x[..., 0] - self.s2[0])]
        return lib.stack(out, -1) if lib is torch else np.stack(out, -1)


class Elementwise(DeviceFunction):
    """Elementwise sin / cos / exp / arctan on R^dim."""

    _KINDS = {"sin": N.MODEL_SIN, "cos": N.MODEL_COS, "exp": N.MODEL_EXP, "arctan": N.MODEL_ARCTAN}

    def __init__(self, name: str, dim: int):
        super().__init__()
        if name not in self._KINDS:
            raise ValueError(f"unknown elementwise function {name!r}")
        self.name = name
        self.kind = self._KINDS[name]
        self.nin = self.nout = int(dim)

    def __call__(self, x):
        lib = torch if isinstance(x, torch.Tensor) else np
        return {"sin": lib.sin, "cos": lib.cos, "exp": lib.exp, "arctan": lib.arctan}[self.name](x)


def make_parameters(qc, qw, r, dt, s1, s2):
    """Q, R and the two model functions of the coordinated-turn bearings-only
    problem (tests/bearings/bearings_utils.py:72-115)."""
    Q = np.array([[qc * dt ** 3 / 3, 0, qc * dt ** 2 / 2, 0, 0],
                  [0, qc * dt ** 3 / 3, 0, qc * dt ** 2 / 2, 0],
                  [qc * dt ** 2 / 2, 0, qc * dt, 0, 0],
                  [0, qc * dt ** 2 / 2, 0, qc * dt, 0],
                  [0, 0, 0, 0, dt * qw]])
    R = r ** 2 * np.eye(2)
    return Q, R, BearingsObservation(s1, s2), CoordinatedTurnTransition(dt)


def population_parameters(lam, Q):
    """Conditional moments of the stochastic Ricker map with Poisson counts
    (notebooks/population_model.py:132-169): x' = log(44.7) + x - exp(x) + q,
    q ~ N(0, Q); y | x ~ Poisson(lam exp(x)).  Returns, in the reference's
    order, torch callables of the state (shape (1,)) for the transition mean,
    transition covariance, observation mean, observation covariance,
    transition chol and observation chol -- the pieces of the two
    ``ConditionalMomentsModel``s."""
    Qm = np.asarray(Q, dtype=np.float64).reshape(1, 1)
    lam = float(lam)
    c = lambda a, x: torch.as_tensor(a, dtype=x.dtype, device=x.device)   # noqa: E731

    def mean_transition(x):
        return float(np.log(44.7)) + x - torch.exp(x)

    def cov_transition(x):
        return c(Qm, x)

    def chol_transition(x):
        return c(np.sqrt(Qm), x)

    def mean_observation(x):
        return lam * torch.exp(x)

    def cov_observation(x):
        return (lam * torch.exp(x)).reshape(1, 1)

    def chol_observation(x):
        return torch.sqrt(lam * torch.exp(x)).reshape(1, 1)

    return (mean_transition, cov_transition, mean_observation, cov_observation, chol_transition,
            chol_observation)


def is_device_function(f) -> bool:
    return isinstance(f, DeviceFunction)


def _eval_dtype(dtype):
    """Host evaluation dtype of torch-callable models: FP32 problems evaluate in
    FP64 and round the table (as the kernels linearize FP32 problems in FP64:
    Jacobians such as d/dw (1 - cos(w dt)) / w cancel in FP32)."""
    return torch.float64 if dtype == torch.float32 else dtype


class TorchFunction(DeviceFunction):
    """A model function evaluated by torch on the host side of the boundary.

    ``fn`` maps a tensor of shape (..., nin) to (..., nout) with torch
    operations (for the extended linearization it must also work under
    ``torch.func.vmap`` / ``jacfwd``).  ``tabulate`` evaluates it at the
    linearization points: f(m) and the Jacobian per point, or f at every
    sigma point m + L xi_p."""

    kind = N.MODEL_TABULATED

    def __init__(self, fn, nin: int, nout: int):
        super().__init__()
        if isinstance(fn, DeviceFunction):
            raise TypeError("already a device function")
        self.fn = fn
        self.nin = int(nin)
        self.nout = int(nout)

    def __call__(self, x):
        return self.fn(x)

    def row_width(self, npts: int) -> int:
        """Values per table row (npts = 0: extended linearization)."""
        return npts * self.nout if npts else self.nout * (1 + self.nin)

    def _eval(self, x):
        y = self.fn(x)
        if not isinstance(y, torch.Tensor):
            y = torch.as_tensor(y, dtype=x.dtype, device=x.device)
        return y.reshape(x.shape[:-1] + (self.nout,))

    def tabulate(self, m: torch.Tensor, L: torch.Tensor, xi, out: torch.Tensor = None):
        """Rows of the value table at points m (n, nin) with factors L (n, nin, nin):
        xi None -> [f(m), J(m)] (n, nout + nout nin); else f(m + L xi_p) (n, npts nout).
        ``out`` (n, row width), when given, receives the rows in place."""
        n = m.shape[0]
        out_dtype = m.dtype
        m, L = m.to(_eval_dtype(out_dtype)), L.to(_eval_dtype(out_dtype))
        if xi is None:
            f1 = lambda v: self._eval(v.reshape(1, self.nin)).reshape(self.nout)
            y = self._eval(m)
            J = torch.func.vmap(torch.func.jacfwd(f1))(m)
            rows = torch.cat([y.reshape(n, self.nout), J.reshape(n, self.nout * self.nin)], dim=1)
        else:
            xi = xi.to(dtype=m.dtype, device=m.device)
            chi = m[:, None, :] + torch.einsum("nij,pj->npi", L, xi)
            rows = self._eval(chi.reshape(-1, self.nin)).reshape(n, xi.shape[0] * self.nout)
        rows = rows.to(out_dtype)
        if out is None:
            return rows.contiguous()
        out.copy_(rows)
        return out


def as_device_function(f, nin: int, nout: int) -> DeviceFunction:
    """``f`` itself when the kernels evaluate it, else ``f`` wrapped in a
    TorchFunction of the given dimensions."""
    if isinstance(f, DeviceFunction):
        return f
    if not callable(f):
        raise TypeError(f"model function must be callable, got {f!r}")
    return TorchFunction(f, nin, nout)


class ConditionalMoments(TorchFunction):
    """A ``ConditionalMomentsModel`` (E[y | x], and chol or covariance of
    y | x) evaluated by torch: the table rows carry the conditional mean and,
    after it, the conditional factor at every linearization point (the mean
    for the extended linearization, each sigma point otherwise)."""

    kind = N.MODEL_CONDITIONAL

    def __init__(self, mean_fn, factor_fn, nin: int, nout: int):
        super().__init__(mean_fn, nin, nout)
        self.factor_fn = factor_fn

    def row_width(self, npts: int) -> int:
        v = self.nout * self.nout
        return npts * (self.nout + v) if npts else self.nout * (1 + self.nin) + v

    def _factor(self, x):
        v = torch.func.vmap(lambda z: torch.as_tensor(self.factor_fn(z), dtype=x.dtype, device=x.device)
                            .reshape(self.nout, self.nout))(x)
        return v.reshape(x.shape[0], self.nout * self.nout)

    def _mean(self, x):
        return torch.func.vmap(lambda z: torch.as_tensor(self.fn(z), dtype=x.dtype, device=x.device)
                               .reshape(self.nout))(x)

    def tabulate(self, m: torch.Tensor, L: torch.Tensor, xi, out: torch.Tensor = None):
        n = m.shape[0]
        out_dtype = m.dtype
        m, L = m.to(_eval_dtype(out_dtype)), L.to(_eval_dtype(out_dtype))
        if xi is None:
            f1 = lambda v: torch.as_tensor(self.fn(v), dtype=v.dtype, device=v.device).reshape(self.nout)
            J = torch.func.vmap(torch.func.jacfwd(f1))(m)
            rows = torch.cat([self._mean(m), J.reshape(n, self.nout * self.nin), self._factor(m)], dim=1)
        else:
            xi = xi.to(dtype=m.dtype, device=m.device)
            chi = (m[:, None, :] + torch.einsum("nij,pj->npi", L, xi)).reshape(-1, self.nin)
            rows = torch.cat([self._mean(chi), self._factor(chi)], dim=1).reshape(n, -1)
        rows = rows.to(out_dtype)
        if out is None:
            return rows.contiguous()
        out.copy_(rows)
        return out
