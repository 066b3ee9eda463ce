"""Conditional-moments models of functions that are non-linear in their noise
(parsmooth/linearization/_common.py:17-66 of the reference).

``get_conditional_model(f, q, method)`` turns a model ``y = f(x, q)``,
``q ~ q``, into a ``ConditionalMomentsModel``: E[y | x] and the chol factor (or
covariance) of y given x, obtained by linearizing ``q -> f(x, q)`` under ``q``
with ``method``.  The two moments are torch callables on one point ``x``; the
package evaluates them at every linearization point of the outer method (the
mean for the extended method, every sigma point otherwise) through the
``ConditionalMoments`` model kind, and the HIP kernels do the outer regression,
the elements and the scans.

The reference evaluates ``F m_q + b`` and ``tria([F chol_q, chol_val])`` (or
``F P_q F^T + cov_val``) from the inner linearization.  With the inner
linearization's own definitions these are, for the extended method,
``f(x, m_q)`` and a factor of ``F_q P_q F_q^T`` (``F_q`` the Jacobian in q at
``m_q``, zero added noise), and for a sigma-point method the weighted mean of
``f(x, q_j)`` and a factor of its weighted covariance ``Phi``
(``F P F^T + (Phi - F P F^T) = Phi``).  They are computed in that closed form
here, which leaves no rank-one downdate of a near-zero residual.
"""
from __future__ import annotations

import numpy as np
import torch

from .._base import MVNSqrt, MVNStandard, ConditionalMomentsModel


def _tria(A: torch.Tensor) -> torch.Tensor:
    """Lower factor L (d x d) with non-negative diagonal and L L^T = A A^T for
    A of shape (d, k)."""
    d, k = A.shape
    if k < d:
        A = torch.cat([A, A.new_zeros(d, d - k)], dim=1)
    R = torch.linalg.qr(A.T, mode="reduced")[1]
    L = R.T
    s = torch.where(torch.diagonal(L) < 0, -1.0, 1.0).to(L.dtype)
    return L * s[None, :]


class _NoiseConstants:
    """m_q, its chol factor and covariance, and the method's sigma points of q,
    converted once per (dtype, device)."""

    def __init__(self, q, sqrt: bool, method):
        self.m = np.asarray(torch.as_tensor(q[0]).detach().cpu().double().numpy()).reshape(-1)
        c = np.asarray(torch.as_tensor(q[1]).detach().cpu().double().numpy())
        n = self.m.shape[0]
        c = c.reshape(n, n)
        self.chol = c if sqrt else np.linalg.cholesky(c)
        self.cov = c @ c.T if sqrt else c
        self.table = method.table_fn(n, **method.kwargs) if method.uses_sigma else None
        self._cache = {}

    def get(self, dtype, dev):
        key = (dtype, str(dev))
        hit = self._cache.get(key)
        if hit is None:
            t = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64)).to(device=dev, dtype=dtype)
            hit = {"m": t(self.m), "chol": t(self.chol), "cov": t(self.cov)}
            if self.table is not None:
                xi, wm, wc = self.table
                hit["pts"] = t(self.m[None, :] + np.asarray(xi, np.float64) @ self.chol.T)
                hit["wm"] = t(wm)
                hit["wc"] = t(wc)
            self._cache[key] = hit
        return hit


def get_conditional_model(f, q, linearization_method):
    """parsmooth/linearization/_common.py:17-66.

    f: torch callable ``f(x, q) -> y`` on single points (x and q of the same
    dimension, as the reference requires); q: the noise, ``MVNSqrt`` or
    ``MVNStandard``; linearization_method: ``extended``, ``cubature``,
    ``gauss_hermite`` or ``unscented`` (a ``functools.partial`` with keywords
    allowed) used for the q part.  Returns a ``ConditionalMomentsModel`` whose
    second callable gives a chol factor for ``MVNSqrt`` noise and a covariance
    for ``MVNStandard`` noise."""
    from ._methods import resolve
    if not isinstance(q, (MVNSqrt, MVNStandard)):
        raise TypeError("q must be an MVNSqrt or an MVNStandard")
    sqrt = isinstance(q, MVNSqrt)
    method = resolve(linearization_method)
    consts = _NoiseConstants(q, sqrt, method)
    # the reference's dimension probe (:43-46), on q's device and, if f keeps
    # its constants on the device the kernels use, there too
    devices = [q[0].device if isinstance(q[0], torch.Tensor) else torch.device("cpu")]
    if devices[0].type != "cuda" and torch.cuda.is_available():
        devices.append(torch.device("cuda", torch.cuda.current_device()))
    for dev in devices:
        probe = torch.as_tensor(consts.m, dtype=torch.float64, device=dev)
        try:
            f(probe, probe)
            break
        except Exception:  # noqa: BLE001 - the reference catches everything here
            continue
    else:
        raise NotImplementedError("`x` and `q` with different dimensions are not supported yet.")

    def _values(x):
        c = consts.get(x.dtype, x.device)
        return c, torch.func.vmap(lambda qj: torch.as_tensor(f(x, qj)).reshape(-1))(c["pts"])

    def mean(x):
        if not method.uses_sigma:
            c = consts.get(x.dtype, x.device)
            return torch.as_tensor(f(x, c["m"])).reshape(-1)
        c, fv = _values(x)
        return c["wm"] @ fv

    def chol_or_cov(x):
        if not method.uses_sigma:
            c = consts.get(x.dtype, x.device)
            F = torch.func.jacfwd(lambda qq: torch.as_tensor(f(x, qq)).reshape(-1))(c["m"])
            return _tria(F @ c["chol"]) if sqrt else F @ c["cov"] @ F.T
        c, fv = _values(x)
        dev = fv - (c["wm"] @ fv)[None, :]
        if sqrt:
            return _tria((dev * torch.sqrt(c["wc"])[:, None]).T)
        return dev.T @ (dev * c["wc"][:, None])

    return ConditionalMomentsModel(mean, chol_or_cov)
