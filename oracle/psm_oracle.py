"""FP64 CPU restatement of parsmooth's square-root (and covariance-form) parallel
filtering / smoothing / iterated-smoothing path.

TEST INFRASTRUCTURE ONLY.  This module is the checker: it may be imported by
``tests/``, by ``__graft_entry__.smoke()`` and by the ``cpu_baseline`` leg of
``bench.py`` and by nothing else.  The product (the HIP library and the
``parsmooth`` package that drives it) never calls it.

Every function cites the reference file:line it restates (paths relative to the
reference repository root).  Everything is float64 numpy, vectorised over a
leading batch axis where that is natural.

Conventions
-----------
* ``tria(A)`` returns the lower-triangular ``L`` with ``L L^T = A A^T``.  The
  reference takes ``R^T`` from a LAPACK QR, whose diagonal sign is arbitrary;
  here the diagonal is normalised to be non-negative.  All downstream
  quantities the reference returns are invariant to those column signs except
  the chol factors themselves, which are therefore the unique
  (non-negative-diagonal) Cholesky factors here.
* Filter elements are tuples ``(A, b, U, eta, Z)``; smoother elements are
  ``(g, E, D)``; covariance-form counterparts ``(A, b, C, eta, J)`` and
  ``(g, E, L)``.

Parity status (see tests/test_oracle_*.py for the pins):
* tria, cholesky update, operators, elements, scans, sequential and parallel
  filter/smoother, log-likelihood, iterated smoothing, linearizations,
  sigma-point tables: pinned (closed forms, brute-force joint Gaussians,
  sqrt==std and par==seq relations, reference fixtures).
* fixed_point_vjp (implicit-gradient restatement): pinned by the scalar known
  answer of tests/test_math_utils.py:57-67 and by finite differences.
* seq_filtering / seq_smoothing on non-linear models: pinned by sequential ==
  parallel at a nominal trajectory and, without one, by the fixed points of
  their own linearization points (filtered, predicted, smoothed successor).
* get_conditional_model: pinned by the reference's linear-model test
  (tests/test_linearization.py:238-283), by models affine in q (exact
  conditional moments for every method) and by a log-normal closed form.
"""
from __future__ import annotations

import functools
import math
from typing import Callable, NamedTuple, Optional

import numpy as np


# ----------------------------------------------------------------------------
# containers (parsmooth/_base.py:5-22)
# ----------------------------------------------------------------------------
class MVNStandard(NamedTuple):
    mean: np.ndarray
    cov: np.ndarray


class MVNSqrt(NamedTuple):
    mean: np.ndarray
    chol: np.ndarray


class FunctionalModel(NamedTuple):
    """``function`` is an oracle model object (callable with a ``jac`` method)."""
    function: object
    mvn: object


class ConditionalMomentsModel(NamedTuple):
    """parsmooth/_base.py:20-22.  ``conditional_mean`` is callable on batched
    points (..., n) with a ``jac`` method (for the extended linearization);
    ``conditional_covariance_or_cholesky`` maps (..., n) to (..., d, d)."""
    conditional_mean: object
    conditional_covariance_or_cholesky: object


def _T(x):
    return np.swapaxes(x, -1, -2)


# ----------------------------------------------------------------------------
# numerics (parsmooth/_utils.py)
# ----------------------------------------------------------------------------
def tria(A):
    """parsmooth/_utils.py:22-24 -- lower L (r x min(r, c)) with L L^T = A A^T."""
    A = np.asarray(A, dtype=np.float64)
    _, R = np.linalg.qr(_T(A), mode="reduced")
    L = _T(R)
    d = np.sign(np.diagonal(L, axis1=-2, axis2=-1)).copy()
    d[d == 0] = 1.0
    return L * d[..., None, :]


def solve_lower(L, B):
    """Solve L X = B (L lower triangular; batched)."""
    return np.linalg.solve(L, B)


def solve_lower_T(L, B):
    """Solve L^T X = B."""
    return np.linalg.solve(_T(L), B)


def cholesky_update(chol, v, multiplier=1.0):
    """parsmooth/_utils.py:39-81 -- rank-one update/downdate of a lower factor.

    Column-by-column recurrence (the reference's scan over the columns of the
    factor, ``_cholesky_update``); entries above the diagonal are discarded and
    non-finite entries are set to zero, as the reference does (:78-80).
    """
    L = np.array(chol, dtype=np.float64)
    n = L.shape[0]
    omega = np.array(v, dtype=np.float64)
    b = 1.0
    out = np.zeros_like(L)
    with np.errstate(all="ignore"):
        for j in range(n):
            d = L[j, j]
            wj = omega[j]
            new_d = np.sqrt(d * d + multiplier / b * wj * wj)
            gamma = d * d * b + multiplier * wj * wj
            col = L[:, j].copy()
            omega = omega - (wj / d) * col
            out[:, j] = new_d * (col / d + (multiplier * wj / gamma) * omega)
            out[j, j] = new_d
            b = b + multiplier * (wj / d) ** 2
    out = np.tril(out)
    out[~np.isfinite(out)] = 0.0
    return out


def cholesky_update_many(chol, vectors, multiplier):
    """parsmooth/_utils.py:13-19."""
    for v in vectors:
        chol = cholesky_update(chol, v, multiplier)
    return chol


def mvn_loglikelihood(x, chol):
    """parsmooth/_utils.py:152-160 (batched over leading axes)."""
    dim = chol.shape[-1]
    y = np.linalg.solve(chol, x[..., None])[..., 0]
    const = np.sum(np.log(np.abs(np.diagonal(chol, axis1=-2, axis2=-1))), -1) + dim * np.log(2 * np.pi) / 2.0
    return -0.5 * np.sum(y * y, -1) - const


# ----------------------------------------------------------------------------
# models used on the path (tests/bearings/bearings_utils.py, tests/_lgssm.py)
# ----------------------------------------------------------------------------
class Linear:
    """x -> A x  (tests/_lgssm.py:5-39)."""

    def __init__(self, A):
        self.A = np.asarray(A, dtype=np.float64)

    def __call__(self, x):
        return np.einsum("ij,...j->...i", self.A, x)

    def jac(self, x):
        return np.broadcast_to(self.A, x.shape[:-1] + self.A.shape).copy()


class CoordinatedTurn:
    """Coordinated-turn transition, tests/bearings/bearings_utils.py:7-44.

    Branch at |w| < 1e-6: (cos wt - 1)/w -> 0 and sin wt / w -> dt with zero
    w-derivative of those two terms (the reference's lax.cond branch).
    """

    def __init__(self, dt):
        self.dt = float(dt)

    def _terms(self, w):
        dt = self.dt
        c = np.cos(w * dt)
        s = np.sin(w * dt)
        small = np.abs(w) < 1e-6
        with np.errstate(all="ignore"):
            a = np.where(small, dt, s / w)
            bb = np.where(small, 0.0, (c - 1.0) / w)
            da = np.where(small, 0.0, (dt * c - a) / w)
            db = np.where(small, 0.0, (-dt * s - bb) / w)
        return c, s, a, bb, da, db

    def __call__(self, x):
        x = np.asarray(x, dtype=np.float64)
        c, s, a, bb, _, _ = self._terms(x[..., 4])
        out = np.empty_like(x)
        out[..., 0] = x[..., 0] + a * x[..., 2] - bb * x[..., 3]
        out[..., 1] = x[..., 1] + bb * x[..., 2] + a * x[..., 3]
        out[..., 2] = c * x[..., 2] + s * x[..., 3]
        out[..., 3] = -s * x[..., 2] + c * x[..., 3]
        out[..., 4] = x[..., 4]
        return out

    def jac(self, x):
        x = np.asarray(x, dtype=np.float64)
        dt = self.dt
        c, s, a, bb, da, db = self._terms(x[..., 4])
        dc = -dt * s
        ds = dt * c
        J = np.zeros(x.shape[:-1] + (5, 5))
        x2, x3 = x[..., 2], x[..., 3]
        J[..., 0, 0] = 1.0
        J[..., 0, 2] = a
        J[..., 0, 3] = -bb
        J[..., 0, 4] = da * x2 - db * x3
        J[..., 1, 1] = 1.0
        J[..., 1, 2] = bb
        J[..., 1, 3] = a
        J[..., 1, 4] = db * x2 + da * x3
        J[..., 2, 2] = c
        J[..., 2, 3] = s
        J[..., 2, 4] = dc * x2 + ds * x3
        J[..., 3, 2] = -s
        J[..., 3, 3] = c
        J[..., 3, 4] = -ds * x2 + dc * x3
        J[..., 4, 4] = 1.0
        return J


class Bearings:
    """Two-sensor bearings, tests/bearings/bearings_utils.py:47-69 (no wrapping)."""

    def __init__(self, s1, s2):
        self.s1 = np.asarray(s1, dtype=np.float64)
        self.s2 = np.asarray(s2, dtype=np.float64)

    def __call__(self, x):
        x = np.asarray(x, dtype=np.float64)
        out = np.empty(x.shape[:-1] + (2,))
        out[..., 0] = np.arctan2(x[..., 1] - self.s1[1], x[..., 0] - self.s1[0])
        out[..., 1] = np.arctan2(x[..., 1] - self.s2[1], x[..., 0] - self.s2[0])
        return out

    def jac(self, x):
        x = np.asarray(x, dtype=np.float64)
        J = np.zeros(x.shape[:-1] + (2, x.shape[-1]))
        for r, s in enumerate((self.s1, self.s2)):
            dx = x[..., 0] - s[0]
            dy = x[..., 1] - s[1]
            q = dx * dx + dy * dy
            J[..., r, 0] = -dy / q
            J[..., r, 1] = dx / q
        return J


class Elementwise:
    """Elementwise sin / cos / exp / arctan (tests/test_linearization.py:157)."""

    _F = {"sin": (np.sin, np.cos), "cos": (np.cos, lambda x: -np.sin(x)),
          "exp": (np.exp, np.exp), "arctan": (np.arctan, lambda x: 1.0 / (1.0 + x * x))}

    def __init__(self, kind):
        self.kind = kind
        self.f, self.df = self._F[kind]

    def __call__(self, x):
        return self.f(np.asarray(x, dtype=np.float64))

    def jac(self, x):
        d = self.df(np.asarray(x, dtype=np.float64))
        return d[..., :, None] * np.eye(x.shape[-1])


def bearings_parameters(qc, qw, r, dt):
    """Q and R of tests/bearings/bearings_utils.py:72-115."""
    Q = np.array([[qc * dt ** 3 / 3, 0, qc * dt ** 2 / 2, 0, 0],
                  [0, qc * dt ** 3 / 3, 0, qc * dt ** 2 / 2, 0],
                  [qc * dt ** 2 / 2, 0, qc * dt, 0, 0],
                  [0, qc * dt ** 2 / 2, 0, qc * dt, 0],
                  [0, 0, 0, 0, dt * qw]])
    R = r ** 2 * np.eye(2)
    return Q, R


# ----------------------------------------------------------------------------
# linearization (parsmooth/linearization/*)
# ----------------------------------------------------------------------------
def get_mvnsqrt(x):
    """parsmooth/linearization/_common.py:9-14."""
    if isinstance(x, MVNSqrt):
        return x
    return MVNSqrt(x.mean, np.linalg.cholesky(x.cov))


def extended(model, x):
    """parsmooth/linearization/_extended.py:9-40,59-70; conditional moments
    :43-56 (F = jacobian of the conditional mean, the conditional factor at the
    mean, b = E[y|m] - F m)."""
    if isinstance(model, ConditionalMomentsModel):
        m = np.asarray(x.mean, dtype=np.float64)
        F = model.conditional_mean.jac(m)
        b = model.conditional_mean(m) - np.einsum("...ij,...j->...i", F, m)
        return F, np.asarray(model.conditional_covariance_or_cholesky(m), np.float64), b
    f, q = model
    m = np.asarray(x.mean, dtype=np.float64)
    F = f.jac(m)
    res = f(m)
    b = res - np.einsum("...ij,...j->...i", F, m) + q.mean
    if isinstance(x, MVNSqrt):
        return F, np.broadcast_to(q.chol, F.shape[:-2] + q.chol.shape).copy(), b
    return F, np.broadcast_to(q.cov, F.shape[:-2] + q.cov.shape).copy(), b


class SigmaTable(NamedTuple):
    xi: np.ndarray   # (npts, n) unit points
    wm: np.ndarray
    wc: np.ndarray


def cubature_table(n):
    """parsmooth/linearization/_cubature.py:63-85."""
    w = np.ones(2 * n) / (2 * n)
    xi = np.concatenate([np.eye(n), -np.eye(n)], axis=0) * np.sqrt(n)
    return SigmaTable(xi, w, w)


def _hermite_poly(p):
    """Physicists' Hermite H_0..H_p, coefficients highest-first (parsmooth/linearization/_gh.py:129-151)."""
    H = [np.array([1.0]), np.array([2.0, 0.0])]
    for i in range(2, p + 1):
        H.append(2 * np.append(H[i - 1], 0.0) - 2 * (i - 1) * np.pad(H[i - 2], (2, 0)))
    return H


def gauss_hermite_table(n, order=3):
    """parsmooth/linearization/_gh.py:73-126.

    1-D nodes are sqrt(2) x roots of H_p, weights
    2^(p-1) p! sqrt(pi) / (p^2 H_{p-1}(x_i)^2) / sqrt(pi); the n-D grid is the
    tensor product with the first coordinate varying fastest (the reference's
    tile table, :110-119).
    """
    p = order
    H = _hermite_poly(p)
    roots = np.sort(np.real(np.roots(H[p])))
    w1 = np.array([2 ** (p - 1) * math.factorial(p) * np.sqrt(np.pi) /
                   (p ** 2 * np.polyval(H[p - 1], r) ** 2) for r in roots])
    idx = np.array(np.unravel_index(np.arange(p ** n), (p,) * n, order="F"))  # (n, p^n)
    wm = np.prod(w1[idx], axis=0) / np.sqrt(np.pi) ** n
    xi = (np.sqrt(2.0) * roots[idx]).T
    return SigmaTable(xi, wm, wm)


def unscented_table(n, alpha=1.0, beta=0.0, kappa=None):
    """parsmooth/linearization/_unscented.py:42-101 (default kappa = 3 + n)."""
    if kappa is None:
        kappa = 3.0 + n
    lam = alpha ** 2 * (n + kappa) - n
    wm = np.full(2 * n + 1, 1.0 / (2 * (n + lam)))
    wm[0] = lam / (n + lam)
    wc = wm.copy()
    wc[0] = lam / (n + lam) + (1 - alpha ** 2 + beta)
    s = np.sqrt(n + lam)
    xi = np.concatenate([np.zeros((1, n)), s * np.eye(n), -s * np.eye(n)], axis=0)
    return SigmaTable(xi, wm, wc)


def cholesky_update_batch(chol, v, multiplier):
    """Batched form of cholesky_update (same recurrence, leading batch axis)."""
    L = np.array(chol, dtype=np.float64)
    n = L.shape[-1]
    omega = np.array(v, dtype=np.float64)
    b = np.ones(L.shape[:-2])
    out = np.zeros_like(L)
    with np.errstate(all="ignore"):
        for j in range(n):
            d = L[..., j, j]
            wj = omega[..., j]
            new_d = np.sqrt(d * d + multiplier / b * wj * wj)
            gamma = d * d * b + multiplier * wj * wj
            col = L[..., :, j].copy()
            omega = omega - (wj / d)[..., None] * col
            out[..., :, j] = new_d[..., None] * (col / d[..., None] + (multiplier * wj / gamma)[..., None] * omega)
            out[..., j, j] = new_d
            b = b + multiplier * (wj / d) ** 2
    out = np.tril(out)
    out[~np.isfinite(out)] = 0.0
    return out


def _sigma_point_conditional(model, x, table: SigmaTable):
    """parsmooth/linearization/_sigma_points.py:25-60 (conditional moments)."""
    xs = get_mvnsqrt(x)
    m, L = np.asarray(xs.mean, np.float64), np.asarray(xs.chol, np.float64)
    pts = m[..., None, :] + table.xi @ _T(L)
    fp = model.conditional_mean(pts)
    mf = np.einsum("p,...pd->...d", table.wm, fp)
    df = fp - mf[..., None, :]
    Psi = np.einsum("...pn,p,...pd->...nd", pts - m[..., None, :], table.wc, df)
    F = _T(np.linalg.solve(_T(L), np.linalg.solve(L, Psi)))
    b = mf - _mv(F, m)
    V = np.asarray(model.conditional_covariance_or_cholesky(pts), np.float64)   # (..., npts, d, d)
    if isinstance(x, MVNSqrt):
        sqrt_phi = _T(np.sqrt(table.wc)[:, None] * df)                           # (..., d, npts)
        temp = np.sqrt(table.wc)[:, None, None] * V                              # (..., npts, d, d)
        temp = np.swapaxes(temp, -3, -2).reshape(temp.shape[:-3] + (temp.shape[-2], -1))
        C = tria(np.concatenate([sqrt_phi, temp], -1))
        FL = F @ L
        for j in range(FL.shape[-1]):
            C = cholesky_update_batch(C, FL[..., :, j], -1.0)
        return F, C, b
    Phi = np.einsum("...pd,p,...pe->...de", df, table.wc, df)
    v_f = np.einsum("p,...pde->...de", table.wc, V)
    temp = F @ L
    return F, Phi - temp @ _T(temp) + v_f, b


def sigma_point_linearize(model, x, table: SigmaTable):
    """parsmooth/linearization/_sigma_points.py:63-100 (leading batch axes allowed)."""
    if isinstance(model, ConditionalMomentsModel):
        return _sigma_point_conditional(model, x, table)
    f, q = model
    xs = get_mvnsqrt(x)
    m, L = np.asarray(xs.mean, np.float64), np.asarray(xs.chol, np.float64)
    pts = m[..., None, :] + table.xi @ _T(L)                        # (..., npts, n)
    fp = f(pts)                                                    # (..., npts, d)
    mf = np.einsum("p,...pd->...d", table.wm, fp)
    dx = pts - m[..., None, :]
    df = fp - mf[..., None, :]
    Psi = np.einsum("...pn,p,...pd->...nd", dx, table.wc, df)
    F = _T(np.linalg.solve(_T(L), np.linalg.solve(L, Psi)))
    qm = np.asarray(q.mean, np.float64)
    b = mf - _mv(F, m) + qm
    if isinstance(x, MVNSqrt):
        sqrt_phi = np.sqrt(table.wc)[:, None] * df                  # (..., npts, d)
        npts, dout = sqrt_phi.shape[-2:]
        if npts >= dout:
            sqrt_phi = tria(_T(sqrt_phi))
        else:
            sqrt_phi = np.concatenate([_T(sqrt_phi), np.zeros(sqrt_phi.shape[:-2] + (dout, dout - npts))], -1)
        cq = np.broadcast_to(np.asarray(q.chol, np.float64), sqrt_phi.shape[:-2] + (dout, dout))
        C = tria(np.concatenate([sqrt_phi, cq], -1))
        FL = F @ L
        for j in range(FL.shape[-1]):
            C = cholesky_update_batch(C, FL[..., :, j], -1.0)
        return F, C, b
    Phi = np.einsum("...pd,p,...pe->...de", df, table.wc, df)
    P = L @ _T(L)
    Lc = Phi - F @ P @ _T(F) + np.asarray(q.cov, np.float64)
    return F, 0.5 * (Lc + _T(Lc)), b


def make_sigma_method(table_fn, **kw):
    def method(model, x):
        n = np.asarray(x.mean).shape[-1]
        return sigma_point_linearize(model, x, table_fn(n, **kw))
    return method


cubature = make_sigma_method(cubature_table)


def gauss_hermite(model, x, order=3):
    n = np.asarray(x.mean).shape[-1]
    return sigma_point_linearize(model, x, gauss_hermite_table(n, order))


def unscented(model, x, alpha=1.0, beta=0.0, kappa=None):
    n = np.asarray(x.mean).shape[-1]
    return sigma_point_linearize(model, x, unscented_table(n, alpha, beta, kappa))


def linearize_many(method, model, traj):
    """Apply a linearization method at every point of a trajectory (vmap).

    Both oracle linearizations broadcast over a leading axis, so this is a
    single batched call."""
    return method(model, type(traj)(np.asarray(traj[0], np.float64), np.asarray(traj[1], np.float64)))


# ----------------------------------------------------------------------------
# doubly non-linear models (parsmooth/linearization/_common.py:17-66)
# ----------------------------------------------------------------------------
def method_table(method, n):
    """The sigma table ``method`` linearizes with at dimension n (None for the
    extended method); a functools.partial passes its keywords on."""
    kw = {}
    while isinstance(method, functools.partial):
        kw = {**method.keywords, **kw}
        method = method.func
    if method is extended:
        return None
    if method is cubature:
        return cubature_table(n)
    if method is gauss_hermite:
        return gauss_hermite_table(n, **kw)
    if method is unscented:
        return unscented_table(n, **kw)
    raise TypeError(f"unknown linearization method {method!r}")


class _AtFixedX:
    """q -> f(x, q) at one x: the function of the inner FunctionalModel
    (parsmooth/linearization/_common.py:53,59)."""

    def __init__(self, f, x):
        self.f, self.x = f, x

    def __call__(self, q):
        return self.f(self.x, q)

    def jac(self, q):
        return self.f.jac_q(self.x, q)


class _NestedMoment:
    """The conditional mean (``which`` 'mean', :52-56) or chol / covariance
    (:58-64) of f(x, q) given x, by linearizing q -> f(x, q) under q with
    ``method``.  Callable on (..., n) points, one inner linearization per point."""

    def __init__(self, f, q, method, which):
        self.f, self.q, self.method, self.which = f, q, method, which
        self.sqrt = isinstance(q, MVNSqrt)
        zm = np.zeros_like(np.asarray(q.mean, np.float64))
        zc = np.zeros_like(np.asarray(q[1], np.float64))
        self.noise = MVNSqrt(zm, zc) if self.sqrt else MVNStandard(zm, zc)

    def _one(self, x):
        F, val, bias = self.method(FunctionalModel(_AtFixedX(self.f, x), self.noise), self.q)
        if self.which == "mean":
            return F @ np.asarray(self.q.mean, np.float64) + bias
        if self.sqrt:
            return tria(np.concatenate([F @ np.asarray(self.q.chol, np.float64), val], axis=1))
        return F @ np.asarray(self.q.cov, np.float64) @ F.T + val

    def __call__(self, x):
        x = np.asarray(x, np.float64)
        outs = [self._one(p) for p in x.reshape(-1, x.shape[-1])]
        out = np.stack(outs)
        return out.reshape(x.shape[:-1] + out.shape[1:])

    def jac(self, x):
        """Jacobian of the conditional mean in x.  F m_q + b with
        b = (weighted mean of f(x, q_j)) - F m_q + 0 is that weighted mean, so the
        Jacobian is the same weighted mean of df/dx(x, q_j) (f(x, m_q) alone for
        the extended method)."""
        x = np.asarray(x, np.float64)
        qs = get_mvnsqrt(self.q)
        m, L = np.asarray(qs.mean, np.float64), np.asarray(qs.chol, np.float64)
        table = method_table(self.method, m.shape[-1])
        if table is None:
            pts, w = m[None], np.ones(1)
        else:
            pts, w = m[None] + table.xi @ L.T, table.wm
        outs = []
        for p in x.reshape(-1, x.shape[-1]):
            outs.append(sum(wj * self.f.jac_x(p, qj) for wj, qj in zip(w, pts)))
        out = np.stack(outs)
        return out.reshape(x.shape[:-1] + out.shape[1:])


def get_conditional_model(f, q, linearization_method):
    """parsmooth/linearization/_common.py:17-66.  ``f`` is callable as f(x, q)
    (q may carry leading batch axes) with ``jac_x`` and ``jac_q`` methods."""
    try:
        f(np.asarray(q.mean, np.float64), np.asarray(q.mean, np.float64))
    except Exception:  # noqa: BLE001 - the reference catches everything here (:43-46)
        raise NotImplementedError("`x` and `q` with different dimensions are not supported yet.")
    return ConditionalMomentsModel(_NestedMoment(f, q, linearization_method, "mean"),
                                   _NestedMoment(f, q, linearization_method, "chol_or_cov"))


# ----------------------------------------------------------------------------
# associative operators (parsmooth/parallel/_operators.py)
# ----------------------------------------------------------------------------
def _mm(a, b):
    return a @ b


def _mv(a, v):
    return (a @ v[..., None])[..., 0]


def sqrt_filtering_operator(e1, e2):
    """parsmooth/parallel/_operators.py:43-77 (batched)."""
    A1, b1, U1, eta1, Z1 = e1
    A2, b2, U2, eta2, Z2 = e2
    nx = Z2.shape[-2]
    I = np.broadcast_to(np.eye(nx), A1.shape)
    Xi = np.concatenate([np.concatenate([_T(U1) @ Z2, I], -1),
                         np.concatenate([Z2, np.zeros_like(A1)], -1)], -2)
    tx = tria(Xi)
    Xi11 = tx[..., :nx, :nx]
    Xi21 = tx[..., nx:, :nx]
    Xi22 = tx[..., nx:, nx:]
    S = _T(solve_lower(Xi11, _T(U1) @ _T(A2)))       # A2 U1 Xi11^-T
    A = A2 @ A1 - S @ _T(Xi21) @ A1
    M1 = _T(solve_lower(Xi11, _T(U1)))                # U1 Xi11^-T
    v = b1 + _mv(U1 @ _T(U1), eta2)
    b = _mv(A2 @ (I - M1 @ _T(Xi21)), v) + b2
    U = tria(np.concatenate([S, U2], -1))
    N = _T(solve_lower_T(Xi11, _T(Xi21)))             # Xi21 Xi11^-1
    w = eta2 - _mv(Z2 @ _T(Z2), b1)
    eta = _mv(_T(A1) @ (I - N @ _T(U1)), w) + eta1
    Z = tria(np.concatenate([_T(A1) @ Xi22, Z1], -1))
    return A, b, U, eta, Z


def standard_filtering_operator(e1, e2):
    """parsmooth/parallel/_operators.py:7-40 (batched)."""
    A1, b1, C1, eta1, J1 = e1
    A2, b2, C2, eta2, J2 = e2
    nx = b1.shape[-1]
    I = np.broadcast_to(np.eye(nx), A1.shape)
    IpCJ = I + C1 @ J2
    IpJC = I + J2 @ C1
    AIpCJ_inv = _T(np.linalg.solve(_T(IpCJ), _T(A2)))
    AIpJC_inv = _T(np.linalg.solve(_T(IpJC), A1))
    A = AIpCJ_inv @ A1
    b = _mv(AIpCJ_inv, b1 + _mv(C1, eta2)) + b2
    C = AIpCJ_inv @ C1 @ _T(A2) + C2
    eta = _mv(AIpJC_inv, eta2 - _mv(J2, b1)) + eta1
    J = AIpJC_inv @ J2 @ A1 + J1
    return A, b, C, eta, J


def sqrt_smoothing_operator(e1, e2):
    """parsmooth/parallel/_operators.py:104-125 (e1 = later aggregate, e2 = earlier)."""
    g1, E1, D1 = e1
    g2, E2, D2 = e2
    return _mv(E2, g1) + g2, E2 @ E1, tria(np.concatenate([E2 @ D1, D2], -1))


def standard_smoothing_operator(e1, e2):
    """parsmooth/parallel/_operators.py:80-101."""
    g1, E1, L1 = e1
    g2, E2, L2 = e2
    return _mv(E2, g1) + g2, E2 @ E1, E2 @ L1 @ _T(E2) + L2


# ----------------------------------------------------------------------------
# associative scans
# ----------------------------------------------------------------------------
def _take(elems, sl):
    return tuple(e[sl] for e in elems)


def associative_scan(op, elems, reverse=False):
    """Inclusive scan with the odd/even recursive-doubling schedule of
    jax.lax.associative_scan (called at parsmooth/parallel/_filtering.py:34 and
    parsmooth/parallel/_smoothing.py:33).  ``op(a, b)``: ``a`` precedes ``b`` in
    scan order; ``reverse=True`` scans from the end (``a`` is the later aggregate).
    """
    elems = tuple(np.asarray(e, dtype=np.float64) for e in elems)
    if reverse:
        elems = tuple(e[::-1] for e in elems)

    def rec(xs):
        n = xs[0].shape[0]
        if n < 2:
            return xs
        reduced = op(_take(xs, slice(0, -1, 2)), _take(xs, slice(1, None, 2)))
        odd = rec(reduced)
        if n % 2 == 0:
            even = op(_take(odd, slice(0, -1)), _take(xs, slice(2, None, 2)))
        else:
            even = op(odd, _take(xs, slice(2, None, 2)))
        even = tuple(np.concatenate([x[:1], r]) for x, r in zip(xs, even))
        out = []
        for ev, od in zip(even, odd):
            res = np.empty((n,) + ev.shape[1:])
            res[0::2] = ev
            res[1::2] = od
            out.append(res)
        return tuple(out)

    res = rec(elems)
    if reverse:
        res = tuple(e[::-1] for e in res)
    return res


def sequential_fold_scan(op, elems, reverse=False):
    """Same inclusive scan as a left fold (order independence check)."""
    elems = tuple(np.asarray(e, dtype=np.float64) for e in elems)
    n = elems[0].shape[0]
    order = range(n - 1, -1, -1) if reverse else range(n)
    out = [np.empty_like(e) for e in elems]
    acc = None
    for k in order:
        cur = tuple(e[k:k + 1] for e in elems)
        acc = cur if acc is None else op(acc, cur)
        for o, a in zip(out, acc):
            o[k] = a[0]
    return tuple(out)


# ----------------------------------------------------------------------------
# parallel filter (parsmooth/parallel/_filtering.py)
# ----------------------------------------------------------------------------
def sqrt_filter_element(F, cq, bt, H, cr, c, m, L, y):
    """parsmooth/parallel/_filtering.py:115-146 (batched)."""
    nx = F.shape[-1]
    ny = H.shape[-2]
    m1 = _mv(F, m) + bt
    N = tria(np.concatenate([F @ L, cq], -1))
    Psi = np.concatenate([np.concatenate([H @ N, cr], -1),
                          np.concatenate([N, np.zeros(N.shape[:-1] + (ny,))], -1)], -2)
    tp = tria(Psi)
    P11 = tp[..., :ny, :ny]
    P21 = tp[..., ny:ny + nx, :ny]
    U = tp[..., ny:ny + nx, ny:]
    K = _T(solve_lower_T(P11, _T(P21)))
    HF = H @ F
    A = F - K @ HF
    b = m1 + _mv(K, y - _mv(H, m1) - c)
    Z = _T(solve_lower(P11, HF))
    eta = _mv(_T(solve_lower_T(P11, _T(Z))), y - _mv(H, bt) - c)
    if nx > ny:
        Z = np.concatenate([Z, np.zeros(Z.shape[:-1] + (nx - ny,))], -1)
    else:
        Z = tria(Z)
    return A, b, U, eta, Z


def std_filter_element(F, Q, bt, H, R, c, m, P, y):
    """parsmooth/parallel/_filtering.py:78-97 (batched)."""
    m = _mv(F, m) + bt
    P = F @ P @ _T(F) + Q
    S = H @ P @ _T(H) + R
    S_invH = np.linalg.solve(S, H)
    K = _T(S_invH @ P)
    A = F - K @ H @ F
    b_std = m + _mv(K, y - _mv(H, m) - c)
    C = P - K @ S @ _T(K)
    temp = _T(S_invH @ F)
    eta = _mv(temp, y - _mv(H, bt) - c)
    J = temp @ H @ F
    return A, b_std, C, eta, J


def _nominal_default(x0, n_states):
    m0 = np.asarray(x0[0], dtype=np.float64)
    nx = m0.shape[-1]
    return type(x0)(np.zeros((n_states, nx)), np.repeat(np.eye(nx)[None], n_states, 0))


def _linearized_ssm(method, trans, obs, nominal, sqrt):
    cls = MVNSqrt if sqrt else MVNStandard
    prev = cls(nominal[0][:-1], nominal[1][:-1])
    cur = cls(nominal[0][1:], nominal[1][1:])
    F, cq, bt = linearize_many(method, trans, prev)
    H, cr, c = linearize_many(method, obs, cur)
    return F, cq, bt, H, cr, c


def par_filtering(ys, x0, trans, obs, method, nominal=None, return_loglikelihood=False,
                  scan=associative_scan):
    """parsmooth/parallel/_filtering.py:13-61."""
    ys = np.asarray(ys, dtype=np.float64)
    T = ys.shape[0]
    sqrt = isinstance(x0, MVNSqrt)
    if nominal is None:
        nominal = _nominal_default(x0, T + 1)
    F, cq, bt, H, cr, c = _linearized_ssm(method, trans, obs, nominal, sqrt)
    m0, L0 = (np.asarray(v, dtype=np.float64) for v in x0)
    nx = m0.shape[-1]
    ms = np.concatenate([m0[None], np.zeros((T - 1, nx))])
    Ls = np.concatenate([L0[None], np.zeros((T - 1, nx, nx))])
    if sqrt:
        elems = sqrt_filter_element(F, cq, bt, H, cr, c, ms, Ls, ys)
        res = scan(sqrt_filtering_operator, elems)
    else:
        elems = std_filter_element(F, cq, bt, H, cr, c, ms, Ls, ys)
        res = scan(standard_filtering_operator, elems)
    fm = np.concatenate([m0[None], res[1]])
    fc = np.concatenate([L0[None], res[2]])
    out = type(x0)(fm, fc)
    if return_loglikelihood:
        if sqrt:
            ells = sqrt_loglikelihood(F, cq, bt, H, cr, c, fm[:-1], fc[:-1], ys)
        else:
            ells = standard_loglikelihood(F, cq, bt, H, cr, c, fm[:-1], fc[:-1], ys)
        return out, float(np.sum(ells))
    return out


def sqrt_loglikelihood(F, cq, bt, H, cr, c, m, L, y):
    """parsmooth/parallel/_filtering.py:149-154 (batched)."""
    pm = _mv(F, m) + bt
    pL = tria(np.concatenate([F @ L, cq], -1))
    om = _mv(H, pm) + c
    oL = tria(np.concatenate([H @ pL, cr], -1))
    return mvn_loglikelihood(y - om, oL)


def standard_loglikelihood(F, Q, bt, H, R, c, m, P, y):
    """parsmooth/parallel/_filtering.py:157-163 (batched)."""
    pm = _mv(F, m) + bt
    pP = F @ P @ _T(F) + Q
    om = _mv(H, pm) + c
    oP = H @ pP @ _T(H) + R
    return mvn_loglikelihood(y - om, np.linalg.cholesky(oP))


# ----------------------------------------------------------------------------
# parallel smoother (parsmooth/parallel/_smoothing.py)
# ----------------------------------------------------------------------------
def sqrt_smoother_element(F, cq, bt, m, L):
    """parsmooth/parallel/_smoothing.py:72-85 (batched)."""
    nx = F.shape[-1]
    Phi = np.concatenate([np.concatenate([F @ L, cq], -1),
                          np.concatenate([L, np.zeros_like(cq)], -1)], -2)
    tp = tria(Phi)
    P11 = tp[..., :nx, :nx]
    P21 = tp[..., nx:, :nx]
    D = tp[..., nx:, nx:]
    E = _T(np.linalg.solve(_T(P11), _T(P21)))
    g = m - _mv(E, _mv(F, m) + bt)
    return g, E, D


def std_smoother_element(F, Q, bt, m, P):
    """parsmooth/parallel/_smoothing.py:60-69 (batched)."""
    Pp = F @ P @ _T(F) + Q
    E = _T(np.linalg.solve(Pp, F @ P))
    g = m - _mv(E, _mv(F, m) + bt)
    Lc = P - E @ Pp @ _T(E)
    return g, E, Lc


def sampling(eps, trans, filt, method, nominal):
    """Pathwise samples: parsmooth/_pathwise_sampler.py:41-61 (the sequential
    recursion x_k = gain_k x_{k+1} + inc_k) with the gains and increments of
    :64-127, which are the smoother elements of parallel/_smoothing.py:60-85
    (gain = E, increment mean = g, increment chol = D, or chol(D) in covariance
    form).  eps (T+1, S, nx): eps[k] drives state k, eps[T] the last state (the
    reference's eps[0]).  Returns (T+1, S, nx)."""
    ms = np.asarray(filt[0], dtype=np.float64)
    Ps = np.asarray(filt[1], dtype=np.float64)
    sqrt = isinstance(filt, MVNSqrt)
    cls = MVNSqrt if sqrt else MVNStandard
    F, cq, bt = linearize_many(method, trans, cls(nominal[0][:-1], nominal[1][:-1]))
    if sqrt:
        g, E, D = sqrt_smoother_element(F, cq, bt, ms[:-1], Ps[:-1])
        Lk, LT = D, Ps[-1]
    else:
        g, E, D = std_smoother_element(F, cq, bt, ms[:-1], Ps[:-1])
        Lk, LT = np.linalg.cholesky(D), np.linalg.cholesky(Ps[-1])
    eps = np.asarray(eps, dtype=np.float64)
    T = ms.shape[0] - 1
    out = np.empty(eps.shape)
    x = ms[-1] + eps[T] @ LT.T
    out[T] = x
    for k in range(T - 1, -1, -1):
        x = x @ E[k].T + g[k] + eps[k] @ Lk[k].T
        out[k] = x
    return out


def par_smoothing(trans, filt, method, nominal=None, scan=associative_scan):
    """parsmooth/parallel/_smoothing.py:14-57."""
    ms = np.asarray(filt[0], dtype=np.float64)
    Ps = np.asarray(filt[1], dtype=np.float64)
    sqrt = isinstance(filt, MVNSqrt)
    if nominal is None:
        nominal = _nominal_default(filt, ms.shape[0])
    cls = MVNSqrt if sqrt else MVNStandard
    F, cq, bt = linearize_many(method, trans, cls(nominal[0][:-1], nominal[1][:-1]))
    if sqrt:
        g, E, D = sqrt_smoother_element(F, cq, bt, ms[:-1], Ps[:-1])
        op = sqrt_smoothing_operator
    else:
        g, E, D = std_smoother_element(F, cq, bt, ms[:-1], Ps[:-1])
        op = standard_smoothing_operator
    g = np.concatenate([g, ms[-1:]])
    E = np.concatenate([E, np.zeros_like(Ps[-1:])])
    D = np.concatenate([D, Ps[-1:]])
    res = scan(op, (g, E, D), reverse=True)
    return cls(res[0], res[2])


# ----------------------------------------------------------------------------
# sequential filter / smoother (parsmooth/sequential/*)
# ----------------------------------------------------------------------------
def seq_filtering(ys, x0, trans, obs, method, nominal=None, return_loglikelihood=False):
    """parsmooth/sequential/_filtering.py:11-53,56-108."""
    ys = np.asarray(ys, dtype=np.float64)
    sqrt = isinstance(x0, MVNSqrt)
    cls = type(x0)
    m, P = (np.asarray(v, dtype=np.float64) for v in x0)
    ms, Ps = [m], [P]
    ell = 0.0
    for k in range(ys.shape[0]):
        ref = cls(m, P) if nominal is None else cls(nominal[0][k], nominal[1][k])
        F, cq, bt = method(trans, ref)
        m = F @ m + bt
        P = tria(np.concatenate([F @ P, cq], 1)) if sqrt else cq + F @ P @ F.T
        ref = cls(m, P) if nominal is None else cls(nominal[0][k + 1], nominal[1][k + 1])
        H, cr, c = method(obs, ref)
        y = ys[k]
        ydiff = y - H @ m - c
        if sqrt:
            nx, ny = m.shape[0], y.shape[0]
            M = np.block([[H @ P, cr], [P, np.zeros((nx, ny))]])
            cS = tria(M)
            Pn = cS[ny:, ny:]
            G = cS[ny:, :ny]
            Ic = cS[:ny, :ny]
            m = m + G @ np.linalg.solve(Ic, ydiff)
            ell += float(mvn_loglikelihood(ydiff, Ic))
            P = Pn
        else:
            S = cr + H @ P @ H.T
            cS = np.linalg.cholesky(S)
            G = P @ np.linalg.solve(cS.T, np.linalg.solve(cS, H)).T
            m = m + G @ ydiff
            P = P - G @ S @ G.T
            ell += float(mvn_loglikelihood(ydiff, cS))
        ms.append(m)
        Ps.append(P)
    out = cls(np.stack(ms), np.stack(Ps))
    if return_loglikelihood:
        return out, ell
    return out


def seq_smoothing(trans, filt, method, nominal=None):
    """parsmooth/sequential/_smoothing.py:12-77."""
    sqrt = isinstance(filt, MVNSqrt)
    cls = type(filt)
    fm = np.asarray(filt[0], dtype=np.float64)
    fP = np.asarray(filt[1], dtype=np.float64)
    n = fm.shape[0]
    sm, sP = fm[-1], fP[-1]
    out_m, out_P = [sm], [sP]
    for k in range(n - 2, -1, -1):
        ref = cls(sm, sP) if nominal is None else cls(nominal[0][k], nominal[1][k])
        F, cq, bt = method(trans, ref)
        mf, Pf = fm[k], fP[k]
        if sqrt:
            nx = F.shape[0]
            Phi = np.block([[F @ Pf, cq], [Pf, np.zeros_like(F)]])
            tp = tria(Phi)
            P11, P21, P22 = tp[:nx, :nx], tp[nx:, :nx], tp[nx:, nx:]
            gain = np.linalg.solve(P11.T, P21.T).T
            sm = mf + gain @ (sm - (bt + F @ mf))
            sP = tria(np.concatenate([P22, gain @ sP], 1))
        else:
            S = F @ Pf @ F.T + cq
            gain = Pf @ np.linalg.solve(S, F).T
            sm_new = mf + gain @ (sm - (bt + F @ mf))
            sP = Pf + gain @ (sP - S) @ gain.T
            sm = sm_new
        out_m.append(sm)
        out_P.append(sP)
    return cls(np.stack(out_m[::-1]), np.stack(out_P[::-1]))


# ----------------------------------------------------------------------------
# methods (parsmooth/methods.py)
# ----------------------------------------------------------------------------
def filtering(ys, x0, trans, obs, method, nominal=None, parallel=True, return_loglikelihood=False):
    """parsmooth/methods.py:14-26."""
    if parallel:
        return par_filtering(ys, x0, trans, obs, method, nominal, return_loglikelihood)
    return seq_filtering(ys, x0, trans, obs, method, nominal, return_loglikelihood)


def smoothing(trans, filt, method, nominal=None, parallel=True):
    """parsmooth/methods.py:29-35."""
    if parallel:
        return par_smoothing(trans, filt, method, nominal)
    return seq_smoothing(trans, filt, method, nominal)


def filter_smoother(ys, x0, trans, obs, method, nominal=None, parallel=True):
    """parsmooth/methods.py:38-47."""
    f = filtering(ys, x0, trans, obs, method, nominal, parallel)
    return smoothing(trans, f, method, nominal, parallel)


def default_criterion(_i, prev, cur):
    """parsmooth/methods.py:50-51."""
    return np.mean((prev.mean - cur.mean) ** 2) > 1e-6


def fixed_point(f, x0, criterion):
    """parsmooth/_utils.py:136-146: carry (1, x0, f(x0)); run body while criterion."""
    i, prev, cur = 1, x0, f(x0)
    while criterion(i, prev, cur):
        i, prev, cur = i + 1, cur, f(cur)
    return cur, i


def iterated_smoothing(ys, x0, trans, obs, method, init_nominal=None, parallel=True,
                       criterion=default_criterion, return_loglikelihood=False):
    """parsmooth/methods.py:54-76."""
    if init_nominal is None:
        init_nominal = filter_smoother(ys, x0, trans, obs, method, None, parallel)
    nominal, _ = fixed_point(lambda n: filter_smoother(ys, x0, trans, obs, method, n, parallel),
                             init_nominal, criterion)
    if return_loglikelihood:
        _, ell = filtering(ys, x0, trans, obs, method, nominal, parallel, return_loglikelihood=True)
        return nominal, ell
    return nominal


# ----------------------------------------------------------------------------
# fixed point with implicit gradient (parsmooth/_utils.py:103-133)
# ----------------------------------------------------------------------------
def fixed_point_vjp(f_x_vjp, f_theta_vjp, x_star_bar, n_iter):
    """Restates _fixed_point_rev (parsmooth/_utils.py:118-125) and _rev_iter
    (:128-133): u starts at x_bar and is mapped u <- x_bar + J_x^T u while the
    criterion ``i < n_iter + 1`` holds (carry starts at (1, x_bar, g(x_bar))),
    i.e. the map is applied n_iter + 1 times; then theta_bar = (df/dtheta)^T u.
    """
    u, _ = fixed_point(lambda v: x_star_bar + f_x_vjp(v), x_star_bar,
                       lambda i, *_: i < n_iter + 1)
    return f_theta_vjp(u)


# ----------------------------------------------------------------------------
# brute-force joint Gaussian for linear models (closed-form pin)
# ----------------------------------------------------------------------------
def joint_gaussian_posterior(ys, m0, P0, F, Q, bt, H, R, c):
    """Exact posterior marginals by conditioning the dense joint Gaussian of
    (x_0..x_T, y_1..y_T) of a linear-Gaussian model.  Returns filtered means/covs
    (conditioning on y_1..y_k), smoothed means/covs and the log marginal
    likelihood log N(y; mu_y, Sigma_y)."""
    ys = np.asarray(ys, dtype=np.float64)
    T, ny = ys.shape
    nx = m0.shape[0]
    mu = [m0]
    for _ in range(T):
        mu.append(F @ mu[-1] + bt)
    mu = np.concatenate(mu)
    Phi = [np.eye(nx)]
    for _ in range(T):
        Phi.append(F @ Phi[-1])
    # cov(x_i, x_j) for j >= i: F^(j-i) Sigma_i
    Sig = [P0]
    for _ in range(T):
        Sig.append(F @ Sig[-1] @ F.T + Q)
    N = (T + 1) * nx
    Cxx = np.zeros((N, N))
    for i in range(T + 1):
        for j in range(i, T + 1):
            blk = Phi[j - i] @ Sig[i]
            Cxx[j * nx:(j + 1) * nx, i * nx:(i + 1) * nx] = blk
            Cxx[i * nx:(i + 1) * nx, j * nx:(j + 1) * nx] = blk.T
    Hbig = np.zeros((T * ny, N))
    for k in range(T):
        Hbig[k * ny:(k + 1) * ny, (k + 1) * nx:(k + 2) * nx] = H
    muy = Hbig @ mu + np.tile(c, T)
    Cyy = Hbig @ Cxx @ Hbig.T + np.kron(np.eye(T), R)
    Cxy = Cxx @ Hbig.T
    yv = ys.reshape(-1)

    def cond(kmax):
        if kmax == 0:
            return mu, Cxx
        idx = slice(0, kmax * ny)
        G = np.linalg.solve(Cyy[idx, idx], Cxy[:, idx].T).T
        return mu + G @ (yv[idx] - muy[idx]), Cxx - G @ Cxy[:, idx].T

    fm, fP = [], []
    for k in range(T + 1):
        mk, Pk = cond(k)
        fm.append(mk[k * nx:(k + 1) * nx])
        fP.append(Pk[k * nx:(k + 1) * nx, k * nx:(k + 1) * nx])
    ms, Ps = cond(T)
    sm = ms.reshape(T + 1, nx)
    sP = np.stack([Ps[k * nx:(k + 1) * nx, k * nx:(k + 1) * nx] for k in range(T + 1)])
    sign, logdet = np.linalg.slogdet(Cyy)
    r = yv - muy
    ell = -0.5 * r @ np.linalg.solve(Cyy, r) - 0.5 * logdet - 0.5 * T * ny * np.log(2 * np.pi)
    return (np.stack(fm), np.stack(fP)), (sm, sP), float(ell)


# ----------------------------------------------------------------------------
# input generators restated from the reference's harness
# ----------------------------------------------------------------------------
def get_system(dim_x, dim_y):
    """tests/_test_utils.py:6-22 (uses numpy's global RandomState)."""
    m = np.random.randn(dim_x)
    cholP = np.random.rand(dim_x, dim_x)
    cholP[np.triu_indices(dim_x, 1)] = 0.0
    P = cholP @ cholP.T
    cholR = np.random.rand(dim_y, dim_y)
    cholR[np.triu_indices(dim_y, 1)] = 0.0
    R = cholR @ cholR.T
    H = np.eye(dim_y, dim_x)
    c = np.random.randn(dim_y)
    y = np.random.randn(dim_y)
    return MVNStandard(m, P), MVNSqrt(m, cholP), H, R, cholR, c, y


def lgssm_data(x0, A, H, R, Q, b, c, T, random_state=None, chol_R=None):
    """tests/_lgssm.py:42-95 (float32 simulation)."""
    if random_state is None or isinstance(random_state, int):
        random_state = np.random.RandomState(random_state)
    normals = random_state.randn(T, Q.shape[0] + R.shape[0]).astype(np.float32)
    if chol_R is None:
        chol_R = np.linalg.cholesky(R)
    chol_Q = np.linalg.cholesky(Q)
    x = np.copy(x0).astype(np.float32)
    obs = np.empty((T, R.shape[0]), dtype=np.float32)
    states = np.empty((T + 1, Q.shape[0]), dtype=np.float32)
    states[0] = x
    for i in range(T):
        x = A @ x + chol_Q @ normals[i, :Q.shape[0]] + b
        states[i + 1] = x
        obs[i] = H @ x + chol_R @ normals[i, Q.shape[0]:] + c
    return states, obs
