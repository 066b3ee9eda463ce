"""Gradient of the iterated-smoother pseudo-log-likelihood with respect to model
parameters, on the HIP device.

The reference differentiates

    ell(theta) = filtering(ys, x0, models(theta), nominal=x*, return_loglikelihood=True)[1],
    x* = fixed_point(f(., theta), init, criterion),   f = filter_smoother

with the implicit-function rule of its fixed-point custom VJP
(parsmooth/_utils.py:118-133): with x_bar = d ell / d x*, the adjoint
u = sum_{j=0}^{N+1} (J_x^T)^j x_bar (N = number of forward applications of f,
the carry of ``_rev_iter`` runs while i < N + 1), and
theta_bar = (d f / d theta)^T u + d ell / d theta; the initial nominal gets no
gradient (:125).

Two evaluations of that same quantity are provided.

Reverse mode (``value_and_grad``, the default): the reference's reverse sum
itself, through the hand-written adjoint of one filter + smoother pass
(psm_adjoint_prepare / psm_filter_vjp / psm_filter_smoother_vjp):

    x_bar, theta_ell = d ell / d (x*, theta)           (psm_filter_vjp, Fisher's identity)
    u <- x_bar + J_x^T u, N + 1 times from u = x_bar  (psm_filter_smoother_vjp)
    theta_bar = (d f / d theta)^T u + theta_ell        (one more adjoint pass)

N + 3 adjoint passes at x* whatever the number of parameters; the primal at
x* is checkpointed once (filtered / smoothed covariances, gains, filter-element
matrices) and every adjoint pass reads it.

Forward mode (``value_and_grad(..., mode="forward")``): one parameter direction
e_p at a time, through the tangent (dual-number) instantiation of the same
kernels:

    v      = (d f / d theta) e_p                 (one tangent pass at x*, dx = 0)
    dx    <- v + J_x dx, N + 1 times from dx = v (tangent passes at x*)
    d ell  = (d ell / d x*) dx + (d ell / d theta) e_p   (one tangent filter pass)

which is the transpose of the reverse sum: P (N + 3) tangent passes.

The parameterisation theta -> model arrays is a small torch function
evaluated on the host; its Jacobian (torch.func.jacfwd) maps the model-array
cotangents (reverse) or supplies their tangents (forward).
"""
from __future__ import annotations

import ctypes
from typing import Callable, Dict, Sequence

import numpy as np
import torch

from . import _native as N
from ._engine import Problem, to_dev
from .linearization._methods import resolve

KEYS = ("trans_params", "trans_q_mean", "trans_q_chol", "obs_params", "obs_q_mean", "obs_q_chol",
        "m0", "L0")


def _dual(value: torch.Tensor, tangent: torch.Tensor) -> torch.Tensor:
    return torch.stack([value, tangent], dim=-1).contiguous()


class _ModelSet:
    """ctypes model structs (value or dual) for one evaluation."""

    def __init__(self, kinds, dims, arrays, dev, dtype):
        (tk, tn), (ok, on) = kinds
        nx, ny = dims
        self.keep = {k: to_dev(v, dtype, dev) if not isinstance(v, torch.Tensor) or v.device != dev
                     or v.dtype != dtype else v.contiguous() for k, v in arrays.items()}
        a = self.keep
        self.trans = N.PsmModel(tk, nx, nx, 0, a["trans_params"].data_ptr(), a["trans_q_mean"].data_ptr(),
                                a["trans_q_chol"].data_ptr())
        self.obs = N.PsmModel(ok, nx, ny, 0, a["obs_params"].data_ptr(), a["obs_q_mean"].data_ptr(),
                              a["obs_q_chol"].data_ptr())


class IteratedLoglikelihood:
    """Pseudo-log-likelihood of the iterated parallel smoother and its gradient
    for a ragged batch of sequences.

    Parameters
    ----------
    build: theta (torch float64 CPU tensor, shape (P,)) -> dict with the keys of
        ``KEYS``: model parameter vectors (see include/psm.h for each model
        kind), noise means and chol factors, prior means (n_seq, nx) and chols
        (n_seq, nx, nx).  Written with torch ops so that it can be
        differentiated.
    trans_kind, obs_kind: model ids (parsmooth._native.MODEL_*).
    nx, ny, lengths, observations (concatenated (n_obs, ny)).
    linearization_method, n_iter, init_nominal ((n_states, nx), (n_states, nx, nx)).
    dtype: torch.float64 or torch.float32 (value arithmetic; tangents use the
        same base type, log-likelihood terms accumulate in FP64).
    """

    def __init__(self, build: Callable, trans_kind: int, obs_kind: int, nx: int, ny: int,
                 lengths: Sequence[int], observations, linearization_method, n_iter: int,
                 init_nominal, dtype=torch.float64):
        self.build = build
        self.kinds = ((trans_kind, nx), (obs_kind, ny))
        self.nx, self.ny = nx, ny
        self.dev = N.device()
        self.dtype = dtype
        self.method = resolve(linearization_method)
        self.n_iter = int(n_iter)
        self.prob = Problem(nx, ny, lengths, dtype, True, self.dev)
        self.lengths = lengths
        self._tprob = None
        self.ys = to_dev(observations, dtype, self.dev).reshape(self.prob.n_obs, ny)
        self._ys_dual = None
        self.init = (to_dev(init_nominal[0], dtype, self.dev).reshape(self.prob.n_states, nx),
                     to_dev(init_nominal[1], dtype, self.dev).reshape(self.prob.n_states, nx, nx))

    @property
    def tprob(self):
        """The tangent problem of forward mode, made on first use: the tangent
        kernels exist for fewer sizes than the value kernels, and reverse mode
        does not need them."""
        if self._tprob is None:
            tcode = N.PSM_F64_TANGENT if self.dtype == torch.float64 else N.PSM_F32_TANGENT
            self._tprob = Problem(self.nx, self.ny, self.lengths, self.dtype, True, self.dev, tangent_code=tcode)
        return self._tprob

    @property
    def ys_dual(self):
        if self._ys_dual is None:
            self._ys_dual = _dual(self.ys, torch.zeros_like(self.ys))
        return self._ys_dual

    # -------------------------------------------------------------- passes
    def _sigma(self, tangent):
        # the sigma table is FP64 for value and tangent passes alike (constant:
        # no tangent)
        return self.method.sigma(self.nx, self.dtype, self.dev)

    def _filter_smoother(self, prob, models, sigma, ys, m0, L0, nom, out, filt):
        rc = N.lib().psm_filter_smoother(
            ctypes.byref(prob.struct), ctypes.byref(models.trans), ctypes.byref(models.obs),
            ctypes.byref(sigma) if sigma is not None else None, N.ptr(ys), N.ptr(m0), N.ptr(L0),
            N.ptr(nom[0]), N.ptr(nom[1]), N.ptr(filt[0]), N.ptr(filt[1]), N.ptr(out[0]), N.ptr(out[1]),
            N.ptr(prob.workspace), prob.ws_bytes, 0, N.stream_ptr())
        N.check(rc, "psm_filter_smoother")

    def _filter_ell(self, prob, models, sigma, ys, m0, L0, nom, filt, ell):
        rc = N.lib().psm_filter(
            ctypes.byref(prob.struct), ctypes.byref(models.trans), ctypes.byref(models.obs),
            ctypes.byref(sigma) if sigma is not None else None, N.ptr(ys), N.ptr(m0), N.ptr(L0),
            N.ptr(nom[0]), N.ptr(nom[1]), N.ptr(filt[0]), N.ptr(filt[1]), N.ptr(ell),
            N.ptr(prob.workspace), prob.ws_bytes, N.stream_ptr())
        N.check(rc, "psm_filter")

    def _new(self, tangent):
        p = self.tprob if tangent else self.prob
        shape_m = (p.n_states, self.nx) + ((2,) if tangent else ())
        shape_L = (p.n_states, self.nx, self.nx) + ((2,) if tangent else ())
        return (torch.empty(shape_m, dtype=self.dtype, device=self.dev),
                torch.empty(shape_L, dtype=self.dtype, device=self.dev))

    # -------------------------------------------------------------- public
    def value(self, theta: torch.Tensor):
        """(per-sequence ell (float64), x* mean, x* chol) at theta."""
        vals = self.build(theta.detach().to(torch.float64).cpu())
        models = _ModelSet(self.kinds, (self.nx, self.ny), vals, self.dev, self.dtype)
        sigma, _keep = self._sigma(False)
        m0, L0 = models.keep["m0"], models.keep["L0"]
        filt = self._new(False)
        cur = self.init
        for k in range(self.n_iter):
            out = self._new(False)
            self._filter_smoother(self.prob, models, sigma, self.ys, m0, L0, cur, out, filt)
            cur = out
        ell = torch.empty(self.prob.n_seq, dtype=torch.float64, device=self.dev)
        self._filter_ell(self.prob, models, sigma, self.ys, m0, L0, cur, filt, ell)
        return ell, cur[0], cur[1]

    def fixed_point_residual(self, theta: torch.Tensor) -> np.ndarray:
        """Per sequence, max |f(x*) - x*| over the means of its states: how far
        the n_iter-th iterate is from a fixed point.  The implicit gradient is the
        derivative of the converged log-likelihood only where this is small (an
        iteration caught in a cycle has no fixed point to differentiate)."""
        theta = theta.detach().to(torch.float64).cpu()
        _, xm, xL = self.value(theta)
        models = _ModelSet(self.kinds, (self.nx, self.ny), self.build(theta), self.dev, self.dtype)
        sigma, _keep = self._sigma(False)
        out, filt = self._new(False), self._new(False)
        self._filter_smoother(self.prob, models, sigma, self.ys, models.keep["m0"], models.keep["L0"],
                              (xm, xL), out, filt)
        d = (out[0] - xm).abs().amax(1).double().cpu().numpy()
        return np.maximum.reduceat(d, self.prob.state_offsets_host[:-1])

    def _dual_models(self, vals, tans):
        dual = {k: _dual(to_dev(vals[k], self.dtype, self.dev), to_dev(tans[k], self.dtype, self.dev))
                for k in KEYS}
        return _ModelSet(self.kinds, (self.nx, self.ny), dual, self.dev, self.dtype)

    def _tangent_pass(self, models, sigma, nominal, dnominal):
        """One filter-smoother pass at ``nominal`` carrying the tangents; returns
        the (value, tangent) trajectories as (n_states, nx[, nx], 2) tensors."""
        nom = (_dual(nominal[0], dnominal[0]), _dual(nominal[1], dnominal[1]))
        out = self._new(True)
        filt = self._new(True)
        m0, L0 = models.keep["m0"], models.keep["L0"]
        self._filter_smoother(self.tprob, models, sigma, self.ys_dual, m0, L0, nom, out, filt)
        return out

    def jvp_pass(self, theta, dtheta, nominal, dnominal):
        """Jacobian-vector product of one filter-smoother pass (the fixed-point
        map of iterated smoothing): returns ((mean, chol), (d mean, d chol)) for
        the parameter direction ``dtheta`` and nominal-trajectory direction
        ``dnominal`` (tensors or arrays of the trajectory shapes)."""
        theta = torch.as_tensor(theta, dtype=torch.float64).detach().cpu()
        dtheta = torch.as_tensor(dtheta, dtype=torch.float64).detach().cpu()
        vals, tans = torch.func.jvp(self.build, (theta,), (dtheta,))
        models = self._dual_models(vals, tans)
        sigma, _keep = self._sigma(True)
        nom = tuple(to_dev(a, self.dtype, self.dev) for a in nominal)
        dnom = tuple(to_dev(a, self.dtype, self.dev) for a in dnominal)
        out = self._tangent_pass(models, sigma, nom, dnom)
        return ((out[0][..., 0], out[1][..., 0]), (out[0][..., 1], out[1][..., 1]))

    def value_and_grad(self, theta: torch.Tensor, mode: str = "reverse"):
        """(ell (n_seq,), grad (n_seq, P)) with the reference's implicit gradient,
        by the reverse-mode adjoint (default) or by forward-mode tangents
        (``mode="forward"``)."""
        if mode == "reverse":
            return self._value_and_grad_reverse(theta)
        if mode != "forward":
            raise ValueError(f"mode must be 'reverse' or 'forward', not {mode!r}")
        theta = theta.detach().to(torch.float64).cpu()
        ell, xm, xL = self.value(theta)
        vals = self.build(theta)
        jac = torch.func.jacfwd(self.build)(theta)
        P = theta.numel()
        sigma, _keep = self._sigma(True)
        grads = torch.empty(self.prob.n_seq, P, dtype=torch.float64, device=self.dev)
        filt = self._new(True)
        for p in range(P):
            models = self._dual_models(vals, {k: jac[k][..., p] for k in KEYS})
            m0, L0 = models.keep["m0"], models.keep["L0"]
            # dx <- (d f / d theta) e_p + J_x dx, N + 2 passes from dx = 0
            dm, dL = torch.zeros_like(xm), torch.zeros_like(xL)
            for _ in range(self.n_iter + 2):
                out = self._tangent_pass(models, sigma, (xm, xL), (dm, dL))
                dm, dL = out[0][..., 1].contiguous(), out[1][..., 1].contiguous()
            ell_d = torch.empty(self.prob.n_seq, 2, dtype=torch.float64, device=self.dev)
            self._filter_ell(self.tprob, models, sigma, self.ys_dual, m0, L0,
                             (_dual(xm, dm), _dual(xL, dL)), filt, ell_d)
            grads[:, p] = ell_d[:, 1]
        return ell, grads


    # -------------------------------------------------------------- reverse mode
    def _adjoint_problem(self):
        """The problem the adjoint passes run on: this one in FP64; for FP32 an
        FP64 twin over the same FP32-rounded observations and model arrays.
        The adjoint is FP64 either way, but its checkpoint (filtered and
        smoothed moments at x*) comes from the primal pass of the problem it is
        given, and FP32 smoothed means carry an absolute rounding of ~1e-7 of
        the track's scale into every step's residual (y - c - H m): summed over
        16,384 steps that moves the sensor-position gradients by ~5e-3 against
        central differences, where the FP64 checkpoint of the same FP32 fixed
        point stays within 1e-5."""
        if self.dtype == torch.float64:
            return self.prob, self.ys
        twin = getattr(self, "_twin", None)
        if twin is None:
            p = self.prob
            twin = (Problem(self.nx, self.ny, p.lengths, torch.float64, True, self.dev), self.ys.double())
            self._twin = twin
        return twin

    def _adjoint_models(self, vals):
        """Model arrays of the adjoint problem: the FP32-rounded arrays of the
        primal, in FP64, for an FP32 objective."""
        if self.dtype == torch.float64:
            return _ModelSet(self.kinds, (self.nx, self.ny), vals, self.dev, torch.float64)
        rounded = {k: to_dev(v, self.dtype, self.dev).double() for k, v in vals.items()}
        return _ModelSet(self.kinds, (self.nx, self.ny), rounded, self.dev, torch.float64)

    def _adjoint_workspace(self, prob):
        nbytes = int(N.lib().psm_adjoint_workspace_bytes(ctypes.byref(prob.struct)))
        if nbytes == 0:
            N.check(-1, "psm_adjoint_workspace_bytes")
        ws = getattr(self, "_aws", None)
        if ws is None or ws.numel() < nbytes:
            ws = torch.empty(nbytes, dtype=torch.uint8, device=self.dev)
            self._aws = ws
        return ws

    def _model_cotangents(self, th: torch.Tensor, vals, jac) -> torch.Tensor:
        """Per-sequence parameter gradients from the adjoint's model-array
        cotangents th (n_seq, NTH; layout of PSM_ADJ_NTH in include/psm.h)
        and the Jacobian of the parameterisation."""
        nx, ny = self.nx, self.ny
        n_seq = th.shape[0]
        P = next(iter(jac.values())).shape[-1]
        sizes = (("trans_params", nx * nx), ("trans_q_mean", nx), ("trans_q_chol", nx * nx),
                 ("obs_params", ny * nx), ("obs_q_mean", ny), ("obs_q_chol", ny * ny), ("m0", nx), ("L0", nx * nx))
        off, o = {}, 0
        for key, k in sizes:
            off[key] = o
            o += k
        th = th.double().cpu()
        grad = torch.zeros(n_seq, P, dtype=torch.float64)
        for key in KEYS:
            v = torch.as_tensor(vals[key])
            J = jac[key].to(torch.float64)
            per_seq = key in ("m0", "L0") and v.dim() == (2 if key == "m0" else 3)
            k = (v[0] if per_seq else v).numel()
            cot = th[:, off[key]:off[key] + k]                      # (n_seq, k)
            if per_seq:
                grad += torch.einsum("sk,skp->sp", cot, J.reshape(n_seq, k, P))
            else:
                grad += cot @ J.reshape(k, P)
        return grad

    def _value_and_grad_reverse(self, theta: torch.Tensor):
        theta = theta.detach().to(torch.float64).cpu()
        ell, xm, xL = self.value(theta)
        vals = self.build(theta)
        jac = torch.func.jacfwd(self.build)(theta)
        p, ys = self._adjoint_problem()
        models = self._adjoint_models(vals)
        sigma, _keep = self.method.sigma(self.nx, torch.float64, self.dev)
        sref = ctypes.byref(sigma) if sigma is not None else None
        m0, L0 = models.keep["m0"], models.keep["L0"]
        aws = self._adjoint_workspace(p)
        f64 = dict(dtype=torch.float64, device=self.dev)
        filt = (torch.empty((p.n_states, self.nx), **f64), torch.empty((p.n_states, self.nx, self.nx), **f64))
        sm = (torch.empty_like(filt[0]), torch.empty_like(filt[1]))
        xm, xL = xm.double().contiguous(), xL.double().contiguous()
        base = (ctypes.byref(p.struct), ctypes.byref(models.trans), ctypes.byref(models.obs), sref,
                N.ptr(ys), N.ptr(m0), N.ptr(L0), N.ptr(xm), N.ptr(xL))
        rc = N.lib().psm_adjoint_prepare(*base, N.ptr(filt[0]), N.ptr(filt[1]), N.ptr(sm[0]), N.ptr(sm[1]),
                                         N.ptr(p.workspace), p.ws_bytes, N.ptr(aws), aws.numel(), N.stream_ptr())
        N.check(rc, "psm_adjoint_prepare")
        M, nx, ny = p.n_states, self.nx, self.ny
        nth = 3 * nx * nx + 2 * nx + ny * nx + ny + ny * ny
        xb = (torch.empty((M, nx), **f64), torch.empty((M, nx, nx), **f64))
        th_ell = torch.empty((p.n_seq, nth), **f64)
        rc = N.lib().psm_filter_vjp(*base, N.ptr(sm[0]), N.ptr(sm[1]), N.ptr(xb[0]), N.ptr(xb[1]), N.ptr(th_ell),
                                    N.ptr(aws), aws.numel(), N.stream_ptr())
        N.check(rc, "psm_filter_vjp")
        u = (xb[0].clone(), xb[1].clone())
        nxt = (torch.empty_like(u[0]), torch.empty_like(u[1]))

        def vjp(cot, add, out, th):
            rc_ = N.lib().psm_filter_smoother_vjp(*base, N.ptr(sm[0]), N.ptr(sm[1]), N.ptr(cot[0]), N.ptr(cot[1]),
                                                  N.ptr(add[0] if add else None), N.ptr(add[1] if add else None),
                                                  N.ptr(out[0]), N.ptr(out[1]), N.ptr(th), N.ptr(aws),
                                                  aws.numel(), N.stream_ptr())
            N.check(rc_, "psm_filter_smoother_vjp")

        # u <- x_bar + J_x^T u, N + 1 times (the reference's _rev_iter loop)
        for _ in range(self.n_iter + 1):
            vjp(u, xb, nxt, None)
            u, nxt = nxt, u
        th_f = torch.empty((p.n_seq, nth), **f64)
        vjp(u, None, nxt, th_f)
        return ell, self._model_cotangents(th_ell + th_f, vals, jac).to(self.dev)

    def vjp_pass(self, theta, nominal, cot_mean, cot_chol):
        """Vector-Jacobian product of one filter-smoother pass at ``nominal``:
        for cotangents of its smoothed (mean, chol), returns the cotangents of
        the nominal (mean, chol) and the per-sequence parameter gradient
        (n_seq, P)."""
        theta = torch.as_tensor(theta, dtype=torch.float64).detach().cpu()
        vals = self.build(theta)
        jac = torch.func.jacfwd(self.build)(theta)
        p, ys = self._adjoint_problem()
        models = self._adjoint_models(vals)
        sigma, _keep = self.method.sigma(self.nx, torch.float64, self.dev)
        sref = ctypes.byref(sigma) if sigma is not None else None
        aws = self._adjoint_workspace(p)
        nm = tuple(to_dev(a, self.dtype, self.dev).double().contiguous() for a in nominal)
        f64 = dict(dtype=torch.float64, device=self.dev)
        filt = (torch.empty((p.n_states, self.nx), **f64), torch.empty((p.n_states, self.nx, self.nx), **f64))
        sm = (torch.empty_like(filt[0]), torch.empty_like(filt[1]))
        base = (ctypes.byref(p.struct), ctypes.byref(models.trans), ctypes.byref(models.obs), sref,
                N.ptr(ys), N.ptr(models.keep["m0"]), N.ptr(models.keep["L0"]), N.ptr(nm[0]), N.ptr(nm[1]))
        rc = N.lib().psm_adjoint_prepare(*base, N.ptr(filt[0]), N.ptr(filt[1]), N.ptr(sm[0]), N.ptr(sm[1]),
                                         N.ptr(p.workspace), p.ws_bytes, N.ptr(aws), aws.numel(), N.stream_ptr())
        N.check(rc, "psm_adjoint_prepare")
        cot = (to_dev(cot_mean, torch.float64, self.dev).contiguous(),
               to_dev(cot_chol, torch.float64, self.dev).contiguous())
        out = (torch.empty((p.n_states, self.nx), **f64), torch.empty((p.n_states, self.nx, self.nx), **f64))
        nth = 3 * self.nx ** 2 + 2 * self.nx + self.ny * self.nx + self.ny + self.ny ** 2
        th = torch.empty((p.n_seq, nth), **f64)
        rc = N.lib().psm_filter_smoother_vjp(*base, N.ptr(sm[0]), N.ptr(sm[1]), N.ptr(cot[0]), N.ptr(cot[1]),
                                             None, None, N.ptr(out[0]), N.ptr(out[1]), N.ptr(th), N.ptr(aws),
                                             aws.numel(), N.stream_ptr())
        N.check(rc, "psm_filter_smoother_vjp")
        return (out[0], out[1]), self._model_cotangents(th, vals, jac)


def bearings_parameterisation(dt: float, m0, L0):
    """theta -> model arrays for the bearings-only coordinated-turn model with
    theta = (log qc, log qw, log sigma, s1x, s1y, s2x, s2y): chol Q from the
    white-noise-acceleration covariance (notebooks/bearing_data.py:93-134), R =
    sigma^2 I, sensors at (s1x, s1y), (s2x, s2y).  m0 (n_seq, 5), L0 (n_seq, 5, 5)
    are fixed priors."""
    m0 = torch.as_tensor(np.asarray(m0, dtype=np.float64))
    L0 = torch.as_tensor(np.asarray(L0, dtype=np.float64))

    def build(theta):
        qc, qw, sig = torch.exp(theta[0]), torch.exp(theta[1]), torch.exp(theta[2])
        z = torch.zeros((), dtype=theta.dtype)
        a, b, c = qc * dt ** 3 / 3, qc * dt ** 2 / 2, qc * dt
        Q = torch.stack([torch.stack([a, z, b, z, z]), torch.stack([z, a, z, b, z]),
                         torch.stack([b, z, c, z, z]), torch.stack([z, b, z, c, z]),
                         torch.stack([z, z, z, z, dt * qw])])
        return {"trans_params": torch.tensor([dt], dtype=theta.dtype),
                "trans_q_mean": torch.zeros(5, dtype=theta.dtype), "trans_q_chol": torch.linalg.cholesky(Q),
                "obs_params": theta[3:7], "obs_q_mean": torch.zeros(2, dtype=theta.dtype),
                "obs_q_chol": sig * torch.eye(2, dtype=theta.dtype), "m0": m0.to(theta.dtype),
                "L0": L0.to(theta.dtype)}
    return build


class _LoglikFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, obj):
        ell, grad = obj.value_and_grad(theta)
        ctx.save_for_backward(grad.sum(0).to(theta.device, theta.dtype))
        return ell.sum().to(theta.device, theta.dtype)

    @staticmethod
    def backward(ctx, g):
        (grad,) = ctx.saved_tensors
        return g * grad, None


def loglikelihood(theta: torch.Tensor, obj: IteratedLoglikelihood) -> torch.Tensor:
    """Sum over sequences of the pseudo-log-likelihood, differentiable in theta
    through torch.autograd (the backward returns the implicit gradient)."""
    return _LoglikFn.apply(theta, obj)
