"""The helpers of the reference's parsmooth/_utils.py: tria, the rank-one
Cholesky updates, trajectory shifts, fixed_point with the implicit-function
gradient, and mvn_loglikelihood.

tria, the Cholesky updates and mvn_loglikelihood run in the HIP library
(psm_tria, psm_cholesky_update, psm_mvn_loglikelihood: one lane per batch item)
on float32 / float64 tensors, which may carry leading batch dimensions; numpy
inputs are moved to the device as float64.  none_or_shift / none_or_concat are
indexing, and fixed_point is host control flow around the caller's function,
differentiated by torch.autograd as the reference's custom VJP prescribes.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _native as N


def _dev(x, like=None):
    dev = N.device()
    if isinstance(x, torch.Tensor):
        dt = x.dtype if x.dtype in (torch.float32, torch.float64) else torch.float64
        return x.to(device=dev, dtype=dt)
    dt = like.dtype if isinstance(like, torch.Tensor) and like.dtype in (torch.float32, torch.float64) \
        else torch.float64
    return torch.as_tensor(np.asarray(x, dtype=np.float64), device=dev).to(dt)


def tria(A):
    """Lower-triangular L (rows x min(rows, cols)) with L L^T = A A^T
    (parsmooth/_utils.py:22-24).  The reference takes L from a QR decomposition
    and keeps its signs; this L has a non-negative diagonal, so the two agree up
    to the signs of L's columns (and exactly in L L^T)."""
    A = _dev(A)
    if A.dim() < 2:
        raise ValueError("tria needs a matrix")
    rows, cols = A.shape[-2], A.shape[-1]
    batch = A.shape[:-2]
    work = A.reshape(-1, rows, cols).contiguous().clone()
    nb = work.shape[0]
    N.check(N.lib().psm_tria(N.dtype_code(work.dtype), nb, rows, cols, N.ptr(work), N.stream_ptr()),
            "psm_tria")
    return work[..., :min(rows, cols)].reshape(*batch, rows, min(rows, cols))


def cholesky_update_many(chol_init, update_vectors, multiplier):
    """Factor of L L^T + multiplier sum_k v_k v_k^T, one vector after another
    (parsmooth/_utils.py:13-20).  chol_init (..., n, n) lower, update_vectors
    (..., k, n); the batch dimensions broadcast."""
    L = _dev(chol_init)
    V = _dev(update_vectors, like=L).to(L.dtype)
    if L.dim() < 2 or L.shape[-1] != L.shape[-2]:
        raise ValueError("chol_init must be square")
    n = L.shape[-1]
    if V.dim() < 2 or V.shape[-1] != n:
        raise ValueError("update_vectors must be (..., k, n)")
    nv = V.shape[-2]
    batch = torch.broadcast_shapes(L.shape[:-2], V.shape[:-2])
    out = L.expand(*batch, n, n).reshape(-1, n, n).contiguous().clone()
    Vb = V.expand(*batch, nv, n).reshape(-1, nv, n).contiguous()
    nb = out.shape[0]
    work = torch.empty(nb, n, dtype=out.dtype, device=out.device)
    N.check(N.lib().psm_cholesky_update(N.dtype_code(out.dtype), nb, n, nv, N.ptr(out), N.ptr(Vb),
                                        float(multiplier), N.ptr(work), N.stream_ptr()),
            "psm_cholesky_update")
    return out.reshape(*batch, n, n)


def _cholesky_update(chol, update_vector, multiplier=1.0):
    """Factor of L L^T + multiplier v v^T (parsmooth/_utils.py:39-81): entries
    above the diagonal are zero and non-finite entries are set to 0."""
    v = _dev(update_vector, like=chol if isinstance(chol, torch.Tensor) else None)
    return cholesky_update_many(chol, v.unsqueeze(-2), multiplier)


def mvn_loglikelihood(x, chol_cov):
    """log N(x; 0, chol_cov chol_cov^T) (parsmooth/_utils.py:152-160), over the
    broadcast batch dimensions of x (..., n) and chol_cov (..., n, n); computed
    and returned in float64."""
    L = _dev(chol_cov)
    x = _dev(x, like=L).to(L.dtype)
    n = L.shape[-1]
    if L.dim() < 2 or L.shape[-2] != n or x.shape[-1] != n:
        raise ValueError("x (..., n) and chol_cov (..., n, n) do not match")
    batch = torch.broadcast_shapes(x.shape[:-1], L.shape[:-2])
    xb = x.expand(*batch, n).reshape(-1, n).contiguous()
    Lb = L.expand(*batch, n, n).reshape(-1, n, n).contiguous()
    nb = xb.shape[0]
    out = torch.empty(nb, dtype=torch.float64, device=L.device)
    work = torch.empty(nb, n, dtype=torch.float64, device=L.device)
    N.check(N.lib().psm_mvn_loglikelihood(N.dtype_code(L.dtype), nb, n, N.ptr(xb), N.ptr(Lb), N.ptr(out),
                                          N.ptr(work), N.stream_ptr()), "psm_mvn_loglikelihood")
    return out.reshape(batch)


# ---------------------------------------------------------------- trajectories
def _tree_map(fn, *trees):
    first = trees[0]
    if isinstance(first, tuple):
        vals = [fn(*leaves) for leaves in zip(*trees)]
        return type(first)(*vals) if hasattr(first, "_fields") else tuple(vals)
    return fn(*trees)


def none_or_shift(x, shift):
    """x[shift:] for shift > 0, else x[:shift], over every field of a trajectory
    (parsmooth/_utils.py:84-89); None stays None."""
    if x is None:
        return None
    if shift > 0:
        return _tree_map(lambda z: z[shift:], x)
    return _tree_map(lambda z: z[:shift], x)


def _cat(parts):
    if all(isinstance(p, torch.Tensor) for p in parts):
        dev, dt = parts[0].device, parts[0].dtype
        return torch.cat([p.to(device=dev, dtype=dt) for p in parts])
    return np.concatenate([np.asarray(p) for p in parts])


def none_or_concat(x, y, position=1):
    """y (one state) prepended to the trajectory x for position 1, appended
    otherwise (parsmooth/_utils.py:92-98); None if either is None."""
    if x is None or y is None:
        return None
    if position == 1:
        return _tree_map(lambda b, a: _cat([a[None, ...], b]), x, y)
    return _tree_map(lambda b, a: _cat([b, a[None, ...]]), x, y)


# ---------------------------------------------------------------- fixed point
def _leaves(x):
    if isinstance(x, torch.Tensor):
        return [x], lambda ls: ls[0]
    if isinstance(x, tuple):
        subs = [_leaves(v) for v in x]
        counts = [len(s[0]) for s in subs]

        def rebuild(ls):
            vals, k = [], 0
            for (_, rb), c in zip(subs, counts):
                vals.append(rb(ls[k:k + c]))
                k += c
            return type(x)(*vals) if hasattr(x, "_fields") else tuple(vals)

        return [leaf for s in subs for leaf in s[0]], rebuild
    raise TypeError(f"fixed_point carries tensors or tuples of tensors, not {type(x)}")


def fixed_point(f, x0, criterion):
    """x_{i+1} = f(x_i) from x0 while criterion(i, x_prev, x) holds, with f
    applied once before the first test (i = 1), as parsmooth/_utils.py:103-149:
    with criterion i < n, f runs n times.

    Gradient (the reference's custom VJP): for the cotangent x_bar of the
    result, u = x_bar + J_x^T u is iterated n_iter + 1 times from u = x_bar at
    the fixed point, and the parameters that f closes over receive
    (d f / d theta)^T u; x0 receives zero.  x is a tensor or a (named) tuple
    of tensors."""
    with torch.no_grad():
        x_prev, x = x0, f(x0)
        i = 1
        while criterion(i, x_prev, x):
            x_prev, x = x, f(x)
            i += 1
    n_iter = i
    if not torch.is_grad_enabled():
        return x
    leaves, rebuild = _leaves(x)
    xs = [leaf.detach().requires_grad_(True) for leaf in leaves]
    fx = f(rebuild(xs))
    f_leaves, _ = _leaves(fx)
    if not any(leaf.requires_grad for leaf in f_leaves):
        return x
    flat = torch.cat([leaf.reshape(-1) for leaf in f_leaves])
    star = torch.cat([leaf.detach().reshape(-1) for leaf in leaves])

    def implicit(x_bar):
        if x_bar is None:  # an undefined cotangent stays undefined
            return None
        u = x_bar
        outs = [leaf for leaf in f_leaves if leaf.requires_grad]
        for _ in range(n_iter + 1):
            parts, k = [], 0
            for leaf in f_leaves:
                parts.append(u[k:k + leaf.numel()].reshape(leaf.shape))
                k += leaf.numel()
            gouts = [p for p, leaf in zip(parts, f_leaves) if leaf.requires_grad]
            jt = torch.autograd.grad(outs, xs, gouts, retain_graph=True, allow_unused=True)
            jt = torch.cat([(g if g is not None else torch.zeros_like(xx)).reshape(-1) for g, xx in zip(jt, xs)])
            u = x_bar + jt
        return u

    flat.register_hook(implicit)
    out = flat - flat.detach() + star
    parts, k = [], 0
    for leaf in leaves:
        parts.append(out[k:k + leaf.numel()].reshape(leaf.shape))
        k += leaf.numel()
    return rebuild(parts)


def _step_arrays(*arrays):
    """Device tensors of one common dtype for the single-step helpers: float32
    when every tensor argument is float32, float64 otherwise."""
    tens = [a for a in arrays if isinstance(a, torch.Tensor)]
    dt = torch.float32 if tens and all(t.dtype == torch.float32 for t in tens) and \
        len(tens) == len(arrays) else torch.float64
    dev = N.device()
    return [a.to(device=dev, dtype=dt) if isinstance(a, torch.Tensor)
            else torch.as_tensor(np.asarray(a, dtype=np.float64), device=dev).to(dt) for a in arrays]
