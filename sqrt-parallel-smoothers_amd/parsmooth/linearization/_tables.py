"""Unit sigma-point tables (points xi_p and weights) of the sigma-point
linearizations.  Points in state space are m + L xi_p.

* cubature: parsmooth/linearization/_cubature.py:63-85 (2n points +-sqrt(n) e_i,
  weights 1/(2n));
* Gauss-Hermite: parsmooth/linearization/_gh.py:73-126 (tensor grid of the 1-D
  rule with order^n points, the first coordinate varying fastest).  The 1-D
  nodes and weights come from the Golub-Welsch eigenvalue problem of the
  Hermite (probabilists') Jacobi matrix, which gives the same rule as the
  reference's polynomial-root formula;
* unscented: parsmooth/linearization/_unscented.py:42-101 (2n+1 points, default
  alpha=1, beta=0, kappa=3+n).
"""
from __future__ import annotations

import functools

import numpy as np


@functools.lru_cache(maxsize=None)
def cubature_table(n: int):
    w = np.full(2 * n, 1.0 / (2 * n))
    xi = np.concatenate([np.eye(n), -np.eye(n)], axis=0) * np.sqrt(n)
    return xi, w, w.copy()


@functools.lru_cache(maxsize=None)
def _gh_1d(order: int):
    # Jacobi matrix of the monic probabilists' Hermite recurrence
    # He_{k+1} = x He_k - k He_{k-1}: off-diagonal sqrt(k).
    k = np.arange(1, order)
    J = np.diag(np.sqrt(k), 1) + np.diag(np.sqrt(k), -1)
    nodes, vecs = np.linalg.eigh(J)
    weights = vecs[0, :] ** 2          # standard normal weights (sum to 1)
    return nodes, weights


@functools.lru_cache(maxsize=None)
def gauss_hermite_table(n: int, order: int = 3):
    if order < 1:
        raise ValueError("order must be >= 1")
    nodes, weights = _gh_1d(order)
    idx = np.array(np.unravel_index(np.arange(order ** n), (order,) * n, order="F"))
    xi = nodes[idx].T
    w = np.prod(weights[idx], axis=0)
    return xi, w, w.copy()


@functools.lru_cache(maxsize=None)
def unscented_table(n: int, alpha: float = 1.0, beta: float = 0.0, kappa=None):
    if kappa is None:
        kappa = 3.0 + n
    lam = alpha ** 2 * (n + kappa) - n
    wm = np.full(2 * n + 1, 1.0 / (2 * (n + lam)))
    wm[0] = lam / (n + lam)
    wc = wm.copy()
    wc[0] = lam / (n + lam) + (1 - alpha ** 2 + beta)
    s = np.sqrt(n + lam)
    xi = np.concatenate([np.zeros((1, n)), s * np.eye(n), -s * np.eye(n)], axis=0)
    return xi, wm, wc
