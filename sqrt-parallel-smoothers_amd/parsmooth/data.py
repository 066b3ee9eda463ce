"""Synthetic inputs of the reference's experiments (numpy, seeded).

* ``bearings_data``: coordinated-turn target seen by two bearing sensors,
  restating notebooks/bearing_data.py:137-198 (float32 simulation; the turn rate
  follows 1 + q dt cumsum(randn); the state is propagated with the matrix
  exponential of the 4x4 turn generator).
* ``lgssm_data``: linear-Gaussian simulation, tests/_lgssm.py:42-95.
* ``bearings_data_batch``: many such targets at once (uneven lengths), with the
  turn propagated by the closed form of the same matrix exponential
  (the velocity rotates by -a dt and the position integrates it).
* ``cv_model``: the nx=4 constant-velocity tracking model (white acceleration).
"""
from __future__ import annotations

import numpy as np
import scipy.linalg as linalg


def bearings_data(x0, dt, r, T, s1, s2, q=10.0, random_state=None):
    """Returns (ts, true_states (T+1, 5), observations (T, 2)), all float32."""
    if random_state is None or isinstance(random_state, (int, np.integer)):
        random_state = np.random.RandomState(random_state)
    a_s = (1 + q * dt * np.cumsum(random_state.randn(T))).astype(np.float32)
    s1 = np.asarray(s1, dtype=np.float32)
    s2 = np.asarray(s2, dtype=np.float32)
    x = np.copy(x0).astype(np.float32)
    observations = np.empty((T, 2), dtype=np.float32)
    true_states = np.zeros((T + 1, 5), dtype=np.float32)
    ts = np.linspace(dt, (T + 1) * dt, T).astype(np.float32)
    true_states[0, :4] = x
    normals = random_state.randn(T, 2).astype(np.float32)
    for i, a in enumerate(a_s):
        G = np.array([[0, 0, 1, 0], [0, 0, 0, 1], [0, 0, 0, a], [0, 0, -a, 0]], dtype=np.float32)
        x = linalg.expm(G * dt) @ x
        y1 = np.arctan2(x[1] - s1[1], x[0] - s1[0]) + r * normals[i, 0]
        y2 = np.arctan2(x[1] - s2[1], x[0] - s2[0]) + r * normals[i, 1]
        observations[i] = [y1, y2]
        true_states[i + 1] = np.concatenate((x, np.array([a])))
    return ts, true_states, observations


def population_data(x0, T, Q, lam, random_state=None):
    """Ricker-map states and Poisson counts (notebooks/population_model.py:172-208):
    x_{k+1} = log(44.7) + x_k - exp(x_k) + chol(Q) eps_k, y_k ~ Poisson(lam exp(x_{k+1})).
    Returns (true_states (T+1, 1) with x0 first, observations (T, 1) as float).
    The reference draws from JAX keys; this draws from numpy's
    RandomState(random_state), so the same seed gives other samples."""
    if random_state is None or isinstance(random_state, (int, np.integer)):
        random_state = np.random.RandomState(random_state)
    x = np.asarray(x0, dtype=np.float64).reshape(-1)
    cQ = np.linalg.cholesky(np.asarray(Q, dtype=np.float64).reshape(x.size, x.size))
    noises = random_state.randn(T, x.size) @ cQ.T
    states = np.empty((T + 1, x.size))
    obs = np.empty((T, x.size))
    states[0] = x
    for k in range(T):
        x = np.log(44.7) + x - np.exp(x) + noises[k]
        states[k + 1] = x
        obs[k] = random_state.poisson(lam * np.exp(x)).astype(np.float64)
    return states, obs


def lgssm_data(x0, A, H, R, Q, b, c, T, random_state=None, chol_R=None):
    """Returns (true_states (T+1, nx), observations (T, ny)), float32."""
    if random_state is None or isinstance(random_state, (int, np.integer)):
        random_state = np.random.RandomState(random_state)
    nq, nr = Q.shape[0], R.shape[0]
    normals = random_state.randn(T, nq + nr).astype(np.float32)
    if chol_R is None:
        chol_R = np.linalg.cholesky(R)
    chol_Q = np.linalg.cholesky(Q)
    x = np.copy(x0).astype(np.float32)
    obs = np.empty((T, nr), dtype=np.float32)
    states = np.empty((T + 1, nq), dtype=np.float32)
    states[0] = x
    for i in range(T):
        x = A @ x + chol_Q @ normals[i, :nq] + b
        states[i + 1] = x
        obs[i] = H @ x + chol_R @ normals[i, nq:] + c
    return states, obs


def cv_model(dt=0.1, q=1.0, r=0.5):
    """Constant-velocity model on (x, y, vx, vy): F, Q (white acceleration of
    intensity q), H = [I 0], R = r^2 I."""
    F = np.array([[1, 0, dt, 0], [0, 1, 0, dt], [0, 0, 1, 0], [0, 0, 0, 1]], dtype=np.float64)
    Q = q * np.array([[dt ** 3 / 3, 0, dt ** 2 / 2, 0], [0, dt ** 3 / 3, 0, dt ** 2 / 2],
                      [dt ** 2 / 2, 0, dt, 0], [0, dt ** 2 / 2, 0, dt]])
    H = np.eye(2, 4)
    R = r ** 2 * np.eye(2)
    return F, Q, H, R


def bearings_data_batch(lengths, x0s, dt, r, s1, s2, q=10.0, seed=0):
    """Simulate len(lengths) bearings-only targets; returns (true_states list,
    observations list) with float32 arrays of shapes (T_s+1, 5) and (T_s, 2).
    x0s: (n_seq, 4) initial positions/velocities."""
    rng = np.random.default_rng(seed)
    lengths = np.asarray(lengths, dtype=np.int64)
    B, Tm = lengths.size, int(lengths.max())
    a = (1 + q * dt * np.cumsum(rng.standard_normal((B, Tm)), axis=1)).astype(np.float32)
    noise = rng.standard_normal((B, Tm, 2)).astype(np.float32)
    x = np.asarray(x0s, dtype=np.float64).reshape(B, 4).copy()
    states = np.zeros((B, Tm + 1, 5), dtype=np.float32)
    obs = np.zeros((B, Tm, 2), dtype=np.float32)
    states[:, 0, :4] = x
    s1 = np.asarray(s1, dtype=np.float64)
    s2 = np.asarray(s2, dtype=np.float64)
    for k in range(Tm):
        w = a[:, k].astype(np.float64)
        c, s = np.cos(w * dt), np.sin(w * dt)
        sw, cw = s / w, (1.0 - c) / w
        px, py, vx, vy = x[:, 0], x[:, 1], x[:, 2], x[:, 3]
        x = np.stack([px + sw * vx + cw * vy, py - cw * vx + sw * vy,
                      c * vx + s * vy, -s * vx + c * vy], axis=1)
        obs[:, k, 0] = np.arctan2(x[:, 1] - s1[1], x[:, 0] - s1[0]) + r * noise[:, k, 0]
        obs[:, k, 1] = np.arctan2(x[:, 1] - s2[1], x[:, 0] - s2[0]) + r * noise[:, k, 1]
        states[:, k + 1, :4] = x
        states[:, k + 1, 4] = w
    return ([states[b, :lengths[b] + 1] for b in range(B)], [obs[b, :lengths[b]] for b in range(B)])
