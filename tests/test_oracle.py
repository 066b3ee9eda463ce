"""Pins the FP64 oracle (oracle/psm_oracle.py) against the reference's own
fixtures and against what the mathematics fixes.  CPU only."""
import numpy as np
import pytest

from oracle import psm_oracle as O


def _rand_lower(n, rng=np.random):
    L = rng.rand(n, n)
    L[np.triu_indices(n, 1)] = 0.0
    return L


# ---------------------------------------------------------------- tria
@pytest.mark.parametrize("shape", [(1, 1), (2, 4), (5, 10), (7, 7), (10, 10), (3, 8)])
def test_tria_is_lower_cholesky_of_gram(shape):
    rng = np.random.RandomState(0)
    A = rng.randn(*shape)
    L = O.tria(A)
    assert np.allclose(np.triu(L, 1), 0.0)
    assert np.all(np.diag(L) >= 0)
    np.testing.assert_allclose(L @ L.T, A @ A.T, atol=1e-12)
    np.testing.assert_allclose(L, np.linalg.cholesky(A @ A.T), atol=1e-10)


def test_tria_rank_deficient_gives_zeros_not_nan():
    A = np.zeros((4, 8))
    A[0, 0] = 2.0
    L = O.tria(A)
    assert np.all(np.isfinite(L))
    np.testing.assert_allclose(L @ L.T, A @ A.T, atol=1e-14)


# ---------------------------------------------------------------- cholesky update
@pytest.mark.parametrize("multiplier", [1.0, -0.1])
@pytest.mark.parametrize("seed", [0, 42, 666])
@pytest.mark.parametrize("dim", [2, 3, 10, 11])
def test_cholesky_update_definition(multiplier, seed, dim):
    """tests/test_math_utils.py:16-34 (checked against the definition)."""
    np.random.seed(seed)
    L = _rand_lower(dim)
    v = np.random.randn(dim)
    expected = L @ L.T + multiplier * np.outer(v, v)
    if np.linalg.eigvalsh(expected).min() <= 1e-6:
        pytest.skip("not positive definite")
    out = O.cholesky_update(L, v, multiplier)
    np.testing.assert_allclose(out @ out.T, expected, rtol=1e-4, atol=1e-10)
    outb = O.cholesky_update_batch(L[None], v[None], multiplier)[0]
    np.testing.assert_allclose(outb, out, atol=1e-14)


def test_cholesky_downdate_failure_masks_to_zero():
    L = np.eye(3)
    out = O.cholesky_update(L, np.array([2.0, 0.0, 0.0]), -1.0)
    assert np.all(np.isfinite(out))


# ---------------------------------------------------------------- operators
def _filter_elems(dim, rng):
    A1, A2 = rng.randn(dim, dim), rng.randn(dim, dim)
    b1, b2 = rng.randn(dim), rng.randn(dim)
    U1, U2 = _rand_lower(dim, rng), _rand_lower(dim, rng)
    eta1, eta2 = rng.randn(dim), rng.randn(dim)
    Z1, Z2 = _rand_lower(dim, rng), _rand_lower(dim, rng)
    return (A1, b1, U1, eta1, Z1), (A2, b2, U2, eta2, Z2)


@pytest.mark.parametrize("dim", [1, 2, 3, 5])
@pytest.mark.parametrize("seed", [0, 42])
def test_sqrt_vs_std_filtering_operator(dim, seed):
    """tests/test_parallel_operators.py:17-58."""
    rng = np.random.RandomState(seed)
    e1, e2 = _filter_elems(dim, rng)
    s1 = (e1[0], e1[1], e1[2] @ e1[2].T, e1[3], e1[4] @ e1[4].T)
    s2 = (e2[0], e2[1], e2[2] @ e2[2].T, e2[3], e2[4] @ e2[4].T)
    A, b, C, eta, J = O.standard_filtering_operator(s1, s2)
    A_, b_, U, eta_, Z = O.sqrt_filtering_operator(e1, e2)
    np.testing.assert_allclose(A, A_, atol=1e-9, rtol=1e-9)
    np.testing.assert_allclose(b, b_, atol=1e-9, rtol=1e-9)
    np.testing.assert_allclose(eta, eta_, atol=1e-9, rtol=1e-9)
    np.testing.assert_allclose(C, U @ U.T, atol=1e-9, rtol=1e-9)
    np.testing.assert_allclose(J, Z @ Z.T, atol=1e-9, rtol=1e-9)


@pytest.mark.parametrize("dim", [1, 2, 3, 5])
@pytest.mark.parametrize("seed", [0, 42])
def test_sqrt_vs_std_smoothing_operator(dim, seed):
    """tests/test_parallel_operators.py:61-89."""
    rng = np.random.RandomState(seed)
    g1, g2 = rng.randn(dim), rng.randn(dim)
    E1, E2 = rng.randn(dim, dim), rng.randn(dim, dim)
    D1, D2 = _rand_lower(dim, rng), _rand_lower(dim, rng)
    g, E, L = O.standard_smoothing_operator((g1, E1, D1 @ D1.T), (g2, E2, D2 @ D2.T))
    g_, E_, D = O.sqrt_smoothing_operator((g1, E1, D1), (g2, E2, D2))
    np.testing.assert_allclose(g, g_, atol=1e-12)
    np.testing.assert_allclose(E, E_, atol=1e-12)
    np.testing.assert_allclose(L, D @ D.T, atol=1e-12)


def test_filtering_operator_is_associative():
    rng = np.random.RandomState(3)
    a, b = _filter_elems(4, rng)
    c, _ = _filter_elems(4, rng)
    lhs = O.sqrt_filtering_operator(O.sqrt_filtering_operator(a, b), c)
    rhs = O.sqrt_filtering_operator(a, O.sqrt_filtering_operator(b, c))
    for x, y in zip(lhs, rhs):
        if x.ndim == 2 and x.shape[0] == x.shape[1] and np.allclose(np.triu(x, 1), 0):
            np.testing.assert_allclose(x @ x.T, y @ y.T, atol=1e-9)
        else:
            np.testing.assert_allclose(x, y, atol=1e-9)


def test_scan_order_independent():
    rng = np.random.RandomState(5)
    n, dim = 11, 3
    elems = [_filter_elems(dim, rng)[0] for _ in range(n)]
    stacked = tuple(np.stack([e[j] for e in elems]) for j in range(5))
    r1 = O.associative_scan(O.sqrt_filtering_operator, stacked)
    r2 = O.sequential_fold_scan(O.sqrt_filtering_operator, stacked)
    np.testing.assert_allclose(r1[1], r2[1], atol=1e-9)
    np.testing.assert_allclose(r1[2] @ np.swapaxes(r1[2], 1, 2), r2[2] @ np.swapaxes(r2[2], 1, 2), atol=1e-9)


# ---------------------------------------------------------------- linear-Gaussian closed forms
def _lin_problem(dim_x, dim_y, seed, T):
    np.random.seed(seed)
    x0, chol_x0, F, Q, cholQ, b, _ = O.get_system(dim_x, dim_x)
    _, _, H, R, cholR, c, _ = O.get_system(dim_x, dim_y)
    _, ys = O.lgssm_data(x0.mean, F, H, R, Q, b, c, T, random_state=seed)
    trans_s = O.FunctionalModel(O.Linear(F), O.MVNSqrt(b, cholQ))
    obs_s = O.FunctionalModel(O.Linear(H), O.MVNSqrt(c, cholR))
    trans = O.FunctionalModel(O.Linear(F), O.MVNStandard(b, Q))
    obs = O.FunctionalModel(O.Linear(H), O.MVNStandard(c, R))
    return x0, chol_x0, (F, Q, b, H, R, c), ys, (trans_s, obs_s), (trans, obs)


@pytest.mark.parametrize("dim_x", [1, 2, 3])
@pytest.mark.parametrize("dim_y", [1, 2, 3])
@pytest.mark.parametrize("method", ["extended", "cubature"])
def test_filter_smoother_vs_joint_gaussian(dim_x, dim_y, method):
    """Parallel sqrt/std filter, smoother and log-likelihood against the exact
    posterior of the dense joint Gaussian (closed form)."""
    meth = getattr(O, method)
    x0, cx0, (F, Q, b, H, R, c), ys, (ts, os_), (tt, ot) = _lin_problem(dim_x, dim_y, 0, 6)
    (fm, fP), (sm, sP), ell = O.joint_gaussian_posterior(ys, x0.mean, x0.cov, F, Q, b, H, R, c)
    for x, tr, ob in ((cx0, ts, os_), (x0, tt, ot)):
        for parallel in (True, False):
            filt, e = O.filtering(ys, x, tr, ob, meth, None, parallel, return_loglikelihood=True)
            cov = filt[1] @ np.swapaxes(filt[1], 1, 2) if isinstance(x, O.MVNSqrt) else filt[1]
            np.testing.assert_allclose(filt[0], fm, atol=1e-8)
            np.testing.assert_allclose(cov, fP, atol=1e-8)
            assert e == pytest.approx(ell, rel=1e-9, abs=1e-9)
            sm_ = O.smoothing(tr, filt, meth, None, parallel)
            cov = sm_[1] @ np.swapaxes(sm_[1], 1, 2) if isinstance(x, O.MVNSqrt) else sm_[1]
            np.testing.assert_allclose(sm_[0], sm, atol=1e-8)
            np.testing.assert_allclose(cov, sP, atol=1e-8)


@pytest.mark.parametrize("dim_x", [1, 2])
@pytest.mark.parametrize("dim_y", [1, 2])
@pytest.mark.parametrize("parallel", [True, False])
def test_iterated_linear_equals_one_pass(dim_x, dim_y, parallel):
    """tests/test_iterated_smoother.py:26-69."""
    x0, cx0, (F, Q, b, H, R, c), ys, (ts, os_), _ = _lin_problem(dim_x, dim_y, 0, 5)
    nominal = O.MVNSqrt(np.random.randn(6, dim_x), np.repeat(np.eye(dim_x)[None], 6, 0))
    ref = O.seq_smoothing(ts, O.seq_filtering(ys, cx0, ts, os_, O.extended), O.extended)
    it = O.iterated_smoothing(ys, cx0, ts, os_, O.cubature, nominal, parallel, criterion=lambda i, *_: i < 5)
    np.testing.assert_allclose(it.mean, ref.mean, atol=1e-9)


@pytest.mark.parametrize("dim_x", [1, 3])
@pytest.mark.parametrize("dim_y", [1, 2, 3])
@pytest.mark.parametrize("seed", [0, 42])
def test_elements_sqrt_vs_std_and_first_element_is_kalman_step(dim_x, dim_y, seed):
    """tests/test_parallel_filter.py:26-73 and the element-1 prefix identity."""
    np.random.seed(seed)
    x0, cx0, F, Q, cholQ, b, _ = O.get_system(dim_x, dim_x)
    _, _, H, R, cholR, c, y = O.get_system(dim_x, dim_y)
    A, bb, C, eta, J = O.std_filter_element(F, Q, b, H, R, c, x0.mean, x0.cov, y)
    A_, bb_, U, eta_, Z = O.sqrt_filter_element(F, cholQ, b, H, cholR, c, cx0.mean, cx0.chol, y)
    for u, v in ((A, A_), (bb, bb_), (eta, eta_), (C, U @ U.T), (J, Z @ Z.T)):
        np.testing.assert_allclose(u, v, atol=1e-9)
    # one predict + update from the prior
    m = F @ x0.mean + b
    P = F @ x0.cov @ F.T + Q
    S = H @ P @ H.T + R
    K = P @ H.T @ np.linalg.inv(S)
    np.testing.assert_allclose(bb_, m + K @ (y - H @ m - c), atol=1e-9)
    np.testing.assert_allclose(U @ U.T, P - K @ S @ K.T, atol=1e-9)


@pytest.mark.parametrize("dim_x", [1, 2, 3])
@pytest.mark.parametrize("seed", [0, 42])
def test_smoother_element_sqrt_vs_std(dim_x, seed):
    """tests/test_parallel_smoother.py:26-54."""
    np.random.seed(seed)
    x0, cx0, F, Q, cholQ, b, _ = O.get_system(dim_x, dim_x)
    g, E, L = O.std_smoother_element(F, Q, b, x0.mean, x0.cov)
    g_, E_, D = O.sqrt_smoother_element(F, cholQ, b, cx0.mean, cx0.chol)
    np.testing.assert_allclose(g, g_, atol=1e-9)
    np.testing.assert_allclose(E, E_, atol=1e-9)
    np.testing.assert_allclose(L, D @ D.T, atol=1e-9)


# ---------------------------------------------------------------- linearizations
METHODS = {
    "extended": O.extended,
    "cubature": O.cubature,
    "gauss_hermite": O.gauss_hermite,
    "unscented": O.unscented,
}


@pytest.mark.parametrize("dim_x", [1, 3])
@pytest.mark.parametrize("seed", [0, 42])
@pytest.mark.parametrize("method", list(METHODS))
@pytest.mark.parametrize("sqrt", [True, False])
def test_linearization_exact_on_linear(dim_x, seed, method, sqrt):
    """tests/test_linearization.py:67-107."""
    np.random.seed(seed)
    a = np.random.randn(dim_x, dim_x)
    m_x, m_q = np.random.randn(dim_x), np.random.randn(dim_x)
    cx, cq = _rand_lower(dim_x), _rand_lower(dim_x)
    if sqrt:
        x, q = O.MVNSqrt(m_x, cx), O.MVNSqrt(m_q, cq)
    else:
        x, q = O.MVNStandard(m_x, cx @ cx.T), O.MVNStandard(m_q, cq @ cq.T)
    F, Ql, rem = METHODS[method](O.FunctionalModel(O.Linear(a), q), x)
    if sqrt:
        Ql = Ql @ Ql.T
    np.testing.assert_allclose(F, a, atol=1e-9)
    np.testing.assert_allclose(rem, m_q, atol=1e-9)
    np.testing.assert_allclose(Ql, cq @ cq.T, atol=1e-9)


@pytest.mark.parametrize("dim_x", [1, 3])
@pytest.mark.parametrize("seed", [0, 42])
@pytest.mark.parametrize("method", list(METHODS))
@pytest.mark.parametrize("fun", ["sin", "cos", "exp", "arctan"])
def test_linearization_sqrt_vs_std(dim_x, seed, method, fun):
    """tests/test_linearization.py:154-181."""
    np.random.seed(seed)
    m_x, m_q = np.random.randn(dim_x), np.random.randn(dim_x)
    cx, cq = _rand_lower(dim_x), _rand_lower(dim_x)
    f = O.Elementwise(fun)
    Fs, Cs, bs = METHODS[method](O.FunctionalModel(f, O.MVNSqrt(m_x, cx)), O.MVNSqrt(m_q, cq))
    F, C, b = METHODS[method](O.FunctionalModel(f, O.MVNStandard(m_x, cx @ cx.T)), O.MVNStandard(m_q, cq @ cq.T))
    np.testing.assert_allclose(Fs, F, atol=1e-10)
    np.testing.assert_allclose(bs, b, atol=1e-10)
    np.testing.assert_allclose(C, Cs @ Cs.T, atol=1e-10)


@pytest.mark.parametrize("order", [2, 3, 4, 5])
@pytest.mark.parametrize("n", [1, 2, 3])
def test_gauss_hermite_table_vs_hermgauss(order, n):
    x, w = np.polynomial.hermite.hermgauss(order)
    tab = O.gauss_hermite_table(n, order)
    assert tab.xi.shape == (order ** n, n)
    np.testing.assert_allclose(np.sort(np.unique(np.round(tab.xi[:, 0], 12))), np.sort(np.sqrt(2) * x), atol=1e-10)
    np.testing.assert_allclose(tab.wm.sum(), 1.0, atol=1e-12)
    if n == 1:
        order_ = np.argsort(tab.xi[:, 0])
        np.testing.assert_allclose(tab.wm[order_], (w / np.sqrt(np.pi))[np.argsort(x)], atol=1e-12)
    # exactness: E[x_i^2] = 1, E[x_i x_j] = 0
    np.testing.assert_allclose(tab.xi.T @ (tab.wm[:, None] * tab.xi), np.eye(n), atol=1e-10)


def test_unscented_and_cubature_moments():
    for tab in (O.cubature_table(4), O.unscented_table(4), O.unscented_table(3, 0.5, 2.0, 1.0)):
        np.testing.assert_allclose(tab.wm.sum(), 1.0, atol=1e-12)
        np.testing.assert_allclose(tab.wm @ tab.xi, 0.0, atol=1e-12)
        np.testing.assert_allclose(tab.xi.T @ (tab.wm[:, None] * tab.xi), np.eye(tab.xi.shape[1]), atol=1e-12)


def test_model_jacobians_vs_central_differences():
    rng = np.random.RandomState(1)
    models = [O.CoordinatedTurn(0.01), O.Bearings([-1.5, 0.5], [1.0, 1.0]), O.Linear(rng.randn(3, 5))]
    for f in models:
        for w in (0.7, 1e-8, -2.0):
            x = rng.randn(5)
            x[4] = w
            J = f.jac(x)
            h = 1e-6
            Jfd = np.stack([(f(x + h * e) - f(x - h * e)) / (2 * h) for e in np.eye(5)], axis=1)
            if isinstance(f, O.CoordinatedTurn) and abs(w) < 1e-6:
                # branch semantics: the w-derivative of sin(wt)/w and (cos wt - 1)/w is 0
                Jfd[:2, 4] = J[:2, 4]
            np.testing.assert_allclose(J, Jfd, atol=1e-7)


# ---------------------------------------------------------------- reference fixtures (bearings)
def _bearings(golden_dir):
    ys = np.load(f"{golden_dir}/bearings_ys.npy")
    Q, R = O.bearings_parameters(0.01, 0.1, 0.5, 0.01)
    tm = O.FunctionalModel(O.CoordinatedTurn(0.01), O.MVNSqrt(np.zeros(5), np.linalg.cholesky(Q)))
    om = O.FunctionalModel(O.Bearings([-1.5, 0.5], [1.0, 1.0]), O.MVNSqrt(np.zeros(2), np.linalg.cholesky(R)))
    return ys, O.MVNSqrt(np.array([-1.0, -1.0, 0, 0, 0]), np.eye(5)), tm, om


def test_bearings_ieks_fixture(golden_dir):
    """tests/test_bearings_only.py:21-72 (extended, parallel, 100 iterations, 3 decimals)."""
    ys, x0, tm, om = _bearings(golden_dir)
    exp = np.load(f"{golden_dir}/bearings_ieks.npz")
    res = O.iterated_smoothing(ys, x0, tm, om, O.extended, None, True, criterion=lambda i, *_: i < 100)
    np.testing.assert_array_almost_equal(res.mean[1:], exp["arr_0"], decimal=3)
    np.testing.assert_array_almost_equal(res.chol[1:] @ np.swapaxes(res.chol[1:], 1, 2), exp["arr_1"], decimal=3)


def test_bearings_icks_fixture_is_on_the_cubature_two_cycle(golden_dir):
    """The cubature iteration on this data does not converge in FP64: it
    settles into a 2-cycle.  The reference fixture (generated at FP32) lies on
    that cycle: two applications of the map starting at the fixture return
    close to it, one application does not."""
    ys, x0, tm, om = _bearings(golden_dir)
    exp = np.load(f"{golden_dir}/bearings_icks.npz")
    em, ch = exp["arr_0"], np.linalg.cholesky(exp["arr_1"])
    nom = O.MVNSqrt(np.concatenate([em[:1], em]), np.concatenate([ch[:1], ch]))
    one = O.filter_smoother(ys, x0, tm, om, O.cubature, nom, True)
    two = O.filter_smoother(ys, x0, tm, om, O.cubature, one, True)
    d1 = np.abs(one.mean[1:] - em).max(axis=1)
    d2 = np.abs(two.mean[1:] - em).max(axis=1)
    assert np.median(d2) < 3e-3 and d2.max() < 3e-2
    assert d1.max() > 5 * d2.max()


def test_fixed_point_scalar_known_answer():
    """tests/test_math_utils.py:57-67: x* = 0.5 / 0.3 and its implicit gradient."""
    a, b = 0.7, 0.5
    x_star, n = O.fixed_point(lambda x: a * x + b, 1.0, lambda i, *_: i < 500)
    assert x_star == pytest.approx(0.5 / 0.3, rel=1e-7)
    # d x*/d a = x*/(1-a), d x*/d b = 1/(1-a); implicit VJP with x_bar = 1
    ga, gb = O.fixed_point_vjp(lambda u: a * u, lambda u: (u * x_star, u), 1.0, n)
    assert ga == pytest.approx(x_star / (1 - a), rel=1e-7)
    assert gb == pytest.approx(1 / (1 - a), rel=1e-7)


# ---------------------------------------------------------------- C restatement (CPU baseline)
def _c_oracle():
    import os
    import subprocess
    from oracle import psm_oracle_c as C
    so = os.path.join(os.path.dirname(C.__file__), "liboracle_c.so")
    if not os.path.exists(so):
        subprocess.run(["make", "-C", os.path.dirname(so)], check=True, capture_output=True)
    return C


def test_c_restatement_matches_numpy_oracle_bearings():
    C = _c_oracle()
    from parsmooth.data import bearings_data
    T = 200
    _, _, ys = bearings_data(np.array([0.1, 0.2, 1.0, 0.0]), 0.01, 0.5, T, (-1.5, 0.5), (1.0, 1.0), random_state=4)
    Q, R = O.bearings_parameters(0.01, 0.1, 0.5, 0.01)
    cQ, cR = np.linalg.cholesky(Q), np.linalg.cholesky(R)
    m0 = np.array([0.1, 0.2, 1.0, 0.0, 0.0])
    init = np.tile([-1.0, -1.0, 6.0, 4.0, 2.0], (T + 1, 1))
    to = O.FunctionalModel(O.CoordinatedTurn(0.01), O.MVNSqrt(np.zeros(5), cQ))
    oo = O.FunctionalModel(O.Bearings((-1.5, 0.5), (1.0, 1.0)), O.MVNSqrt(np.zeros(2), cR))
    ref, ell = O.iterated_smoothing(ys, O.MVNSqrt(m0, np.eye(5)), to, oo, O.extended,
                                    O.MVNSqrt(init, np.tile(np.eye(5), (T + 1, 1, 1))), True,
                                    criterion=lambda i, *_: i < 3, return_loglikelihood=True)
    m, L, e = C.iterated_smoother(ys, m0, np.eye(5), init, ("ct", 0.01), ("bearings", ((-1.5, 0.5), (1.0, 1.0))),
                                  np.zeros(5), cQ, np.zeros(2), cR, 3, True)
    np.testing.assert_allclose(m, ref.mean, atol=1e-10)
    np.testing.assert_allclose(L, ref.chol, atol=1e-10)
    assert e == pytest.approx(ell, rel=1e-12)


def test_c_restatement_linear_closed_form():
    C = _c_oracle()
    x0, cx0, (F, Q, b, H, R, c), ys, (ts, os_), _ = _lin_problem(3, 2, 0, 6)
    (fm, fP), (sm, sP), ell = O.joint_gaussian_posterior(ys, x0.mean, x0.cov, F, Q, b, H, R, c)
    m, L, e = C.iterated_smoother(ys, cx0.mean, cx0.chol, np.zeros((7, 3)), ("linear", F), ("linear", H),
                                  b, np.linalg.cholesky(Q), c, np.linalg.cholesky(R), 1, True)
    np.testing.assert_allclose(m, sm, atol=1e-8)
    np.testing.assert_allclose(L @ np.swapaxes(L, 1, 2), sP, atol=1e-8)
    assert e == pytest.approx(ell, rel=1e-9)


def _bearings_c_case(T, seed):
    from parsmooth.data import bearings_data
    _, _, ys = bearings_data(np.array([0.1, 0.2, 1.0, 0.0]), 0.01, 0.5, T, (-1.5, 0.5), (1.0, 1.0),
                             random_state=seed)
    Q, R = O.bearings_parameters(0.01, 0.1, 0.5, 0.01)
    return np.asarray(ys, np.float64), np.linalg.cholesky(Q), np.linalg.cholesky(R)


@pytest.mark.parametrize("table", ["cubature", "unscented", "gauss_hermite"])
def test_c_restatement_sigma_points_match_numpy_oracle(table):
    """The C sigma-point regression (a full-size checker of cubature IPLS) against
    the numpy restatement of parsmooth/linearization/_sigma_points.py:63-100:
    three iterations of the bearings-only IPLS from a nominal with a non-identity
    chol, smoothed trajectory and log-likelihood."""
    C = _c_oracle()
    T = 60
    ys, cQ, cR = _bearings_c_case(T, 3)
    tab = {"cubature": O.cubature_table, "unscented": O.unscented_table,
           "gauss_hermite": O.gauss_hermite_table}[table](5)
    meth = {"cubature": O.cubature, "unscented": O.unscented, "gauss_hermite": O.gauss_hermite}[table]
    m0 = np.array([0.1, 0.2, 1.0, 0.0, 0.0])
    rng = np.random.RandomState(1)
    init_m = m0 + 0.05 * rng.randn(T + 1, 5)
    init_L = np.tile(np.diag([0.3, 0.4, 0.5, 0.6, 0.2]), (T + 1, 1, 1))
    to = O.FunctionalModel(O.CoordinatedTurn(0.01), O.MVNSqrt(np.zeros(5), cQ))
    oo = O.FunctionalModel(O.Bearings((-1.5, 0.5), (1.0, 1.0)), O.MVNSqrt(np.zeros(2), cR))
    ref, ell = O.iterated_smoothing(ys, O.MVNSqrt(m0, np.eye(5)), to, oo, meth, O.MVNSqrt(init_m, init_L), True,
                                    criterion=lambda i, *_: i < 3, return_loglikelihood=True)
    m, L, e, _ = C.iterated_smoother_batch([T], ys, m0[None], np.eye(5)[None], init_m, init_L, ("ct", 0.01),
                                           ("bearings", ((-1.5, 0.5), (1.0, 1.0))), np.zeros(5), cQ, np.zeros(2),
                                           cR, 3, sigma=(tab.xi, tab.wm, tab.wc), want_ell=True)
    np.testing.assert_allclose(m, ref.mean, atol=1e-9)
    np.testing.assert_allclose(L @ np.swapaxes(L, 1, 2), ref.chol @ np.swapaxes(ref.chol, 1, 2), atol=1e-9)
    assert e[0] == pytest.approx(ell, rel=1e-10)


def test_c_restatement_cubature_is_exact_on_linear_models():
    """Every sigma-point rule reproduces mean and covariance, so on a linear
    model the regression returns the model itself: cubature IPLS equals the
    dense joint-Gaussian posterior (a dropped weight, a wrong point scale or a
    missed downdate fails this)."""
    C = _c_oracle()
    x0, cx0, (F, Q, b, H, R, c), ys, _, _ = _lin_problem(3, 2, 0, 6)
    (_, _), (sm, sP), ell = O.joint_gaussian_posterior(ys, x0.mean, x0.cov, F, Q, b, H, R, c)
    tab = O.cubature_table(3)
    rng = np.random.RandomState(2)
    m, L, e, _ = C.iterated_smoother_batch([6], ys, cx0.mean[None], cx0.chol[None], rng.randn(7, 3),
                                           np.tile(2.0 * np.eye(3), (7, 1, 1)), ("linear", F), ("linear", H), b,
                                           np.linalg.cholesky(Q), c, np.linalg.cholesky(R), 1,
                                           sigma=(tab.xi, tab.wm, tab.wc), want_ell=True)
    np.testing.assert_allclose(m, sm, atol=1e-8)
    np.testing.assert_allclose(L @ np.swapaxes(L, 1, 2), sP, atol=1e-8)
    assert e[0] == pytest.approx(ell, rel=1e-9)


def test_c_restatement_batch_equals_single_sequences():
    """A ragged batch (three bearings sequences of lengths 40, 7, 25, cubature,
    two iterations, several threads) equals each sequence run alone, bit for
    bit: the batch offsets of states, observations and priors are right."""
    C = _c_oracle()
    lengths = [40, 7, 25]
    tab = O.cubature_table(5)
    cases = [_bearings_c_case(T, 10 + k) for k, T in enumerate(lengths)]
    cQ, cR = cases[0][1], cases[0][2]
    rng = np.random.RandomState(5)
    m0 = np.array([0.1, 0.2, 1.0, 0.0, 0.0]) + 0.1 * rng.randn(3, 5)
    init = [np.tile(m0[k], (T + 1, 1)) for k, T in enumerate(lengths)]
    args = (("ct", 0.01), ("bearings", ((-1.5, 0.5), (1.0, 1.0))), np.zeros(5), cQ, np.zeros(2), cR, 2)
    bm, bL, be, used = C.iterated_smoother_batch(lengths, np.concatenate([c[0] for c in cases]), m0,
                                                 np.tile(np.eye(5), (3, 1, 1)), np.concatenate(init), None, *args,
                                                 sigma=(tab.xi, tab.wm, tab.wc), want_ell=True, threads=3)
    off = np.concatenate([[0], np.cumsum(np.asarray(lengths) + 1)])
    for k, T in enumerate(lengths):
        m, L, e, _ = C.iterated_smoother_batch([T], cases[k][0], m0[k:k + 1], np.eye(5)[None], init[k], None, *args,
                                               sigma=(tab.xi, tab.wm, tab.wc), want_ell=True, threads=1)
        assert np.array_equal(bm[off[k]:off[k + 1]], m) and np.array_equal(bL[off[k]:off[k + 1]], L)
        assert be[k] == e[0]


# ---------------------------------------------------------------- pathwise sampler
def _linear_smoothing_setup(seed, T=25):
    np.random.seed(seed)
    x0, cx0, F, Q, cholQ, b, _ = O.get_system(2, 2)
    _, _, H, R, cholR, c, _ = O.get_system(2, 1)
    F = 0.9 * F / max(1.0, np.abs(np.linalg.eigvals(F)).max())
    rng = np.random.RandomState(seed)
    ys = rng.randn(T, 1)
    tq = O.FunctionalModel(O.Linear(F), O.MVNSqrt(b, cholQ))
    to = O.FunctionalModel(O.Linear(H), O.MVNSqrt(c, cholR))
    filt = O.filtering(ys, cx0, tq, to, O.extended, None, True)
    smooth = O.smoothing(tq, filt, O.extended, None, True)
    return tq, filt, smooth


def test_sampler_moments_are_the_smoothing_marginals():
    """tests/test_sampler.py:27-67: sample means and covariances of the paths
    match the smoothed marginals (here within 5 standard errors / 5 %)."""
    tq, filt, smooth = _linear_smoothing_setup(3)
    S = 40000
    eps = np.random.RandomState(7).randn(filt[0].shape[0], S, 2)
    xs = O.sampling(eps, tq, filt, O.extended, smooth)
    P = smooth.chol @ np.swapaxes(smooth.chol, -1, -2)
    se = np.sqrt(np.diagonal(P, axis1=1, axis2=2) / S)
    assert np.all(np.abs(xs.mean(1) - smooth.mean) <= 5 * se)
    C = np.einsum("tsi,tsj->tij", xs - xs.mean(1, keepdims=True), xs - xs.mean(1, keepdims=True)) / (S - 1)
    np.testing.assert_allclose(C, P, rtol=0.05, atol=0.05 * np.abs(P).max())


def test_sampler_sqrt_equals_std_for_the_same_normals():
    tq, filt, smooth = _linear_smoothing_setup(11)
    eps = np.random.RandomState(1).randn(filt[0].shape[0], 16, 2)
    xs = O.sampling(eps, tq, filt, O.extended, smooth)
    gram = lambda L: L @ np.swapaxes(L, -1, -2)
    tq_std = O.FunctionalModel(tq.function, O.MVNStandard(tq.mvn.mean, gram(tq.mvn.chol)))
    xs2 = O.sampling(eps, tq_std, O.MVNStandard(filt.mean, gram(filt.chol)), O.extended,
                     O.MVNStandard(smooth.mean, gram(smooth.chol)))
    np.testing.assert_allclose(xs2, xs, rtol=1e-9, atol=1e-9)


# ---------------------------------------------------------------- conditional moments
class _CondLinear:
    """E[y | x] = a x + b m_q + c (tests/test_linearization.py:25-26)."""

    def __init__(self, a, bq):
        self.a, self.bq = a, bq

    def __call__(self, x):
        return np.asarray(x) @ self.a.T + self.bq

    def jac(self, x):
        return np.broadcast_to(self.a, np.asarray(x).shape[:-1] + self.a.shape).copy()


class _CondConst:
    def __init__(self, M):
        self.M = M

    def __call__(self, x):
        return np.broadcast_to(self.M, np.asarray(x).shape[:-1] + self.M.shape).copy()


@pytest.mark.parametrize("dim_x", [1, 3])
@pytest.mark.parametrize("dim_q", [1, 2, 3])
@pytest.mark.parametrize("method", list(METHODS))
@pytest.mark.parametrize("sqrt", [True, False])
def test_linear_conditional_is_exact(dim_x, dim_q, method, sqrt):
    """tests/test_linearization.py:110-149."""
    np.random.seed(dim_x * 7 + dim_q)
    a, b, c = np.random.randn(dim_x, dim_x), np.random.randn(dim_x, dim_q), np.random.randn(dim_x)
    m_x, m_q = np.random.randn(dim_x), np.random.randn(dim_q)
    chol_x = np.tril(np.random.rand(dim_x, dim_x))
    chol_q = np.tril(np.random.rand(dim_q, dim_q))
    Bq = b @ chol_q
    chol_f = np.concatenate([Bq, np.zeros((dim_x, dim_x - dim_q))], 1) if dim_x > dim_q else O.tria(Bq)
    mean = _CondLinear(a, b @ m_q + c)
    if sqrt:
        model = O.ConditionalMomentsModel(mean, _CondConst(chol_f))
        x = O.MVNSqrt(m_x, chol_x)
    else:
        model = O.ConditionalMomentsModel(mean, _CondConst(Bq @ Bq.T))
        x = O.MVNStandard(m_x, chol_x @ chol_x.T)
    F, Ql, rem = METHODS[method](model, x)
    if sqrt:
        Ql = Ql @ Ql.T
    xp = np.random.randn(dim_x)
    np.testing.assert_allclose(F, a, atol=1e-7)
    np.testing.assert_allclose(F @ xp + rem, a @ xp + b @ m_q + c, atol=1e-7)
    np.testing.assert_allclose(Ql, Bq @ Bq.T, atol=1e-7)


class _PopMean:
    """Population model: log(44.7) + x - exp(x) (tests/test_linearization.py:45-46)."""

    def __call__(self, x):
        return np.log(44.7) + np.asarray(x) - np.exp(x)

    def jac(self, x):
        x = np.asarray(x)
        return (1.0 - np.exp(x))[..., :, None] * np.eye(x.shape[-1])


class _PoissonMean:
    """lam exp(x) (tests/test_linearization.py:57-58)."""

    def __init__(self, lam):
        self.lam = lam

    def __call__(self, x):
        return self.lam * np.exp(np.asarray(x))

    def jac(self, x):
        x = np.asarray(x)
        return (self.lam * np.exp(x))[..., :, None] * np.eye(x.shape[-1])


class _PoissonFactor:
    def __init__(self, lam, sqrt):
        self.lam, self.sqrt = lam, sqrt

    def __call__(self, x):
        v = self.lam * np.exp(np.asarray(x))
        return (np.sqrt(v) if self.sqrt else v)[..., None]


@pytest.mark.parametrize("seed", [0, 42])
@pytest.mark.parametrize("method", list(METHODS))
def test_conditional_sqrt_vs_std(seed, method):
    """tests/test_linearization.py:180-233: the population model's transition
    (constant noise 0.3^2) and Poisson observation (lam = 10)."""
    np.random.seed(seed)
    m_x = np.random.randn(1)
    chol_x = np.random.rand(1, 1)
    pairs = [(O.ConditionalMomentsModel(_PopMean(), _CondConst(np.array([[0.3]]))),
              O.ConditionalMomentsModel(_PopMean(), _CondConst(np.array([[0.09]])))),
             (O.ConditionalMomentsModel(_PoissonMean(10.0), _PoissonFactor(10.0, True)),
              O.ConditionalMomentsModel(_PoissonMean(10.0), _PoissonFactor(10.0, False)))]
    for sq_model, std_model in pairs:
        Fs, Ls, bs = METHODS[method](sq_model, O.MVNSqrt(m_x, chol_x))
        F, P, b = METHODS[method](std_model, O.MVNStandard(m_x, chol_x @ chol_x.T))
        np.testing.assert_allclose(Fs, F, atol=1e-10)
        np.testing.assert_allclose(bs, b, atol=1e-10)
        np.testing.assert_allclose(Ls @ Ls.T, P, atol=1e-10)


# ------------------------------------------------- models non-linear in the noise
class _LinXQ:
    """f(x, q) = a x + b q + c (tests/test_linearization.py:243-252)."""

    def __init__(self, a, b, c):
        self.a, self.b, self.c = a, b, c

    def __call__(self, x, q):
        # matmul, not einsum: a dimension mismatch must raise, as a @ x does in the reference
        return np.asarray(x) @ self.a.T + np.asarray(q) @ self.b.T + self.c

    def jac_x(self, x, q):
        return self.a.copy()

    def jac_q(self, x, q):
        return self.b.copy()


class _AffineInQ:
    """f(x, q) = g(x) + B(x) q on R^2: g(x) = (sin x0 + x1, 0.5 x0 x1),
    B(x) = [[1 + 0.1 x1^2, 0.2], [0.3 x0, 1]].  Every method reproduces the
    conditional moments g(x) + B(x) m_q and B(x) P_q B(x)^T exactly."""

    @staticmethod
    def g(x):
        return np.array([np.sin(x[0]) + x[1], 0.5 * x[0] * x[1]])

    @staticmethod
    def B(x):
        return np.array([[1.0 + 0.1 * x[1] ** 2, 0.2], [0.3 * x[0], 1.0]])

    def __call__(self, x, q):
        return self.g(x) + np.einsum("ij,...j->...i", self.B(x), q)

    def jac_x(self, x, q):
        Jg = np.array([[np.cos(x[0]), 1.0], [0.5 * x[1], 0.5 * x[0]]])
        dB = np.array([[0.0, 0.2 * x[1] * q[0]], [0.3 * q[0], 0.0]])   # d(B(x) q)/dx
        return Jg + dB

    def jac_q(self, x, q):
        return self.B(x)


class _MultNoise:
    """f(x, q) = g(x) * exp(s q) + r q on R^2 with g(x) = (x0 + 0.1 x1,
    0.95 x1 + 0.1 sin x0), s = 0.2, r = 0.1: noise that is multiplicative and
    additive at once."""
    s, r = 0.2, 0.1

    @staticmethod
    def g(x):
        return np.array([x[0] + 0.1 * x[1], 0.95 * x[1] + 0.1 * np.sin(x[0])])

    def __call__(self, x, q):
        return self.g(x) * np.exp(self.s * np.asarray(q)) + self.r * np.asarray(q)

    def jac_x(self, x, q):
        Jg = np.array([[1.0, 0.1], [0.1 * np.cos(x[0]), 0.95]])
        return np.exp(self.s * np.asarray(q))[:, None] * Jg

    def jac_q(self, x, q):
        return np.diag(self.s * self.g(x) * np.exp(self.s * np.asarray(q)) + self.r)


class _LogNormal:
    """f(x, q) = x exp(q) in one dimension: E[f | x] = x exp(m + s^2 / 2),
    Var[f | x] = x^2 (exp(s^2) - 1) exp(2 m + s^2) for q ~ N(m, s^2)."""

    def __call__(self, x, q):
        return np.asarray(x) * np.exp(np.asarray(q))

    def jac_x(self, x, q):
        return np.exp(np.asarray(q)).reshape(1, 1)

    def jac_q(self, x, q):
        return (np.asarray(x) * np.exp(np.asarray(q))).reshape(1, 1)


@pytest.mark.parametrize("dim_x", [1, 2])
@pytest.mark.parametrize("dim_q", [1, 2])
@pytest.mark.parametrize("seed", [0, 42])
@pytest.mark.parametrize("method", list(METHODS))
@pytest.mark.parametrize("sqrt", [True, False])
def test_get_conditional_model_linear(dim_x, dim_q, seed, method, sqrt):
    """tests/test_linearization.py:238-283: the outer linearization of the
    conditional model of a x + b q + c recovers a, the mean and (b chol_q)(b chol_q)^T."""
    np.random.seed(seed)
    a, b, c = np.random.randn(dim_x, dim_x), np.random.randn(dim_x, dim_q), np.random.randn(dim_x)
    m_x, m_q = np.random.randn(dim_x), np.random.randn(dim_q)
    chol_x, chol_q = _rand_lower(dim_x), _rand_lower(dim_q)
    if sqrt:
        x_mvn, q_mvn = O.MVNSqrt(m_x, chol_x), O.MVNSqrt(m_q, chol_q)
    else:
        x_mvn, q_mvn = O.MVNStandard(m_x, chol_x @ chol_x.T), O.MVNStandard(m_q, chol_q @ chol_q.T)
    f = _LinXQ(a, b, c)
    if dim_x != dim_q:
        with pytest.raises(NotImplementedError):
            O.get_conditional_model(f, q_mvn, METHODS[method])
        return
    model = O.get_conditional_model(f, q_mvn, METHODS[method])
    F_x, Q_lin, rem = METHODS[method](model, x_mvn)
    if sqrt:
        Q_lin = Q_lin @ Q_lin.T
    x_prime = np.random.randn(dim_x)
    np.testing.assert_allclose(F_x, a, atol=1e-9)
    np.testing.assert_allclose(F_x @ x_prime + rem, a @ x_prime + b @ m_q + c, atol=1e-9)
    np.testing.assert_allclose(Q_lin, (b @ chol_q) @ (b @ chol_q).T, atol=1e-9)


@pytest.mark.parametrize("method", list(METHODS))
@pytest.mark.parametrize("sqrt", [True, False])
def test_get_conditional_model_affine_in_q(method, sqrt):
    """Non-linear in x, affine in q: the conditional moments are exact for every
    inner method, and the conditional mean's Jacobian is dg/dx + d(B(x) m_q)/dx."""
    rng = np.random.RandomState(7)
    m_q = rng.randn(2)
    cq = _rand_lower(2, rng) + np.eye(2) * 0.2
    q = O.MVNSqrt(m_q, cq) if sqrt else O.MVNStandard(m_q, cq @ cq.T)
    f = _AffineInQ()
    model = O.get_conditional_model(f, q, METHODS[method])
    xs = rng.randn(6, 2)
    got_m = model.conditional_mean(xs)
    got_c = model.conditional_covariance_or_cholesky(xs)
    got_J = model.conditional_mean.jac(xs)
    for k, x in enumerate(xs):
        B = f.B(x)
        np.testing.assert_allclose(got_m[k], f.g(x) + B @ m_q, rtol=1e-12, atol=1e-12)
        cov = got_c[k] @ got_c[k].T if sqrt else got_c[k]
        np.testing.assert_allclose(cov, B @ cq @ cq.T @ B.T, rtol=1e-11, atol=1e-12)
        if sqrt:
            assert np.allclose(got_c[k], np.tril(got_c[k])) and np.all(np.diag(got_c[k]) >= 0)
        np.testing.assert_allclose(got_J[k], f.jac_x(x, m_q), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("sqrt", [True, False])
def test_get_conditional_model_lognormal(sqrt):
    """x exp(q) under Gauss-Hermite with 12 nodes: the log-normal mean and
    variance to 1e-11 (the variance is mostly the non-linear residual
    Phi - F P F^T, so a dropped residual fails here)."""
    import functools
    m, s = 0.3, 0.4
    q = O.MVNSqrt(np.array([m]), np.array([[s]])) if sqrt else O.MVNStandard(np.array([m]), np.array([[s * s]]))
    model = O.get_conditional_model(_LogNormal(), q, functools.partial(O.gauss_hermite, order=12))
    xs = np.array([[0.5], [-1.3], [2.0]])
    mean = model.conditional_mean(xs)[:, 0]
    fac = model.conditional_covariance_or_cholesky(xs)[:, 0, 0]
    var = fac * fac if sqrt else fac
    x = xs[:, 0]
    np.testing.assert_allclose(mean, x * np.exp(m + s * s / 2), rtol=1e-11)
    np.testing.assert_allclose(var, x * x * np.expm1(s * s) * np.exp(2 * m + s * s), rtol=1e-10)
    np.testing.assert_allclose(model.conditional_mean.jac(xs)[:, 0, 0], np.exp(m + s * s / 2), rtol=1e-11)


# ------------------------------------- sequential passes on non-linear models
def _nonlinear_problem(kind, sqrt, T=25, seed=3):
    """Simulated data and (x0, transition, observation) for a non-linear model:
    'ct_bearings' (coordinated turn + two bearings), 'ct_linear' (coordinated
    turn observed in position) or 'cv_bearings' (constant velocity + bearings)."""
    rng = np.random.RandomState(seed)
    dt = 0.1
    Q5, R = O.bearings_parameters(0.1, 0.2, 0.05, dt)
    if kind == "cv_bearings":
        F = np.eye(4)
        F[0, 2] = F[1, 3] = dt
        Q = Q5[:4, :4]
        f = O.Linear(F)
        m0 = np.array([0.1, 0.2, 1.0, 0.5])
    else:
        Q = Q5
        f = O.CoordinatedTurn(dt)
        m0 = np.array([0.1, 0.2, 1.0, 0.5, 0.8])
    nx = m0.size
    if kind == "ct_linear":
        H = np.zeros((2, nx))
        H[0, 0] = H[1, 1] = 1.0
        h = O.Linear(H)
    else:
        h = O.Bearings((-1.5, 0.5), (1.0, 1.0))
    cQ, cR = np.linalg.cholesky(Q), np.linalg.cholesky(R)
    x = m0.copy()
    ys = np.empty((T, 2))
    for k in range(T):
        x = f(x) + cQ @ rng.randn(nx)
        ys[k] = h(x) + cR @ rng.randn(2)
    P0 = 0.1 * np.eye(nx)
    if sqrt:
        mk = lambda M, C: O.MVNSqrt(M, np.linalg.cholesky(C))
    else:
        mk = O.MVNStandard
    return ys, mk(m0, P0), O.FunctionalModel(f, mk(np.zeros(nx), Q)), O.FunctionalModel(h, mk(np.zeros(2), R))


def _noisy_nominal(rng, traj, sqrt):
    m = traj[0] + 0.05 * rng.randn(*traj[0].shape)
    L = np.tile(0.3 * np.eye(m.shape[1]), (m.shape[0], 1, 1))
    return O.MVNSqrt(m, L) if sqrt else O.MVNStandard(m, L @ L)


@pytest.mark.parametrize("method", list(METHODS))
@pytest.mark.parametrize("sqrt", [True, False])
def test_sequential_equals_parallel_at_a_nominal(method, sqrt):
    """At a given nominal trajectory both formulations linearize at the same
    points, so the sequential recursion and the parallel scan give the same
    filter, log-likelihood and smoother on a non-linear model."""
    ys, x0, tr, ob = _nonlinear_problem("ct_bearings", sqrt)
    meth = METHODS[method]
    rng = np.random.RandomState(1)
    nom = _noisy_nominal(rng, O.filtering(ys, x0, tr, ob, meth, None, False), sqrt)
    fs, es = O.filtering(ys, x0, tr, ob, meth, nom, False, return_loglikelihood=True)
    fp, ep = O.filtering(ys, x0, tr, ob, meth, nom, True, return_loglikelihood=True)
    np.testing.assert_allclose(fs[0], fp[0], rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(fs[1], fp[1], rtol=1e-9, atol=1e-9)
    assert es == pytest.approx(ep, rel=1e-9)
    ss = O.smoothing(tr, fs, meth, nom, False)
    sp = O.smoothing(tr, fs, meth, nom, True)
    np.testing.assert_allclose(ss[0], sp[0], rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(ss[1], sp[1], rtol=1e-9, atol=1e-9)


def _predicted(tr, meth, traj, sqrt):
    """The predicted distributions F x_{k-1} + b (chol tria([F L, cq]) or
    covariance F P F^T + Q) of a linear transition from every state but the last."""
    F, cq, b = O.linearize_many(meth, tr, type(traj)(traj[0][:-1], traj[1][:-1]))
    m = np.einsum("kij,kj->ki", F, traj[0][:-1]) + b
    if sqrt:
        L = O.tria(np.concatenate([F @ traj[1][:-1], cq], -1))
    else:
        L = F @ traj[1][:-1] @ np.swapaxes(F, -1, -2) + cq
    return m, L


@pytest.mark.parametrize("method", list(METHODS))
@pytest.mark.parametrize("sqrt", [True, False])
def test_sequential_linearization_points(method, sqrt):
    """Without a nominal trajectory the sequential filter linearizes the
    transition at the filtered state k-1 and the observation at the predicted
    state k, and the smoother the transition out of k at the smoothed state k+1
    (parsmooth/sequential/_filtering.py:36-41, _smoothing.py:27-33).  Each
    result is therefore reproduced by the same pass at the nominal trajectory
    made of those points: with a linear observation the filter at its own
    output; with a linear transition the filter at its own predictions; the
    smoother at its output shifted by one state."""
    meth = METHODS[method]
    # transition points: a linear observation makes the observation point irrelevant
    ys, x0, tr, ob = _nonlinear_problem("ct_linear", sqrt)
    filt = O.filtering(ys, x0, tr, ob, meth, None, False)
    again = O.filtering(ys, x0, tr, ob, meth, filt, False)
    np.testing.assert_allclose(again[0], filt[0], rtol=1e-10, atol=1e-11)
    np.testing.assert_allclose(again[1], filt[1], rtol=1e-10, atol=1e-11)
    # observation points: a linear transition makes the transition point irrelevant
    ys, x0, tr, ob = _nonlinear_problem("cv_bearings", sqrt)
    filt = O.filtering(ys, x0, tr, ob, meth, None, False)
    pm, pL = _predicted(tr, meth, filt, sqrt)
    nom = type(filt)(np.concatenate([filt[0][:1], pm]), np.concatenate([filt[1][:1], pL]))
    again = O.filtering(ys, x0, tr, ob, meth, nom, False)
    np.testing.assert_allclose(again[0], filt[0], rtol=1e-10, atol=1e-11)
    np.testing.assert_allclose(again[1], filt[1], rtol=1e-10, atol=1e-11)
    # and at the filtered states instead the result changes (the points matter here)
    other = O.filtering(ys, x0, tr, ob, meth, filt, False)
    assert np.abs(other[0] - filt[0]).max() > 1e-6
    # smoother points: the smoothed successor
    ys, x0, tr, ob = _nonlinear_problem("ct_bearings", sqrt)
    filt = O.filtering(ys, x0, tr, ob, meth, None, False)
    sm = O.smoothing(tr, filt, meth, None, False)
    shifted = type(sm)(np.concatenate([sm[0][1:], sm[0][-1:]]), np.concatenate([sm[1][1:], sm[1][-1:]]))
    again = O.smoothing(tr, filt, meth, shifted, False)
    np.testing.assert_allclose(again[0], sm[0], rtol=1e-10, atol=1e-11)
    np.testing.assert_allclose(again[1], sm[1], rtol=1e-10, atol=1e-11)
    other = O.smoothing(tr, filt, meth, sm, False)
    assert np.abs(other[0] - sm[0]).max() > 1e-6
