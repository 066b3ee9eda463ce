"""``parsmooth.linearization.get_conditional_model`` (the conditional moments of
a model non-linear in its noise, parsmooth/linearization/_common.py:17-66):
the torch callables it returns, evaluated on the CPU, against the oracle's
restatement of the reference.  The device path through them is covered by
tests/test_gpu_conditional.py."""
import functools

import numpy as np
import pytest
import torch

from oracle import psm_oracle as O
from parsmooth import MVNSqrt, MVNStandard
from parsmooth.linearization import cubature, extended, gauss_hermite, get_conditional_model, unscented

from test_oracle import _AffineInQ, _LinXQ, _MultNoise, _rand_lower

METHODS = {"extended": (extended, O.extended), "cubature": (cubature, O.cubature),
           "gauss_hermite": (gauss_hermite, O.gauss_hermite), "unscented": (unscented, O.unscented)}


def _c(a, like):
    return torch.as_tensor(np.asarray(a, np.float64), dtype=like.dtype, device=like.device)


def lin_xq_torch(a, b, c):
    return lambda x, q: _c(a, x) @ x + _c(b, q) @ q + _c(c, x)


def affine_in_q_torch(x, q):
    g = torch.stack([torch.sin(x[0]) + x[1], 0.5 * x[0] * x[1]])
    B0 = torch.stack([1.0 + 0.1 * x[1] ** 2, torch.full_like(x[0], 0.2)])
    B1 = torch.stack([0.3 * x[0], torch.ones_like(x[0])])
    return g + torch.stack([B0 @ q, B1 @ q])


def mult_noise_torch(x, q):
    s, r = _MultNoise.s, _MultNoise.r
    g = torch.stack([x[0] + 0.1 * x[1], 0.95 * x[1] + 0.1 * torch.sin(x[0])])
    return g * torch.exp(s * q) + r * q


def lognormal_torch(x, q):
    return x * torch.exp(q)


CASES = {
    "affine_in_q": (affine_in_q_torch, _AffineInQ()),
    "mult_noise": (mult_noise_torch, _MultNoise()),
}


def _noise(rng, n, sqrt):
    m = rng.randn(n)
    L = _rand_lower(n, rng) + 0.2 * np.eye(n)
    if sqrt:
        return MVNSqrt(torch.as_tensor(m), torch.as_tensor(L)), O.MVNSqrt(m, L)
    return MVNStandard(torch.as_tensor(m), torch.as_tensor(L @ L.T)), O.MVNStandard(m, L @ L.T)


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("method", list(METHODS))
@pytest.mark.parametrize("sqrt", [True, False])
def test_moments_match_oracle(case, method, sqrt):
    rng = np.random.RandomState(11)
    tf, of = CASES[case]
    q_t, q_o = _noise(rng, 2, sqrt)
    mt = get_conditional_model(tf, q_t, METHODS[method][0])
    mo = O.get_conditional_model(of, q_o, METHODS[method][1])
    xs = rng.randn(5, 2)
    want_m = mo.conditional_mean(xs)
    want_c = mo.conditional_covariance_or_cholesky(xs)
    for k, x in enumerate(xs):
        xt = torch.as_tensor(x)
        np.testing.assert_allclose(mt.conditional_mean(xt).numpy(), want_m[k], rtol=1e-12, atol=1e-13)
        got = mt.conditional_covariance_or_cholesky(xt).numpy()
        np.testing.assert_allclose(got, want_c[k], rtol=1e-11, atol=1e-12)


@pytest.mark.parametrize("method", list(METHODS))
@pytest.mark.parametrize("sqrt", [True, False])
def test_batched_evaluation_matches_pointwise(method, sqrt):
    """The model kind evaluates the callables under torch.func.vmap (and the
    mean's Jacobian with jacfwd): same values as one point at a time, and the
    Jacobian equals the oracle's."""
    rng = np.random.RandomState(2)
    q_t, q_o = _noise(rng, 2, sqrt)
    mt = get_conditional_model(mult_noise_torch, q_t, METHODS[method][0])
    mo = O.get_conditional_model(_MultNoise(), q_o, METHODS[method][1])
    xs = torch.as_tensor(rng.randn(7, 2))
    vm = torch.func.vmap(mt.conditional_mean)(xs)
    vc = torch.func.vmap(mt.conditional_covariance_or_cholesky)(xs)
    J = torch.func.vmap(torch.func.jacfwd(mt.conditional_mean))(xs)
    for k in range(xs.shape[0]):
        np.testing.assert_allclose(vm[k].numpy(), mt.conditional_mean(xs[k]).numpy(), rtol=1e-14, atol=1e-15)
        np.testing.assert_allclose(vc[k].numpy(), mt.conditional_covariance_or_cholesky(xs[k]).numpy(),
                                   rtol=1e-13, atol=1e-14)
    np.testing.assert_allclose(J.numpy(), mo.conditional_mean.jac(xs.numpy()), rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("dim_x", [1, 2])
@pytest.mark.parametrize("dim_q", [1, 2])
@pytest.mark.parametrize("method", list(METHODS))
@pytest.mark.parametrize("sqrt", [True, False])
def test_linear_model(dim_x, dim_q, method, sqrt):
    """tests/test_linearization.py:238-283 at the level of the conditional
    moments: mean a x + b m_q + c, factor of b P_q b^T; different dimensions raise."""
    rng = np.random.RandomState(dim_x * 3 + dim_q)
    a, b, c = rng.randn(dim_x, dim_x), rng.randn(dim_x, dim_q), rng.randn(dim_x)
    q_t, q_o = _noise(rng, dim_q, sqrt)
    if dim_x != dim_q:
        with pytest.raises(NotImplementedError):
            get_conditional_model(lin_xq_torch(a, b, c), q_t, METHODS[method][0])
        with pytest.raises(NotImplementedError):
            O.get_conditional_model(_LinXQ(a, b, c), q_o, METHODS[method][1])
        return
    mt = get_conditional_model(lin_xq_torch(a, b, c), q_t, METHODS[method][0])
    mo = O.get_conditional_model(_LinXQ(a, b, c), q_o, METHODS[method][1])
    x = rng.randn(dim_x)
    np.testing.assert_allclose(mt.conditional_mean(torch.as_tensor(x)).numpy(), mo.conditional_mean(x),
                               rtol=1e-12, atol=1e-13)
    got = mt.conditional_covariance_or_cholesky(torch.as_tensor(x)).numpy()
    np.testing.assert_allclose(got, mo.conditional_covariance_or_cholesky(x), rtol=1e-11, atol=1e-12)


@pytest.mark.parametrize("sqrt", [True, False])
def test_lognormal_gauss_hermite_order(sqrt):
    """The partial's keywords reach the inner method: 12 Gauss-Hermite nodes give
    the log-normal moments to 1e-11, the default 3 nodes do not."""
    m, s = 0.3, 0.4
    t = lambda a: torch.tensor(a, dtype=torch.float64)
    q = MVNSqrt(t([m]), t([[s]])) if sqrt else MVNStandard(t([m]), t([[s * s]]))
    x = torch.tensor([1.7], dtype=torch.float64)
    want_m = 1.7 * np.exp(m + s * s / 2)
    want_v = 1.7 ** 2 * np.expm1(s * s) * np.exp(2 * m + s * s)
    for order, tol in ((12, 1e-11), (3, None)):
        mt = get_conditional_model(lognormal_torch, q, functools.partial(gauss_hermite, order=order))
        fac = float(mt.conditional_covariance_or_cholesky(x).reshape(-1)[0])
        var = fac * fac if sqrt else fac
        got = float(mt.conditional_mean(x)[0])
        if tol is None:
            assert abs(var - want_v) / want_v > 1e-6
        else:
            assert abs(got - want_m) / want_m < tol and abs(var - want_v) / want_v < 10 * tol


def test_rejects_non_gaussian_noise():
    with pytest.raises(TypeError):
        get_conditional_model(lognormal_torch, (torch.zeros(1), torch.ones(1, 1)), extended)


def test_dimension_probe_runs_on_the_noise_device():
    """The probe f(q.mean, q.mean) runs where q lives: a function whose constants
    sit on that device passes it, and a dimension mismatch still raises."""
    A = torch.eye(2, dtype=torch.float64)
    f = lambda x, q: x @ A.T + q   # noqa: E731
    q = MVNSqrt(torch.zeros(2, dtype=torch.float64), torch.eye(2, dtype=torch.float64))
    get_conditional_model(f, q, cubature)
    with pytest.raises(NotImplementedError):
        get_conditional_model(f, MVNSqrt(torch.zeros(3, dtype=torch.float64), torch.eye(3, dtype=torch.float64)),
                              cubature)
