"""Single steps of the sequential filter and smoother on the device, after the
reference's tests/test_sequential_filter.py:24-118 and
tests/test_sequential_smoother.py:26-66 (same systems, same closed forms), and
against the oracle's one-step log-likelihood."""
import numpy as np
import pytest
import torch
from scipy.linalg import solve

from oracle import psm_oracle as O

pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("needs a HIP device", allow_module_level=True)

from parsmooth.sequential._filtering import _sqrt_predict, _sqrt_update, _standard_predict, _standard_update  # noqa: E402
from parsmooth.sequential._smoothing import _sqrt_smooth, _standard_smooth  # noqa: E402

npy = lambda t: t.detach().cpu().numpy().astype(np.float64) if isinstance(t, torch.Tensor) else np.asarray(t)


@pytest.mark.parametrize("dim_x", [1, 2, 3, 5])
@pytest.mark.parametrize("seed", [0, 42])
def test_predict_standard_vs_sqrt(dim_x, seed):
    np.random.seed(seed)
    x, chol_x, F, Q, cholQ, b, _ = O.get_system(dim_x, dim_x)
    nc = _sqrt_predict(F, cholQ, b, chol_x)
    ns = _standard_predict(F, Q, b, x)
    np.testing.assert_allclose(npy(ns.mean), npy(nc.mean), atol=1e-12)
    np.testing.assert_allclose(npy(ns.cov), npy(nc.chol) @ npy(nc.chol).T, atol=1e-12)


@pytest.mark.parametrize("dim_x", [1, 2, 3])
@pytest.mark.parametrize("seed", [0, 42])
@pytest.mark.parametrize("sqrt", [True, False])
def test_predict_value(dim_x, seed, sqrt):
    np.random.seed(seed)
    x, chol_x, F, Q, cholQ, b, _ = O.get_system(dim_x, dim_x)
    nx_ = _sqrt_predict(F, cholQ, b, chol_x) if sqrt else _standard_predict(F, Q, b, x)
    cov = npy(nx_.chol) @ npy(nx_.chol).T if sqrt else npy(nx_.cov)
    np.testing.assert_allclose(npy(nx_.mean), F @ x.mean + b, atol=1e-12)
    np.testing.assert_allclose(cov, Q + F @ x.cov @ F.T, atol=1e-12)


@pytest.mark.parametrize("dim_x", [1, 2, 3])
@pytest.mark.parametrize("dim_y", [1, 2, 3])
@pytest.mark.parametrize("seed", [0, 42])
@pytest.mark.parametrize("sqrt", [True, False])
def test_update_value(dim_x, dim_y, seed, sqrt):
    np.random.seed(seed)
    x, chol_x, H, R, cholR, c, y = O.get_system(dim_x, dim_y)
    if sqrt:
        nx_, ell = _sqrt_update(H, cholR, c, chol_x, y)
        cov = npy(nx_.chol) @ npy(nx_.chol).T
    else:
        nx_, ell = _standard_update(H, R, c, x, y)
        cov = npy(nx_.cov)
    res = y - H @ x.mean - c
    S = H @ x.cov @ H.T + R
    K = x.cov @ solve(S, H, assume_a="pos").T
    np.testing.assert_allclose(npy(nx_.mean), x.mean + K @ res, atol=1e-10)
    np.testing.assert_allclose(cov, x.cov - K @ H @ x.cov, atol=1e-10)
    np.testing.assert_allclose(float(ell), O.mvn_loglikelihood(res, np.linalg.cholesky(S)), atol=1e-10)


@pytest.mark.parametrize("dim_x", [1, 3])
@pytest.mark.parametrize("dim_y", [1, 2, 3])
@pytest.mark.parametrize("seed", [0, 42])
def test_update_standard_vs_sqrt(dim_x, dim_y, seed):
    np.random.seed(seed)
    x, chol_x, H, R, cholR, c, y = O.get_system(dim_x, dim_y)
    xs, ell = _standard_update(H, R, c, x, y)
    xc, chol_ell = _sqrt_update(H, cholR, c, chol_x, y)
    np.testing.assert_allclose(npy(xs.cov), npy(xc.chol) @ npy(xc.chol).T, atol=1e-10)
    np.testing.assert_allclose(npy(xs.mean), npy(xc.mean), atol=1e-10)
    np.testing.assert_allclose(float(ell), float(chol_ell), atol=1e-10)


@pytest.mark.parametrize("dim_x", [1, 3])
@pytest.mark.parametrize("dim_y", [1, 2, 3])
@pytest.mark.parametrize("seed", [0, 42])
def test_update_standard_vs_sqrt_no_info(dim_x, dim_y, seed):
    """Observation noise x1e8 (chol x1e4): the update changes nothing (atol
    1e-3, the reference's)."""
    np.random.seed(seed)
    x, chol_x, H, R, cholR, c, y = O.get_system(dim_x, dim_y)
    R, cholR = R * 1e8, cholR * 1e4
    nxs, ell = _standard_update(H, R, c, x, y)
    nxc, chol_ell = _sqrt_update(H, cholR, c, chol_x, y)
    np.testing.assert_allclose(float(ell), float(chol_ell), atol=1e-3)
    np.testing.assert_allclose(x.cov, npy(nxs.cov), atol=1e-3)
    np.testing.assert_allclose(x.mean, npy(nxs.mean), atol=1e-3)
    np.testing.assert_allclose(chol_x.chol @ chol_x.chol.T, npy(nxc.chol) @ npy(nxc.chol).T, atol=1e-3)
    np.testing.assert_allclose(x.mean, npy(nxc.mean), atol=1e-3)


@pytest.mark.parametrize("dim_x", [2, 4])
@pytest.mark.parametrize("dim_y", [1, 2])
def test_predict_update_log_likelihood_matches_oracle(dim_x, dim_y):
    """Prediction then update: the oracle's one-step sqrt log-likelihood term."""
    np.random.seed(dim_x * 10 + dim_y)
    _, chol_x, F, _, cholQ, b, _ = O.get_system(dim_x, dim_x)
    _, _, H, _, cholR, c, y = O.get_system(dim_x, dim_y)
    pred = _sqrt_predict(F, cholQ, b, chol_x)
    _, ell = _sqrt_update(H, cholR, c, pred, y)
    ref = O.sqrt_loglikelihood(F, cholQ, b, H, cholR, c, chol_x.mean, chol_x.chol, y)
    assert float(ell) == pytest.approx(float(ref), rel=1e-12, abs=1e-12)


@pytest.mark.parametrize("dim_x", [1, 2, 3])
@pytest.mark.parametrize("seed", [0, 42])
def test_smooth_one_standard_vs_sqrt(dim_x, seed):
    np.random.seed(seed)
    xf, chol_xf, F, Q, cholQ, b, _ = O.get_system(dim_x, dim_x)
    xs, chol_xs, *_ = O.get_system(dim_x, dim_x)
    ns = _standard_smooth(F, Q, b, xf, xs)
    nc = _sqrt_smooth(F, cholQ, b, chol_xf, chol_xs)
    np.testing.assert_allclose(npy(ns.mean), npy(nc.mean), atol=1e-10)
    np.testing.assert_allclose(npy(ns.cov), npy(nc.chol) @ npy(nc.chol).T, atol=1e-10)


@pytest.mark.parametrize("dim_x", [1, 2, 3])
@pytest.mark.parametrize("seed", [0, 42])
@pytest.mark.parametrize("sqrt", [True, False])
def test_smooth_one_value(dim_x, seed, sqrt):
    np.random.seed(seed)
    xf, chol_xf, F, Q, cholQ, b, _ = O.get_system(dim_x, dim_x)
    xs, chol_xs, *_ = O.get_system(dim_x, dim_x)
    nx_ = _sqrt_smooth(F, cholQ, b, chol_xf, chol_xs) if sqrt else _standard_smooth(F, Q, b, xf, xs)
    cov = npy(nx_.chol) @ npy(nx_.chol).T if sqrt else npy(nx_.cov)
    m_ = F @ xf.mean + b
    P_ = F @ xf.cov @ F.T + Q
    G = xf.cov @ solve(P_.T, F, assume_a="pos").T
    np.testing.assert_allclose(npy(nx_.mean), xf.mean + G @ (xs.mean - m_), atol=1e-10)
    np.testing.assert_allclose(cov, xf.cov + G @ (xs.cov - P_) @ G.T, atol=1e-10)


def test_steps_accept_fp32_tensors():
    """All-float32 tensor arguments stay float32 on the device."""
    np.random.seed(1)
    x, chol_x, F, Q, cholQ, b, _ = O.get_system(3, 3)
    t = lambda a: torch.as_tensor(np.asarray(a), dtype=torch.float32, device="cuda")
    out = _sqrt_predict(t(F), t(cholQ), t(b), (t(chol_x.mean), t(chol_x.chol)))
    assert out.mean.dtype == torch.float32 and out.chol.dtype == torch.float32
    np.testing.assert_allclose(npy(out.chol) @ npy(out.chol).T, Q + F @ x.cov @ F.T, rtol=1e-5, atol=1e-5)
