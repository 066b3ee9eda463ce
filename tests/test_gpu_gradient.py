"""Gradient of the iterated-smoother pseudo-log-likelihood (tangent kernels)
against central finite differences of the FP64 oracle."""
import numpy as np
import pytest
import torch

from oracle import psm_oracle as O

pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("needs a HIP device", allow_module_level=True)

from parsmooth import _native as N  # noqa: E402
from parsmooth.gradient import IteratedLoglikelihood, loglikelihood  # noqa: E402
from parsmooth.linearization import cubature, extended  # noqa: E402


def _linear_build(theta):
    """theta = (log q, log r, a): F = [[1, a], [0, 1]], Q = q I, R = r, H = [1 0]."""
    q, r, a = torch.exp(theta[0]), torch.exp(theta[1]), theta[2]
    F = torch.stack([torch.stack([torch.ones((), dtype=theta.dtype), a]),
                     torch.tensor([0.0, 1.0], dtype=theta.dtype)])
    return {"trans_params": F.reshape(-1), "trans_q_mean": torch.zeros(2, dtype=theta.dtype),
            "trans_q_chol": torch.sqrt(q) * torch.eye(2, dtype=theta.dtype),
            "obs_params": torch.tensor([1.0, 0.0], dtype=theta.dtype),
            "obs_q_mean": torch.zeros(1, dtype=theta.dtype),
            "obs_q_chol": torch.sqrt(r).reshape(1, 1),
            "m0": torch.zeros(1, 2, dtype=theta.dtype), "L0": torch.eye(2, dtype=theta.dtype)[None]}


def _oracle_linear_ell(theta, ys):
    v = {k: t.detach().numpy() for k, t in _linear_build(torch.as_tensor(theta)).items()}
    tm = O.FunctionalModel(O.Linear(v["trans_params"].reshape(2, 2)), O.MVNSqrt(v["trans_q_mean"], v["trans_q_chol"]))
    om = O.FunctionalModel(O.Linear(v["obs_params"].reshape(1, 2)), O.MVNSqrt(v["obs_q_mean"], v["obs_q_chol"]))
    _, ell = O.filtering(ys, O.MVNSqrt(v["m0"][0], v["L0"][0]), tm, om, O.extended, None, True,
                         return_loglikelihood=True)
    return ell


def test_gradient_linear_model_matches_finite_differences():
    rng = np.random.RandomState(0)
    T = 200
    theta = np.array([np.log(0.05), np.log(0.3), 0.1])
    v = {k: t.numpy() for k, t in _linear_build(torch.as_tensor(theta)).items()}
    _, ys = O.lgssm_data(np.zeros(2), v["trans_params"].reshape(2, 2), np.array([[1.0, 0.0]]),
                         v["obs_q_chol"] ** 2, v["trans_q_chol"] @ v["trans_q_chol"].T, np.zeros(2), np.zeros(1),
                         T, random_state=rng)
    obj = IteratedLoglikelihood(_linear_build, N.MODEL_LINEAR, N.MODEL_LINEAR, 2, 1, [T], ys, extended, 2,
                                (np.zeros((T + 1, 2)), np.tile(np.eye(2), (T + 1, 1, 1))))
    ell, grad = obj.value_and_grad(torch.as_tensor(theta))
    ell_o = _oracle_linear_ell(theta, ys)
    assert float(ell[0]) == pytest.approx(ell_o, rel=1e-10)
    h = 1e-5
    fd = np.array([(_oracle_linear_ell(theta + h * e, ys) - _oracle_linear_ell(theta - h * e, ys)) / (2 * h)
                   for e in np.eye(3)])
    np.testing.assert_allclose(grad[0].cpu().numpy(), fd, rtol=1e-6, atol=1e-6)


def _bearings_build(theta):
    """theta = (log qc, log qw, log sigma, s1x, s1y, s2x, s2y); dt = 0.01."""
    dt = 0.01
    qc, qw, sig = torch.exp(theta[0]), torch.exp(theta[1]), torch.exp(theta[2])
    z = torch.zeros((), dtype=theta.dtype)
    Q = torch.stack([torch.stack([qc * dt ** 3 / 3, z, qc * dt ** 2 / 2, z, z]),
                     torch.stack([z, qc * dt ** 3 / 3, z, qc * dt ** 2 / 2, z]),
                     torch.stack([qc * dt ** 2 / 2, z, qc * dt, z, z]),
                     torch.stack([z, qc * dt ** 2 / 2, z, qc * dt, z]),
                     torch.stack([z, z, z, z, dt * qw])])
    return {"trans_params": torch.tensor([dt], dtype=theta.dtype), "trans_q_mean": torch.zeros(5, dtype=theta.dtype),
            "trans_q_chol": torch.linalg.cholesky(Q), "obs_params": theta[3:7],
            "obs_q_mean": torch.zeros(2, dtype=theta.dtype), "obs_q_chol": sig * torch.eye(2, dtype=theta.dtype),
            "m0": torch.tensor([[0.1, 0.2, 1.0, 0.0, 0.0]], dtype=theta.dtype),
            "L0": torch.eye(5, dtype=theta.dtype)[None]}


def _oracle_bearings_ell(theta, ys, init, n_iter, method):
    v = {k: t.detach().numpy() for k, t in _bearings_build(torch.as_tensor(theta)).items()}
    tm = O.FunctionalModel(O.CoordinatedTurn(0.01), O.MVNSqrt(v["trans_q_mean"], v["trans_q_chol"]))
    om = O.FunctionalModel(O.Bearings(v["obs_params"][:2], v["obs_params"][2:]), O.MVNSqrt(v["obs_q_mean"], v["obs_q_chol"]))
    _, ell = O.iterated_smoothing(ys, O.MVNSqrt(v["m0"][0], v["L0"][0]), tm, om, method, init, True,
                                  criterion=lambda i, *_: i < n_iter, return_loglikelihood=True)
    return ell


@pytest.mark.parametrize("method", ["extended", "cubature"])
def test_implicit_gradient_bearings_matches_finite_differences(method):
    """At a converged fixed point the implicit gradient equals the derivative of
    the converged log-likelihood (central differences, h = 1e-4, FP64)."""
    from parsmooth.data import bearings_data
    T, n_iter = 120, 50  # both methods are converged to rounding level by then
    theta = np.array([np.log(0.01), np.log(0.1), np.log(0.5), -1.5, 0.5, 1.0, 1.0])
    _, truth, ys = bearings_data(np.array([0.1, 0.2, 1.0, 0.0]), 0.01, 0.5, T, (-1.5, 0.5), (1.0, 1.0),
                                 random_state=2)
    ys = ys.astype(np.float64)
    init = O.MVNSqrt(truth.astype(np.float64), np.tile(0.1 * np.eye(5), (T + 1, 1, 1)))
    om = {"extended": O.extended, "cubature": O.cubature}[method]
    dm = {"extended": extended, "cubature": cubature}[method]
    obj = IteratedLoglikelihood(_bearings_build, N.MODEL_COORDINATED_TURN, N.MODEL_BEARINGS, 5, 2, [T], ys, dm,
                                n_iter, (init.mean, init.chol))
    ell, grad = obj.value_and_grad(torch.as_tensor(theta))
    ell_o = _oracle_bearings_ell(theta, ys, init, n_iter, om)
    assert float(ell[0]) == pytest.approx(ell_o, rel=1e-9)
    h = 1e-4
    fd = np.array([(_oracle_bearings_ell(theta + h * e, ys, init, n_iter, om)
                    - _oracle_bearings_ell(theta - h * e, ys, init, n_iter, om)) / (2 * h) for e in np.eye(7)])
    g = grad[0].cpu().numpy()
    np.testing.assert_allclose(g, fd, rtol=1e-3, atol=1e-3 * np.abs(fd).max())
    # and through torch.autograd
    th = torch.as_tensor(theta, dtype=torch.float64).requires_grad_(True)
    loglikelihood(th, obj).backward()
    np.testing.assert_allclose(th.grad.numpy(), g, rtol=1e-12)


def _oracle_pass(theta, ys, nominal, method):
    v = {k: t.detach().numpy() for k, t in _bearings_build(torch.as_tensor(theta)).items()}
    tm = O.FunctionalModel(O.CoordinatedTurn(0.01), O.MVNSqrt(v["trans_q_mean"], v["trans_q_chol"]))
    om = O.FunctionalModel(O.Bearings(v["obs_params"][:2], v["obs_params"][2:]), O.MVNSqrt(v["obs_q_mean"], v["obs_q_chol"]))
    return O.filter_smoother(ys, O.MVNSqrt(v["m0"][0], v["L0"][0]), tm, om, method, O.MVNSqrt(*nominal), True)


@pytest.mark.parametrize("method", ["extended", "cubature"])
def test_tangent_pass_matches_central_differences(method):
    """One filter-smoother pass in the tangent dtype: the value equals the FP64
    pass and the tangent equals the central difference of the oracle pass along
    a random direction of parameters, nominal mean and nominal chol."""
    from parsmooth.data import bearings_data
    T = 60
    theta = np.array([np.log(0.01), np.log(0.1), np.log(0.5), -1.5, 0.5, 1.0, 1.0])
    _, truth, ys = bearings_data(np.array([0.1, 0.2, 1.0, 0.0]), 0.01, 0.5, T, (-1.5, 0.5), (1.0, 1.0),
                                 random_state=4)
    ys = ys.astype(np.float64)
    rng = np.random.RandomState(3)
    nm = truth.astype(np.float64) + 0.05 * rng.randn(T + 1, 5)
    nL = np.tril(0.02 * rng.randn(T + 1, 5, 5)) + 0.2 * np.eye(5)
    dth, dm, dL = rng.randn(7), rng.randn(T + 1, 5), np.tril(rng.randn(T + 1, 5, 5))
    om = {"extended": O.extended, "cubature": O.cubature}[method]
    dmeth = {"extended": extended, "cubature": cubature}[method]
    obj = IteratedLoglikelihood(_bearings_build, N.MODEL_COORDINATED_TURN, N.MODEL_BEARINGS, 5, 2, [T], ys, dmeth,
                                1, (nm, nL))
    (vm, vL), (tm_, tL) = obj.jvp_pass(theta, dth, (nm, nL), (dm, dL))
    ref = _oracle_pass(theta, ys, (nm, nL), om)
    np.testing.assert_allclose(vm.cpu().numpy(), ref.mean, rtol=0, atol=1e-9 * np.abs(ref.mean).max())
    np.testing.assert_allclose(vL.cpu().numpy(), ref.chol, rtol=0, atol=1e-9 * np.abs(ref.chol).max())
    h = 1e-5  # central-difference error here: ~1e-8 (means), ~1e-6 (chols), relative
    plus = _oracle_pass(theta + h * dth, ys, (nm + h * dm, nL + h * dL), om)
    minus = _oracle_pass(theta - h * dth, ys, (nm - h * dm, nL - h * dL), om)
    fd_m, fd_L = (plus.mean - minus.mean) / (2 * h), (plus.chol - minus.chol) / (2 * h)
    np.testing.assert_allclose(tm_.cpu().numpy(), fd_m, rtol=0, atol=1e-6 * np.abs(fd_m).max())
    np.testing.assert_allclose(tL.cpu().numpy(), fd_L, rtol=0, atol=1e-5 * np.abs(fd_L).max())
